"""CPU pins of the oracle's channel pieces (Case-Channel.f90, the ncl=2
branches of navier.f90 and the stretched-y viscous term of transeq.f90).

* hash_uniform: the SplitMix64 known answer (first output of the generator
  seeded with 0 is 0xE220A8397B1DCDAF, Steele, Lea & Flood 2014 / Vigna's
  reference splitmix64.c) and uniformity moments.
* channel_cfr: the bulk velocity is a quadrature of u dy over the wall-
  normal mesh: linear in u, exact shift property, and convergent to the
  Poiseuille bulk 2/3 as ny grows, on uniform and stretched meshes (a
  metric used the wrong way round does not converge).
* the stretched y viscous term pp2y D2 - pp4y (ppy D1) of momentum_rhs:
  for u = f(y) only, the right-hand side is nu f''(y), checked on
  polynomials against the exact second derivative as ny grows.
* pre_correc_walls / gradp_walls / step_channel: the wall-normal velocity is
  exactly zero on the walls after every step (the Neumann pressure closure
  has a zero normal gradient there), the projected field is divergence-free,
  and the wall slip left by the pressure correction decays over steps (AB3)
  or stays below its first value (RK3); a sign error in the stored wall
  gradient makes it grow.
"""
import math

import numpy as np
import pytest

from oracle import x3d_oracle as O


def _channel(nx=16, ny=33, nz=8, istret=2, **kw):
    p = O.Params(nx, ny, nz, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2,
                 istret=istret, beta=0.259, **kw)
    return p, O.Schemes(p)


def test_hash_uniform_known_answer():
    got = O.hash_uniform(0, 0, np.array([1], dtype=np.uint64))[0]
    assert got == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    r = O.hash_uniform(20261202, 1, np.arange(200000, dtype=np.uint64))
    assert 0.0 <= r.min() and r.max() < 1.0
    assert abs(r.mean() - 0.5) < 3e-3 and abs(r.var() - 1.0 / 12.0) < 2e-3
    # components and seeds give different streams
    assert not np.array_equal(r[:100], O.hash_uniform(20261202, 2, np.arange(100, dtype=np.uint64)))


@pytest.mark.parametrize("istret", [0, 2])
def test_cfr_is_a_bulk_quadrature(istret):
    errs = []
    for ny in (33, 65, 129, 257):
        p, S = _channel(nx=8, ny=ny, nz=8, istret=istret)
        y = O.channel_y(S)
        u = np.broadcast_to((1.0 - y * y)[None, :, None], (p.nz, p.ny, p.nx)).copy()
        _, ub = O.channel_cfr(S, u, 2.0 / 3.0)
        errs.append(abs(ub - 2.0 / 3.0))
        # linearity and the shift property
        _, ub2 = O.channel_cfr(S, 3.0 * u, 0.0)
        assert abs(ub2 - 3.0 * ub) <= 1e-13
        one = np.ones_like(u)
        _, u1 = O.channel_cfr(S, one, 0.0)
        v, _ = O.channel_cfr(S, u, 0.5)
        _, ubv = O.channel_cfr(S, v, 0.0)
        assert abs(ubv - (ub + (0.5 - ub) * u1)) <= 1e-13
    # rectangle rule over ny points: first-order convergence to 2/3
    assert errs[-1] < 0.02 and errs[-1] < 0.6 * errs[0], errs


@pytest.mark.parametrize("istret", [1, 2])
def test_stretched_viscous_term_is_second_derivative(istret):
    errs = []
    for ny in (33, 65, 129):
        p, S = _channel(nx=8, ny=ny, nz=8, istret=istret, re=1.0)
        y = O.channel_y(S)
        f = y ** 3 - 0.5 * y * y
        ux = np.broadcast_to(f[None, :, None], (p.nz, p.ny, p.nx)).copy()
        z = np.zeros_like(ux)
        dux, duy, duz = O.momentum_rhs(S, ux, z, z)
        exact = (6.0 * y - 1.0)[None, :, None]
        e = np.abs(dux - exact)[:, 3:-3, :].max()   # interior: the wall closures are 3rd order
        errs.append(e)
        assert np.abs(duy).max() == 0.0 and np.abs(duz).max() == 0.0
    assert errs[-1] < 1e-3 and errs[-1] < 0.1 * errs[0], errs


@pytest.mark.parametrize("scheme", [3, 5])
def test_channel_step_wall_invariants(scheme):
    p, S = _channel(re=500.0, dt=0.002, itimescheme=scheme)
    u = O.init_channel(S, 0.125, 5)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    wg = O.wall_grad_zero(p)
    slips = []
    for _ in range(6):
        st, wg = O.step_channel(S, st, wg)
        assert np.abs(st.uy[:, 0, :]).max() == 0.0 and np.abs(st.uy[:, -1, :]).max() == 0.0
        assert np.abs(O.divergence(S, st.ux, st.uy, st.uz)).max() < 1e-11
        slips.append(max(np.abs(st.ux[:, 0, :]).max(), np.abs(st.ux[:, -1, :]).max()))
    if scheme == 3:
        assert slips[-1] < 0.3 * slips[0], slips
    else:
        # RK3 rescales the stored gradient by gdt(itr)/gdt(itr-1) between
        # sub-steps: the slip stays bounded by its first value
        assert max(slips[1:]) <= slips[0], slips


@pytest.mark.parametrize("scheme", [3, 5])
def test_poiseuille_is_steady_under_its_pressure_gradient(scheme):
    """Laminar Poiseuille u = 1 - y^2 with the constant pressure gradient
    fcpg = 2 nu (cpg branch) is a steady state: nu u'' + fcpg = 0.  On a
    uniform mesh the compact schemes and their wall closures differentiate
    quadratics exactly, so every step returns the profile to roundoff; a
    wrong sign or a missing factor on fcpg drives the flow off at O(dt)."""
    re = 100.0
    p, S = _channel(nx=16, ny=33, nz=8, istret=0, re=re, dt=0.01, itimescheme=scheme)
    y = O.channel_y(S)
    u0 = np.broadcast_to((1.0 - y * y)[None, :, None], (p.nz, p.ny, p.nx)).copy()
    z = np.zeros_like(u0)
    st = O.State(u0.copy(), z.copy(), z.copy(), [u0.copy()] * 3, [z.copy()] * 3, [z.copy()] * 3, 0)
    wg = O.wall_grad_zero(p)
    for _ in range(5):
        st, wg = O.step_channel(S, st, wg, cfr=None, fcpg=2.0 / re)
    assert np.abs(st.ux - u0).max() < 1e-12
    assert np.abs(st.uy).max() < 1e-12 and np.abs(st.uz).max() < 1e-12
    st2 = O.State(u0.copy(), z.copy(), z.copy(), [u0.copy()] * 3, [z.copy()] * 3, [z.copy()] * 3, 0)
    wg2 = O.wall_grad_zero(p)
    st2, wg2 = O.step_channel(S, st2, wg2, cfr=None, fcpg=-2.0 / re)
    assert np.abs(st2.ux - u0).max() > 1e-5


def test_spinup_rotation_is_a_rotation_of_the_velocity(monkeypatch):
    """momentum_forcing_channel's spin-up term (Case-Channel.f90:413-418)
    with the transport terms switched off: the forcing is the horizontal
    velocity turned by +90 degrees and scaled by wrotation, so it is
    orthogonal to u at every point (it does no work), its magnitude is
    wrotation |(ux, uy)|, its sense is positive (ux fy - uy fx =
    wrotation (ux^2 + uy^2)), and it acts only while itime < spinup_time."""
    p, S = _channel(re=500.0, dt=0.002, itimescheme=3)
    u = O.init_channel(S, 0.125, 5)
    zero = np.zeros_like(u[0])
    monkeypatch.setattr(O, "momentum_rhs", lambda S_, ux, uy, uz: (zero.copy(), zero.copy(), zero.copy()))
    w = 0.12

    def first_rhs(spinup):
        st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [zero] * 3, [zero] * 3, [zero] * 3, 0)
        st, _ = O.step_channel(S, st, O.wall_grad_zero(p), cfr=None, wrotation=w, spinup_time=spinup)
        return st.hx[0], st.hy[0], st.hz[0]

    fx, fy, fz = first_rhs(2)
    ux, uy, uz = u
    scale = w * float(np.max(ux * ux + uy * uy))
    assert np.abs(ux * fx + uy * fy + uz * fz).max() <= 1e-14 * scale
    assert np.abs(np.hypot(fx, fy) - w * np.hypot(ux, uy)).max() <= 1e-14 * w * np.abs(ux).max()
    assert np.abs((ux * fy - uy * fx) - w * (ux * ux + uy * uy)).max() <= 1e-14 * scale
    assert np.abs(fz).max() == 0.0
    for f in first_rhs(1):
        assert np.abs(f).max() == 0.0
