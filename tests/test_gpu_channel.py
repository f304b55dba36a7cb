"""Channel (config 4 family): the device time step with no-slip walls in y,
the stretched-y mesh, the constant flow rate and the constant pressure
gradient, against the oracle's step_channel (Case-Channel.f90, the ncl=2
branches of navier.f90, transeq.f90:337-370).

* init_channel: the counter-based noise and the Poiseuille profile, 1e-14.
* FP64 steps vs the oracle: 1e-11 (AB3 and RK3; uniform and stretched
  meshes; constant flow rate or pressure gradient; with and without the
  spin-up rotation, which stops after its last step).
* FP32 bulk vs the FP64 oracle: 2e-5 relative.
* At the configuration's grid 256x129x128 (istret=2, beta=0.259, RK3,
  dt=0.005): the wall-normal velocity is exactly zero on the walls, the
  projected field is divergence-free to roundoff, the constant flow rate
  holds the bulk velocity at 2/3 and the flow stays bounded.
"""
import math

import numpy as np
import pytest

from tests.helpers import dev_tensor, rel_err

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")


def _kw(nx, ny, nz, istret, scheme, re=4200.0, dt=0.005):
    return dict(nx=nx, ny=ny, nz=nz, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2,
                istret=istret, beta=0.259, re=re, dt=dt, itimescheme=scheme, isecondder=5)


def _oracle(kw):
    from oracle import x3d_oracle as O
    p = O.Params(**kw)
    return O, p, O.Schemes(p)


@pytest.mark.parametrize("istret", [0, 2])
def test_init_channel(istret):
    kw = _kw(16, 33, 16, istret, 5)
    O, p, S = _oracle(kw)
    ref = O.init_channel(S, 0.125, 77)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), 0)
        s.init_channel(0.125, 77)
        got = [t.cpu().numpy() for t in s.velocity()]
        s.close()
    finally:
        X.finalize()
    for g, r in zip(got, ref):
        assert np.abs(g - r).max() <= 1e-14


CASES = [
    # (nx, ny, nz, istret, scheme, forcing)
    (16, 33, 16, 2, 5, "cfr"),
    (16, 33, 16, 2, 3, "cfr"),
    (16, 33, 16, 0, 5, "cfr"),
    (32, 17, 16, 1, 3, "cpg"),
    (16, 33, 16, 3, 5, "cfr"),
    # spin-up rotation (wrotation 0.12) on steps 1 and 2 of 4 (itime < 3)
    (16, 33, 16, 2, 5, "cfr+rot"),
    (32, 17, 16, 1, 3, "cpg+rot"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"-s{c[3]}-t{c[4]}-{c[5]}")
@pytest.mark.parametrize("precision", [0, 1])
def test_channel_steps_vs_oracle(case, precision):
    nx, ny, nz, istret, scheme, forcing = case
    kw = _kw(nx, ny, nz, istret, scheme, re=500.0, dt=0.002)
    O, p, S = _oracle(kw)
    u = O.init_channel(S, 0.125, 5)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    wg = O.wall_grad_zero(p)
    nsteps = 4
    fcpg = 2.0 / 500.0 if forcing.startswith("cpg") else None
    cfr = 2.0 / 3.0 if forcing.startswith("cfr") else None
    wrot, spinup = (0.12, 3) if forcing.endswith("+rot") else (0.0, 0)
    for _ in range(nsteps):
        st, wg = O.step_channel(S, st, wg, cfr=cfr, fcpg=fcpg, wrotation=wrot, spinup_time=spinup)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), precision)
        s.set_velocity(*[dev_tensor(c) for c in u])
        if cfr is not None:
            s.set_forcing(1, 2.0 / 3.0)
        else:
            s.set_forcing(2, fcpg)
        if spinup:
            s.set_rotation(wrot, spinup)
        s.step(nsteps)
        got = [t.cpu().numpy() for t in s.velocity()]
        ub = s.bulk_velocity()
        s.close()
    finally:
        X.finalize()
    tol = 1e-11 if precision == 0 else 2e-5
    for g, r in zip(got, (st.ux, st.uy, st.uz)):
        assert rel_err(g, r) <= tol
    _, ub_ref = O.channel_cfr(S, st.ux, 0.0)
    assert abs(ub - ub_ref) <= (1e-12 if precision == 0 else 1e-6)


@pytest.mark.parametrize("precision", [0, 1])
def test_channel_256x129x128_invariants(precision):
    """The configuration's grid: 60 RK3 steps of the stretched channel at
    Re 4200 (bulk), dt 0.005, constant flow rate."""
    kw = _kw(256, 129, 128, 2, 5)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), precision)
        s.init_channel(0.125, 11)
        s.set_forcing(1, 2.0 / 3.0)
        for _ in range(3):
            s.step(20)
            ux, uy, uz = s.velocity()
            assert float(uy[:, 0, :].abs().max()) == 0.0 and float(uy[:, -1, :].abs().max()) == 0.0
            mx = max(float(t.abs().max()) for t in (ux, uy, uz))
            assert math.isfinite(mx) and mx < 2.0
            dmax, dmean = s.divergence_stats()
            # FP32 bulk: the FP64 divergence of the FP64 velocity carries the
            # FP32 rounding of the pressure gradient (~6e-8) through the
            # derivatives across the finest wall cells (dy ~ 3e-3)
            assert abs(dmax) < (1e-9 if precision == 0 else 1e-4)
            # the flow rate is reset to 2/3 before every sub-step's right-hand side
            assert abs(s.bulk_velocity() - 2.0 / 3.0) < 2e-3
        s.close()
    finally:
        X.finalize()
