"""GPU parity at the benchmark's own sizes.

* The fused time step on the 256^3 grid the benchmark times (the y/z
  momentum and projection sweeps of 32 segments, the 256-point x sweeps with
  4 points per lane, the 256^3 Poisson passes) against the oracle's step
  (transeq.f90:73-470, time_integrators.f90:85-100, navier.f90:257-499) on a
  perturbed Taylor-Green field: FP64 within 1e-11 relative; FP32 bulk within
  1e-5 of the oracle's velocity (each projection removes the divergence of
  the FP32 copy of u, whose derivatives carry eps32 max|u| / h = 2.4e-6).
* On the 512^3 microbenchmark grids, where the oracle takes minutes, size-
  independent properties:
    - div(grad(poisson(f))) = f on band-limited zero-mean fields, with the
      staggered operators of navier.f90:257-431 composed through the C-ABI
      (poisson_000 divides by exactly the eigenvalue of that composition,
      poisson.f90:1706-1783);
    - single Fourier modes times a random cross-line field: every periodic
      derivative returns the closed-form modified wavenumber times the
      shifted mode (schemes.f90:477-480, :658-673) line by line;
    - wall lines of 513 points: the Dirichlet closures of dery_22 / deryy_22
      are exact on polynomials of degree <= 2 (first derivative) and <= 3
      (second derivative), times a random cross-line field.
"""
import math

import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import dev_tensor, perturbed_tgv, rel_err

pytestmark = pytest.mark.gpu

TWOPI = 2.0 * math.acos(-1.0)


def test_fused_step_256_matches_oracle(torch_cuda):
    import incompact3d_amd as X
    n = 256
    kw = dict(nx=n, ny=n, nz=n, re=1600.0, dt=0.002)
    p = O.Params(**kw)
    u = perturbed_tgv(p, seed=1, amp=0.05)
    print("256^3 fused step: device FP64 and FP32 bulk, 2 steps", flush=True)
    X.init(X.make_params(**kw))
    try:
        got = {}
        for prec in (0, 1):
            s = X.Solver(X.make_params(**kw), prec)
            s.set_velocity(*[dev_tensor(c) for c in u])
            s.step(2)
            got[prec] = [t.cpu().numpy() for t in s.velocity()]
            s.close()
    finally:
        X.finalize()
    print("256^3 fused step: oracle, 2 steps", flush=True)
    S = O.Schemes(p)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    for _ in range(2):
        O.step(S, st)
    for c, ref in enumerate((st.ux, st.uy, st.uz)):
        e64 = rel_err(got[0][c], ref)
        e32 = rel_err(got[1][c], ref)
        print(f"component {c}: FP64 {e64:.2e}, FP32 bulk {e32:.2e}", flush=True)
        assert e64 <= 1e-11
        assert e32 <= 1e-5


def _band_limited(shape, seed, kmax=12, nmodes=24):
    """Zero-mean sum of random Fourier modes with |k|_inf <= kmax on the
    (nz, ny, nx) periodic grid of [0, 2 pi)^3, built on the device."""
    import torch
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    x = torch.arange(nx, device="cuda", dtype=torch.float64) * (TWOPI / nx)
    y = torch.arange(ny, device="cuda", dtype=torch.float64) * (TWOPI / ny)
    z = torch.arange(nz, device="cuda", dtype=torch.float64) * (TWOPI / nz)
    f = torch.zeros(shape, device="cuda", dtype=torch.float64)
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not k.any():
            k[0] = 1
        a = float(rng.uniform(0.5, 1.0))
        ph = float(rng.uniform(0.0, TWOPI))
        f += a * torch.sin(k[0] * x[None, None, :] + k[1] * y[None, :, None] + k[2] * z[:, None, None] + ph)
    return f


def _div_grad(X, pr):
    """Sum over d of D_vp,d D_pv,d prod_{e != d} I_vp,e I_pv,e applied to pr
    (gradp navier.f90:386-431 then divergence navier.f90:257-339) on the
    periodic 512^3 grid, where the pressure and velocity grids coincide."""
    import torch
    a, b, c = torch.empty_like(pr), torch.empty_like(pr), torch.empty_like(pr)
    out = torch.zeros_like(pr)
    for d in range(3):
        src = pr
        for e in range(3):
            # along e: the derivative pair when e == d, the interpolation pair otherwise
            op_pv, op_vp, np_vp = ("derpv", "dervp", 0) if e == d else ("interpv", "intervp", 1)
            dst = b if e % 2 == 0 else c
            X.line_op_async(e, op_pv, 1, src, a)
            X.line_op_async(e, op_vp, np_vp, a, dst)
            src = dst
        X.synchronize()
        out += src
    return out


def test_512_div_grad_inverts_poisson(torch_cuda):
    """FP64: div(grad(p)) = f for p = poisson(f).  Roundoff in p sits at every
    mode and div(grad) multiplies mode k by K(k) <= K_max = 6.6e4 on this
    grid, so the bar is 1e-10 (eps64 K_max = 1.5e-11).  FP32 (FP32 FFTs, FP64
    divide): the solution within 1e-5 of the FP64 one, the bar of the
    Poisson parity tests."""
    import torch
    import incompact3d_amd as X
    n = 512
    print("512^3 div(grad(poisson f)) = f, FP64", flush=True)
    X.init(X.make_params(n, n, n))
    try:
        f = _band_limited((n, n, n), seed=4)
        f -= f.mean()
        p64 = f.clone()
        X.poisson(p64)
        back = _div_grad(X, p64)
        err = float((back - f).abs().max() / f.abs().max())
        print(f"max |div grad p - f| / max |f| = {err:.2e}", flush=True)
        assert err <= 1e-10
        del back
        print("512^3 poisson FP32 against FP64", flush=True)
        p32 = f.float()
        X.poisson(p32)
        err32 = float((p32.double() - p64).abs().max() / p64.abs().max())
        print(f"max |p32 - p64| / max |p64| = {err32:.2e}", flush=True)
        assert err32 <= 1e-5
    finally:
        X.finalize()


# (a, b, alpha) of the 6th-order first derivative and (a, b, alpha) of the
# second derivative (schemes.f90:477-480, :658-673)
A1, B1, AL1 = 7.0 / 9.0, 1.0 / 36.0, 1.0 / 3.0
A2, B2, AL2 = 12.0 / 11.0, 3.0 / 44.0, 2.0 / 11.0


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_512_single_mode_derivatives(torch_cuda, dtype):
    import torch
    import incompact3d_amd as X
    n = 512
    h = TWOPI / n
    print(f"512^3 single-mode derivatives, {dtype}", flush=True)
    X.init(X.make_params(n, n, n))
    try:
        tdt = torch.float64 if dtype == "f64" else torch.float32
        c = torch.arange(n, device="cuda", dtype=torch.float64) * h
        g = torch.rand((n, n), device="cuda", dtype=torch.float64, generator=torch.Generator("cuda").manual_seed(3))
        for d, k in ((0, 37), (1, 101), (2, 229)):
            w = k * h
            k1 = (2 * A1 * math.sin(w) + 2 * B1 * math.sin(2 * w)) / (1 + 2 * AL1 * math.cos(w)) / h
            k2 = (2 * A2 * (1 - math.cos(w)) + 2 * B2 * (1 - math.cos(2 * w))) / (1 + 2 * AL2 * math.cos(w)) / h ** 2
            shape = [1, 1, 1]
            shape[2 - d] = n
            s = torch.sin(k * c + 0.3).reshape(shape)
            co = torch.cos(k * c + 0.3).reshape(shape)
            # the cross-line factor varies over the two other directions
            gx = g.reshape([n if i != 2 - d else 1 for i in range(3)])
            u = (s * gx).to(tdt).contiguous()
            t = torch.empty_like(u)
            for op, npaire, ref, scale in (("der1", 0, k1 * co * gx, k1), ("der1", 1, k1 * co * gx, k1),
                                           ("der2", 1, -k2 * s * gx, k2)):
                X.line_op_async(d, op, npaire, u, t)
                X.synchronize()
                err = float((t.double() - ref).abs().max()) / scale
                print(f"dir {d} {op} npaire {npaire}: {err:.2e}", flush=True)
                assert err <= (1e-12 if dtype == "f64" else 2e-6), (d, op, npaire, err)
    finally:
        X.finalize()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_512x513x512_wall_polynomials(torch_cuda, dtype):
    """dery_22 / deryy_22 on the 512x513x512 walls grid (derive.f90:783-854,
    :2829-2919): the wall closures are exact on quadratics (first
    derivative) and cubics (second derivative); the oracle pins the same
    exactness on a short line."""
    import torch
    import incompact3d_amd as X
    from incompact3d_amd import derive as D
    nx, ny, nz = 512, 513, 512
    yly = 2.0
    print(f"512x513x512 wall polynomials, {dtype}", flush=True)
    X.init(X.make_params(nx, ny, nz, yly=yly, ncly1=2, nclyn=2))
    try:
        tdt = torch.float64 if dtype == "f64" else torch.float32
        h = yly / (ny - 1)
        y = (torch.arange(ny, device="cuda", dtype=torch.float64) * h - 1.0).reshape(1, ny, 1)
        g = torch.rand((nz, 1, nx), device="cuda", dtype=torch.float64,
                       generator=torch.Generator("cuda").manual_seed(5)) + 0.5
        poly = 0.3 + 0.7 * y - 1.1 * y ** 2
        dpoly = 0.7 - 2.2 * y
        cub = 0.2 - 0.4 * y + 0.9 * y ** 2 + 0.6 * y ** 3
        d2cub = 1.8 + 3.6 * y
        cy = X.coef
        # roundoff of a k-th derivative scales with max|u| / h^k, so the error
        # is measured in that unit (on the oracle, an exact FP64 input gives
        # 1e-13 (k = 1) and 2.4e-14 (k = 2); rounding the input to FP32 alone
        # gives 2.8e-7 and 6.3e-7).  FP32 arithmetic: the wall rows of the
        # second derivative sum coefficients of up to ~60 in magnitude before
        # a solve whose wall row has alpha = 11, so its bar is 5e-5 (the
        # noise-field parity bar, 2e-5 of max|result|, is looser in this unit)
        for field, exact, kind, k in ((poly, dpoly, "der1", 1), (cub, d2cub, "der2", 2)):
            u = (field * g).to(tdt).contiguous()
            t = torch.empty_like(u)
            if kind == "der1":
                D.dery_22(t, u, None, None, cy.ffy, cy.fsy, cy.fwy, np.ones(ny), nx, ny, nz, 0)
            else:
                D.deryy_22(t, u, None, None, cy.sfy, cy.ssy, cy.swy, nx, ny, nz, 0)
            ref = exact * g
            err = float((t.double() - ref).abs().max() / (u.double().abs().max() / h ** k))
            print(f"{kind}: {err:.2e} of max|u|/h^{k}", flush=True)
            assert err <= (2e-13 if dtype == "f64" else (5e-6 if k == 1 else 5e-5)), (kind, err)
            del u, t
    finally:
        X.finalize()
