"""Shared test helpers (not a test module)."""
import math

import numpy as np

TWOPI = 2.0 * math.acos(-1.0)


def seeded_field(shape, seed, lengths=(TWOPI, TWOPI, TWOPI), nmodes=8, kmax=4):
    """U[-1,1] noise plus `nmodes` random low-wavenumber sine modes
    (|k|_inf <= kmax, amplitude U[0.5,1], phase U[0,2pi)) on a (nz, ny, nx)
    grid spanning `lengths` = (xlx, yly, zlz) without the end point."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    u = rng.uniform(-1.0, 1.0, size=shape)
    x = np.arange(nx) * lengths[0] / nx
    y = np.arange(ny) * lengths[1] / ny
    z = np.arange(nz) * lengths[2] / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        a = rng.uniform(0.5, 1.0)
        ph = rng.uniform(0.0, TWOPI)
        u += a * np.sin(k[0] * X * TWOPI / lengths[0] + k[1] * Y * TWOPI / lengths[1]
                        + k[2] * Z * TWOPI / lengths[2] + ph)
    return u


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def perturbed_tgv(p, seed=0, amp=0.05):
    """TGV plus a divergence-free perturbation u' = curl A of four low modes."""
    import math
    from oracle import x3d_oracle as O
    ux, uy, uz = O.init_tgv(p)
    rng = np.random.default_rng(seed)
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    z = np.arange(p.nz) * p.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    for _ in range(4):
        k = rng.integers(1, 3, size=3) * rng.choice([-1, 1], size=3)
        a = rng.standard_normal(3) * amp
        ph = rng.uniform(0, 2 * math.pi)
        c = np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
        # curl of a sin(k.x + ph) = (k x a) cos(k.x + ph)
        ux = ux + (k[1] * a[2] - k[2] * a[1]) * c
        uy = uy + (k[2] * a[0] - k[0] * a[2]) * c
        uz = uz + (k[0] * a[1] - k[1] * a[0]) * c
    return ux, uy, uz


def dev_tensor(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def les_steps_vs_oracle(X, kw, nsteps, iconserv, cs=0.17, tol=1e-10):
    """nsteps Smagorinsky LES steps on the device (context of kw already
    initialised) against the oracle's step_les, FP64."""
    from oracle import les_oracle as L
    from oracle import x3d_oracle as O
    p = O.Params(**kw)
    S = O.Schemes(p)
    u = perturbed_tgv(p, seed=1)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    fn = L.smag_sgs(cs, iconserv)
    for _ in range(nsteps):
        L.step_les(S, st, fn)
    s = X.Solver(X.make_params(**kw), 0)
    s.set_velocity(*[dev_tensor(c) for c in u])
    s.set_les(1, cs=cs, iconserv=iconserv)
    s.step(nsteps)
    got = [t.cpu().numpy() for t in s.velocity()]
    s.close()
    for g, r in zip(got, (st.ux, st.uy, st.uz)):
        assert rel_err(g, r) <= tol
