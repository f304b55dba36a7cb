"""GPU parity on a stretched y mesh (istret = 1, 2, 3; stretching.f90) against
the oracle, through the C-ABI: the metrics x3d_init computes, the y
derivatives that multiply by them (dery * ppy, deryvp * ppyi, derypv * ppy),
and the stretched poisson_010 (matrice_refinement + inversion5_v1 / _v2
applied as stored pentadiagonal factors between the z transforms).

Bar: 1e-12 relative in FP64; 2e-5 (derivatives) and 1e-5 (Poisson) for the
FP32 entries.  deryy applies no metric, as in the reference.
"""
import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import rel_err, seeded_field

pytestmark = pytest.mark.gpu

CASES = [
    ("s2_walls_32x33x16", dict(nx=32, ny=33, nz=16, ncly1=2, nclyn=2, istret=2, beta=0.259)),
    ("s1_freeslip_64x65x32", dict(nx=64, ny=65, nz=32, ncly1=1, nclyn=1, istret=1, beta=0.7)),
    ("s2_mixed_16x129x64", dict(nx=16, ny=129, nz=64, ncly1=2, nclyn=1, istret=2, beta=0.5)),
    ("s3_half_32x65x16", dict(nx=32, ny=65, nz=16, ncly1=2, nclyn=1, istret=3, beta=0.4)),
    # the channel grid's x extent (256: the FP32 strided passes then give
    # each block one of the nx/2+1-th lines) and its y extent
    ("s2_walls_256x129x16", dict(nx=256, ny=129, nz=16, ncly1=2, nclyn=2, istret=2, beta=0.259)),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def ctx(request, torch_cuda):
    import incompact3d_amd as X
    name, kw = request.param
    p = O.Params(**kw)
    X.init(X.make_params(**kw))
    yield name, p, O.Schemes(p), X
    X.finalize()


def _tol(dtype, f32):
    return 1e-12 if dtype == "f64" else f32


def test_metrics(ctx):
    name, p, S, X = ctx
    for key in ("yp", "ypi", "ppy", "ppyi"):
        a = getattr(X.coef, key)
        b = S.stret[key]
        assert a.shape == b.shape
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max(), f"{name} {key}"


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("npaire", [0, 1])
def test_dery(ctx, dtype, npaire):
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    u = seeded_field((p.nz, p.ny, p.nx), seed=17 + npaire, lengths=(p.xlx, p.yly, p.zlz))
    ref = O.der1(S, u, "y", npaire)
    c = S.d1["y"]
    ff, fs, fw = (c["ffp"], c["fsp"], c["fwp"]) if npaire == 1 else (c["ff"], c["fs"], c["fw"])
    ud = torch.tensor(u, dtype=tdt, device="cuda")
    td = torch.empty_like(ud)
    fn = getattr(D, f"dery_{S.bc['y'][0]}{S.bc['y'][1]}")
    fn(td, ud, None, None, ff, fs, fw, X.coef.ppy, p.nx, p.ny, p.nz, npaire)
    err = rel_err(td.double().cpu().numpy(), ref)
    assert err <= _tol(dtype, 2e-5), f"{name} npaire={npaire}: {err:.3e}"


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_dery_caller_metric(ctx, dtype):
    """dery multiplies by the ppy array its caller passes, as the reference
    does, also when that array differs from x3d_init's (another libm, or a
    different mapping): twice in a row, the second call from the cache."""
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    u = seeded_field((p.nz, p.ny, p.nx), seed=31, lengths=(p.xlx, p.yly, p.zlz))
    ppy = S.stret["ppy"]
    mine = ppy * (1.0 + 1e-3 * np.cos(np.arange(p.ny)))
    if dtype == "f32":
        mine = mine.astype(np.float32).astype(np.float64)
    ref = O.der1(S, u, "y", 1) / ppy[None, :, None] * mine[None, :, None]
    c = S.d1["y"]
    ud = torch.tensor(u, dtype=tdt, device="cuda")
    fn = getattr(D, f"dery_{S.bc['y'][0]}{S.bc['y'][1]}")
    for _ in range(2):
        td = torch.empty_like(ud)
        fn(td, ud, None, None, c["ffp"], c["fsp"], c["fwp"], mine, p.nx, p.ny, p.nz, 1)
        err = rel_err(td.double().cpu().numpy(), ref)
        assert err <= _tol(dtype, 2e-5), f"{name}: {err:.3e}"


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_staggered_y(ctx, dtype):
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    st = S.st["y"]
    ny, nym = p.ny, p.nym
    # deryvp: velocity -> pressure grid, times ppyi
    u = seeded_field((p.nz, ny, p.nx), seed=5, lengths=(p.xlx, p.yly, p.zlz))
    ref = O.dervp(S, u, "y", 0)
    ud = torch.tensor(u, dtype=tdt, device="cuda")
    td = torch.empty((p.nz, nym, p.nx), dtype=tdt, device="cuda")
    D.deryvp(td, ud, None, None, st["cfx6"], st["csx6"], st["cwx6"], X.coef.ppyi[:nym], p.nx, ny, nym, p.nz, 0)
    err = rel_err(td.double().cpu().numpy(), ref)
    assert err <= _tol(dtype, 2e-5), f"deryvp {name}: {err:.3e}"
    # derypv: pressure -> velocity grid, times ppy
    q = seeded_field((p.nz, nym, p.nx), seed=6, lengths=(p.xlx, p.yly, p.zlz))
    ref = O.derpv(S, q, "y", 1)
    qd = torch.tensor(q, dtype=tdt, device="cuda")
    td = torch.empty((p.nz, ny, p.nx), dtype=tdt, device="cuda")
    D.derypv(td, qd, None, None, st["cfip6"], st["csip6"], st["cwip6"], st["cfx6"], st["csx6"], st["cwx6"],
             X.coef.ppy, p.nx, nym, ny, p.nz, 1)
    err = rel_err(td.double().cpu().numpy(), ref)
    assert err <= _tol(dtype, 2e-5), f"derypv {name}: {err:.3e}"


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_poisson_stretched(ctx, dtype):
    import torch
    name, p, S, X = ctx
    f = seeded_field((p.nz, p.nym, p.nx), seed=3, lengths=(p.xlx, p.yly, p.zlz))
    f -= f.mean()
    ref = O.poisson(S, f)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    d = torch.tensor(f, dtype=tdt, device="cuda")
    X.poisson(d)
    err = rel_err(d.double().cpu().numpy(), ref)
    assert err <= _tol(dtype, 1e-5), f"{name} {dtype}: {err:.3e}"


def test_poisson_stretched_inverts_div_grad(ctx):
    """Size-independent property: on fields without the Nyquist modes the
    device solve returns p up to a constant from div(grad p) of the stretched
    operators (the divergence is formed by the oracle)."""
    import torch
    import scipy.fft as sf
    name, p, S, X = ctx
    if p.istret == 3:
        pytest.skip("istret = 3: the reference's first matrix row is not the exact operator's")
    pe = np.random.default_rng(9).standard_normal((p.nz, p.nym, p.nx))
    F = sf.dct(np.fft.fft(np.fft.fft(pe, axis=0), axis=2), type=2, axis=1)
    F[p.nz // 2, :, :] = 0.0
    F[:, :, p.nx // 2] = 0.0
    pe = np.fft.ifft(np.fft.ifft(sf.idct(F, type=2, axis=1), axis=0), axis=2).real
    rhs = O.divergence(S, *O.gradp(S, pe))
    d = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    X.poisson(d)
    pr = d.cpu().numpy()
    assert np.abs((pr - pr.mean()) - (pe - pe.mean())).max() <= 1e-10 * np.abs(pe).max(), name


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_deryy_stretched_is_plain(ctx, dtype):
    """deryy applies no metric on a stretched mesh (its caller combines
    pp2y and pp4y, transeq.f90:337-346): the plain second derivative."""
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    u = seeded_field((p.nz, p.ny, p.nx), seed=23, lengths=(p.xlx, p.yly, p.zlz))
    ref = O.der2(S, u, "y", 0)
    c = S.d2["y"]
    ud = torch.tensor(u, dtype=tdt, device="cuda")
    td = torch.empty_like(ud)
    fn = getattr(D, f"deryy_{S.bc['y'][0]}{S.bc['y'][1]}")
    fn(td, ud, None, None, c["sf"], c["ss"], c["sw"], p.nx, p.ny, p.nz, 0)
    err = rel_err(td.double().cpu().numpy(), ref)
    assert err <= _tol(dtype, 2e-5), f"{name}: {err:.3e}"
