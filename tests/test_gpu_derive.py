"""GPU parity of the compact derivatives and staggered operators (derive.f90)
against the CPU oracle, through the C-ABI mirror of the reference signatures.

Bar: max|gpu - oracle| / max|oracle| <= 1e-12 in FP64 (the FP64 parity tolerance),
<= 2e-5 for the FP32 entries (FP32 storage and arithmetic).

The same checks with host (numpy) fields, through the x3d_host_* twins, are
in test_gpu_derive_host.py.
"""
import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import rel_err, seeded_field

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-5

# (name, params): periodic, y walls (config-2 walls/channel layout), free-slip
# everywhere (the reference TGV test case), mixed 1/2 faces.
CONFIGS = [
    ("periodic", dict(nx=64, ny=48, nz=40)),
    ("periodic_small", dict(nx=16, ny=8, nz=24)),
    ("ywalls", dict(nx=32, ny=33, nz=16, ncly1=2, nclyn=2)),
    # 513-point wall lines: blocks of 16 (FP64) / 32 (FP32) lines, so that a
    # wavefront spans several segments of a line
    ("ywalls_long", dict(nx=16, ny=513, nz=8, ncly1=2, nclyn=2)),
    ("freeslip", dict(nx=17, ny=19, nz=21, nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)),
    ("mixed", dict(nx=18, ny=22, nz=20, nclx1=1, nclxn=2, ncly1=2, nclyn=1, nclz1=2, nclzn=2)),
    # the line lengths of the 256^3 step and of the 512^3 / 512x513x512
    # microbenchmarks, with the launch shapes those grids use: x lines of
    # 256 and 512 points (4 and 8 points per lane), y and z lines of 256 and
    # 512 points over several blocks of lines, 513-point wall lines 64 wide
    ("x256_lines", dict(nx=256, ny=16, nz=8)),
    ("x512_lines", dict(nx=512, ny=8, nz=8)),
    ("yz256_lines", dict(nx=64, ny=256, nz=256)),
    ("y512_lines", dict(nx=64, ny=512, nz=8)),
    ("z512_lines", dict(nx=64, ny=8, nz=512)),
    ("ywalls_513x64", dict(nx=64, ny=513, nz=8, ncly1=2, nclyn=2)),
]


def _context(param):
    import incompact3d_amd as X
    name, kw = param
    p = O.Params(**kw)
    X.init(X.make_params(**kw))
    return name, p, O.Schemes(p), X


@pytest.fixture(scope="module", params=CONFIGS, ids=[c[0] for c in CONFIGS])
def ctx(request, torch_cuda):
    c = _context(request.param)
    yield c
    c[3].finalize()


def _dev(a, dtype):
    import torch
    return torch.tensor(a, dtype=dtype, device="cuda")


def _fields(mem, name, u, tdt, oshape=None):
    """(input, output) fields in device or host memory."""
    import torch
    if mem == "host":
        npdt = np.float64 if tdt == torch.float64 else np.float32
        ud = np.ascontiguousarray(u, dtype=npdt)
        return ud, np.empty(oshape if oshape else ud.shape, dtype=npdt)
    ud = _dev(u, tdt)
    return ud, torch.empty(oshape if oshape else ud.shape, dtype=tdt, device="cuda")


def _result(td):
    return td.astype(np.float64) if isinstance(td, np.ndarray) else td.double().cpu().numpy()


def _bcname(S, ax):
    b1, bn = S.bc[ax]
    return f"{b1}{bn}"


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
@pytest.mark.parametrize("kind", ["der1", "der2"])
@pytest.mark.parametrize("npaire", [0, 1])
def test_derivative(ctx, dtype, ax, kind, npaire):
    _check_derivative(ctx, dtype, ax, kind, npaire, "device")


def _check_derivative(ctx, dtype, ax, kind, npaire, mem):
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    shape = (p.nz, p.ny, p.nx)
    u = seeded_field(shape, seed=11 + npaire, lengths=(p.xlx, p.yly, p.zlz))
    if kind == "der1":
        ref = O.der1(S, u, ax, npaire)
        c = S.d1[ax]
        ff, fs, fw = (c["ffp"], c["fsp"], c["fwp"]) if npaire == 1 else (c["ff"], c["fs"], c["fw"])
        fname = f"der{ax}_{_bcname(S, ax)}"
    else:
        ref = O.der2(S, u, ax, npaire)
        c = S.d2[ax]
        ff, fs, fw = (c["sfp"], c["ssp"], c["swp"]) if npaire == 1 else (c["sf"], c["ss"], c["sw"])
        fname = f"der{ax}{ax}_{_bcname(S, ax)}"
    ud, td = _fields(mem, name, u, tdt)
    fn = getattr(D, fname)
    if kind == "der1" and ax == "y":
        fn(td, ud, None, None, ff, fs, fw, np.ones(p.ny), p.nx, p.ny, p.nz, npaire)
    else:
        fn(td, ud, None, None, ff, fs, fw, p.nx, p.ny, p.nz, npaire)
    err = rel_err(_result(td), ref)
    assert err <= (TOL64 if dtype == "f64" else TOL32), f"{fname} {name} npaire={npaire}: {err:.3e}"


def _stag_case(S, p, op, ax, u_shape):
    """Return (oracle result, call) for one staggered operator, with the
    coefficient arrays the reference callers pass (navier.f90:297-431)."""
    st = S.st[ax]
    n, nm = S.n[ax], S.nm[ax]
    return st, n, nm


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
@pytest.mark.parametrize("op", ["dervp0", "dervp1", "intervp", "derpv", "interpv"])
def test_staggered(ctx, dtype, ax, op):
    _check_staggered(ctx, dtype, ax, op, "device")


def _check_staggered(ctx, dtype, ax, op, mem):
    import torch
    from incompact3d_amd import derive as D
    name, p, S, X = ctx
    tdt = torch.float64 if dtype == "f64" else torch.float32
    st = S.st[ax]
    n, nm = S.n[ax], S.nm[ax]
    periodic = S.bc[ax][0] == 0
    dims = {"x": p.nx, "y": p.ny, "z": p.nz}
    if op.startswith("dervp") and not periodic:
        pass
    vp = op in ("dervp0", "dervp1", "intervp")
    in_dims = dict(dims)
    if not vp:
        in_dims[ax] = nm
    shape = (in_dims["z"], in_dims["y"], in_dims["x"])
    u = seeded_field(shape, seed=5, lengths=(p.xlx, p.yly, p.zlz))
    out_dims = dict(in_dims)
    out_dims[ax] = nm if vp else n
    oshape = (out_dims["z"], out_dims["y"], out_dims["x"])
    ud, td = _fields(mem, name, u, tdt, oshape)
    nx, ny, nz = in_dims["x"], in_dims["y"], in_dims["z"]
    ones = np.ones(max(n, nm))
    if op in ("dervp0", "dervp1"):
        npaire = int(op[-1])
        ref = O.dervp(S, u, ax, npaire)
        args = (st["cfx6"], st["csx6"], st["cwx6"])
        if ax == "x":
            D.derxvp(td, ud, None, None, *args, nx, nm, ny, nz, npaire)
        elif ax == "y":
            D.deryvp(td, ud, None, None, *args, ones[:nm], nx, ny, nm, nz, npaire)
        else:
            D.derzvp(td, ud, None, None, *args, nx, ny, nz, nm, npaire)
    elif op == "intervp":
        ref = O.intervp(S, u, ax, 1)
        args = (st["cifxp6"], st["cisxp6"], st["ciwxp6"])
        if ax == "x":
            D.interxvp(td, ud, None, None, *args, nx, nm, ny, nz, 1)
        elif ax == "y":
            D.interyvp(td, ud, None, None, *args, nx, ny, nm, nz, 1)
        else:
            D.interzvp(td, ud, None, None, *args, nx, ny, nz, nm, 1)
    elif op == "derpv":
        ref = O.derpv(S, u, ax, 1)
        args = (st["cfip6"], st["csip6"], st["cwip6"], st["cfx6"], st["csx6"], st["cwx6"])
        if ax == "x":
            D.derxpv(td, ud, None, None, *args, nm, n, ny, nz, 1)
        elif ax == "y":
            D.derypv(td, ud, None, None, *args, ones[:n], nx, nm, n, nz, 1)
        else:
            D.derzpv(td, ud, None, None, *args, nx, ny, nm, n, 1)
    else:
        ref = O.interpv(S, u, ax, 1)
        args = (st["cifip6"], st["cisip6"], st["ciwip6"], st["cifx6"], st["cisx6"], st["ciwx6"])
        if ax == "x":
            D.interxpv(td, ud, None, None, *args, nm, n, ny, nz, 1)
        elif ax == "y":
            D.interypv(td, ud, None, None, *args, nx, nm, n, nz, 1)
        else:
            D.interzpv(td, ud, None, None, *args, nx, ny, nm, n, 1)
    err = rel_err(_result(td), ref)
    assert err <= (TOL64 if dtype == "f64" else TOL32), f"{op} {ax} {name}: {err:.3e}"


def test_library_coefficients_match_oracle(ctx):
    """The arrays x3d_init built are the oracle's (bit-exact), so callers that
    pass the module arrays hit the device copies."""
    name, p, S, X = ctx
    for ax in "xyz":
        assert np.array_equal(getattr(X.coef, f"ff{ax}"), S.d1[ax]["ff"])
        assert np.array_equal(getattr(X.coef, f"fw{ax}p"), S.d1[ax]["fwp"])
        assert np.array_equal(getattr(X.coef, f"ss{ax}"), S.d2[ax]["ss"])
        assert np.array_equal(getattr(X.coef, f"cw{ax}6"), S.st[ax]["cwx6"])
