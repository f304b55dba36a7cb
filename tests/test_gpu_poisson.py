"""GPU parity of the pressure Poisson solver (poisson_000 / poisson_100 /
poisson_010 / poisson_11x with z free-slip or periodic) and
of the pencil transposes against the oracle, through the C-ABI.

Bar: 1e-12 relative (FP64); the FP32 entry (FP32 FFTs, FP64 divide) is
checked at 1e-5 relative.  At sizes the oracle cannot reach in a test, the
size-independent property div(grad(poisson(f))) = f on the resolved modes is
checked instead.
"""
import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import rel_err, seeded_field

pytestmark = pytest.mark.gpu

CASES = [
    ("p000_64x32x16", dict(nx=64, ny=32, nz=16)),
    ("p000_128", dict(nx=128, ny=128, nz=128)),
    ("p010_32x33x16", dict(nx=32, ny=33, nz=16, ncly1=2, nclyn=2)),
    ("p010_64x17x32", dict(nx=64, ny=17, nz=32, ncly1=1, nclyn=1)),
    ("p11x_bcz1_33x17x65", dict(nx=33, ny=17, nz=65, nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)),
    ("p11x_bcz0_65x33x32", dict(nx=65, ny=33, nz=32, nclx1=1, nclxn=1, ncly1=1, nclyn=1)),
    ("p11x_bcz1_65", dict(nx=65, ny=65, nz=65, nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)),
    # poisson_100: x non-periodic (free-slip or walls), y and z periodic
    ("p100_33x16x16", dict(nx=33, ny=16, nz=16, nclx1=1, nclxn=1)),
    ("p100_65x32x8", dict(nx=65, ny=32, nz=8, nclx1=2, nclxn=2)),
    ("p100_17x8x64", dict(nx=17, ny=8, nz=64, nclx1=1, nclxn=1)),
    ("p100_129x16x32", dict(nx=129, ny=16, nz=32, nclx1=2, nclxn=1)),
    # the benchmark's grid and the 256- and 512-point passes of every
    # direction (x: r2c/c2r of 128 and 256 complex points; y, z: the strided
    # passes of 256 and 512 points, the z pass with the divide)
    ("p000_256", dict(nx=256, ny=256, nz=256)),
    ("p000_256x16x16", dict(nx=256, ny=16, nz=16)),
    ("p000_16x256x16", dict(nx=16, ny=256, nz=16)),
    ("p000_16x16x256", dict(nx=16, ny=16, nz=256)),
    ("p000_512x16x16", dict(nx=512, ny=16, nz=16)),
    ("p000_16x512x16", dict(nx=16, ny=512, nz=16)),
    ("p000_16x16x512", dict(nx=16, ny=16, nz=512)),
    # poisson_010 with the 512-point lines of the 512x513x512 walls grid
    ("p010_16x513x16", dict(nx=16, ny=513, nz=16, ncly1=2, nclyn=2)),
    ("p010_512x17x16", dict(nx=512, ny=17, nz=16, ncly1=2, nclyn=2)),
    ("p010_16x17x512", dict(nx=16, ny=17, nz=512, ncly1=2, nclyn=2)),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def ctx(request, torch_cuda):
    import incompact3d_amd as X
    name, kw = request.param
    p = O.Params(**kw)
    X.init(X.make_params(**kw))
    yield name, p, O.Schemes(p), X
    X.finalize()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_poisson(ctx, dtype):
    import torch
    name, p, S, X = ctx
    shape = (p.nzm, p.nym, p.nxm)
    f = seeded_field(shape, seed=3, lengths=(p.xlx, p.yly, p.zlz))
    f -= f.mean()
    ref = O.poisson(S, f)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    d = torch.tensor(f, dtype=tdt, device="cuda")
    X.poisson(d)
    err = rel_err(d.double().cpu().numpy(), ref)
    assert err <= (1e-12 if dtype == "f64" else 1e-5), f"{name} {dtype}: {err:.3e}"


def test_transposes(ctx):
    import torch
    name, p, S, X = ctx
    a = torch.randn((p.nz, p.ny, p.nx), dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    X.transpose_x_to_y(a, b)
    X.transpose_y_to_z(b, c)
    X.transpose_z_to_y(c, b)
    X.transpose_y_to_x(b, c)
    assert torch.equal(a, c)
    r = torch.empty((p.nz, p.nx, p.ny), dtype=torch.float64, device="cuda")
    X.reorder_ji(a, r)
    assert torch.equal(r, a.transpose(1, 2).contiguous())
