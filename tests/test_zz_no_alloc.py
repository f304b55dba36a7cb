"""x3d_poisson_async_noalloc against x3d_poisson_async, through the C ABI.

The two entries run the same kernels; the no-alloc one takes its spectral
work array from the library's fixed device scratch.  Two calls in a row on
different grids, the second smaller, so that the scratch holds the first
call's data when the second starts: the results must still be bit-equal to
the allocating entry's.  A grid whose work array exceeds the scratch is
refused with X3D_ERR_UNSUPPORTED before any GPU work.
"""
import ctypes as C
import math

import numpy as np
import pytest

from tests.helpers import seeded_field

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")
torch = pytest.importorskip("torch")


def _solve_both(kw, dtype, seed):
    lib = X._lib.lib()
    p = X.make_params(**kw)
    X.init(p)
    try:
        nxm = kw["nx"] - (0 if kw.get("nclx1", 0) == 0 else 1)
        nym = kw["ny"] - (0 if kw.get("ncly1", 0) == 0 else 1)
        nzm = kw["nz"] - (0 if kw.get("nclz1", 0) == 0 else 1)
        f = seeded_field((nzm, nym, nxm), seed)
        f -= f.mean()
        a = torch.tensor(f, dtype=dtype, device="cuda")
        b = a.clone()
        prec = 0 if dtype == torch.float64 else 1
        X.check(lib.x3d_poisson_async(prec, C.c_void_p(a.data_ptr()), nxm, nym, nzm), "poisson_async")
        X.synchronize()
        X.check(lib.x3d_poisson_async_noalloc(prec, C.c_void_p(b.data_ptr()), nxm, nym, nzm), "noalloc")
        X.synchronize()
        return a.cpu().numpy(), b.cpu().numpy()
    finally:
        X.finalize()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_matches_allocating_entry_twice(dtype):
    dt = torch.float64 if dtype == "f64" else torch.float32
    big = dict(nx=128, ny=128, nz=64)
    small = dict(nx=32, ny=33, nz=16, ncly1=2, nclyn=2, yly=2.0, istret=2, beta=0.259)
    for kw, seed in ((big, 3), (small, 4)):
        ref, got = _solve_both(kw, dt, seed)
        assert np.isfinite(got).all()
        assert np.array_equal(ref, got), (kw, float(np.abs(ref - got).max()))


def test_noalloc_refuses_beyond_the_scratch():
    lib = X._lib.lib()
    X.init(X.make_params(512, 512, 512))
    try:
        t = torch.zeros(1, dtype=torch.float64, device="cuda")
        r = lib.x3d_poisson_async_noalloc(0, C.c_void_p(t.data_ptr()), 512, 512, 512)
        assert r != 0
        assert b"scratch" in lib.x3d_last_error()
    finally:
        X.finalize()
