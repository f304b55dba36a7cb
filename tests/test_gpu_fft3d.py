"""x3d_fft3d_r2c / _c2r (decomp_2d_fft_3d with physical space in z-pencils)
against numpy's FFT, through the C ABI.

* r2c: the unnormalised forward transform exp(-i k.x) of the real pressure-
  grid field, half of the z modes, in the reference's (nx, ny, nz/2+1)
  layout: numpy.fft.fftn restricted to kz <= nz/2.  FP64 1e-12, FP32 1e-5 of
  the largest mode.
* c2r: the unnormalised backward transform of a real field's spectrum:
  N times the field.  FP64 1e-12, FP32 1e-5.
* The context's boundaries do not change the transform (periodic, walls in
  y, free-slip box: the pressure grid is what is transformed).
* At the benchmark's 256^3: round trip c2r(r2c(f)) = N f and Parseval.
"""
import numpy as np
import pytest

from tests.helpers import seeded_field

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")
torch = pytest.importorskip("torch")

GRIDS = {
    "per64x32x16": dict(nx=64, ny=32, nz=16),
    "per16x8x32": dict(nx=16, ny=8, nz=32),
    "ywalls32x33x16": dict(nx=32, ny=33, nz=16, ncly1=2, nclyn=2),
    "freeslip17": dict(nx=17, ny=17, nz=17, nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1),
}


def _pgrid(kw):
    return tuple(kw[f"n{a}"] - (0 if kw.get(f"ncl{a}1", 0) == 0 else 1) for a in "zyx")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("grid", list(GRIDS))
def test_fft3d_matches_numpy(grid, dtype):
    kw = GRIDS[grid]
    rdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    X.init(X.make_params(**kw))
    try:
        shape = _pgrid(kw)
        nz = shape[0]
        f = seeded_field(shape, 11)
        ref = np.fft.fftn(f)[: nz // 2 + 1]
        spec = X.fft3d_r2c(torch.tensor(f, dtype=rdt, device="cuda"))
        got = spec.cpu().numpy()
        assert got.shape == ref.shape
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= tol * scale
        # backward: the spectrum of another real field
        g = seeded_field(shape, 12)
        sg = np.fft.fftn(g)[: nz // 2 + 1]
        back = X.fft3d_c2r(torch.tensor(sg, dtype=spec.dtype, device="cuda"), nz).cpu().numpy()
        n = g.size
        assert np.abs(back - n * g).max() <= tol * n * np.abs(g).max()
        # the input of c2r is left as it was
        inp = torch.tensor(sg, dtype=spec.dtype, device="cuda")
        X.fft3d_c2r(inp, nz)
        assert np.array_equal(inp.cpu().numpy(), sg.astype(inp.cpu().numpy().dtype))
    finally:
        X.finalize()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fft3d_256_round_trip_and_parseval(dtype):
    n = 256
    rdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 2e-5
    X.init(X.make_params(n, n, n))
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        f = torch.rand((n, n, n), generator=g, device="cuda", dtype=torch.float64).to(rdt) * 2 - 1
        spec = X.fft3d_r2c(f)
        back = X.fft3d_c2r(spec, n)
        N = float(n ** 3)
        err = float((back.double() - N * f.double()).abs().max()) / (N * float(f.double().abs().max()))
        assert err <= tol
        # Parseval over the full spectrum: kz = 0 and nz/2 once, the others twice
        w = torch.full((n // 2 + 1, 1, 1), 2.0, dtype=torch.float64, device="cuda")
        w[0] = 1.0
        w[-1] = 1.0
        e_spec = float((spec.abs().double() ** 2 * w).sum())
        e_phys = N * float((f.double() ** 2).sum())
        assert abs(e_spec - e_phys) <= (1e-12 if dtype == "f64" else 1e-5) * e_phys
    finally:
        X.finalize()


def test_fft3d_rejects_other_extents():
    X.init(X.make_params(32, 32, 32))
    try:
        f = torch.zeros((16, 32, 32), dtype=torch.float64, device="cuda")
        with pytest.raises(X.X3DError):
            X.fft3d_r2c(f)
    finally:
        X.finalize()


def test_host_poisson_fft_transposes():
    """x3d_host_poisson(_f32), x3d_host_fft3d_r2c/c2r and the host transposes
    on the 32 x 33 x 16 box with walls in y: the same results as the device
    entries (the host twins stage the arrays through HBM and call them)."""
    import ctypes as C
    kw = GRIDS["ywalls32x33x16"]
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    nzm, nym, nxm = _pgrid(kw)
    X.init(X.make_params(**kw))
    try:
        lib = X._lib.lib()
        f = seeded_field((nzm, nym, nxm), 21)
        f -= f.mean()
        for dt, fn in ((np.float64, lib.x3d_host_poisson), (np.float32, lib.x3d_host_poisson_f32)):
            h = np.ascontiguousarray(f, dtype=dt)
            ref = torch.tensor(h, device="cuda")
            X.poisson(ref)
            X.check(fn(C.c_void_p(h.ctypes.data), X.iref(nxm), X.iref(nym), X.iref(nzm)), "host_poisson")
            assert np.array_equal(h, ref.cpu().numpy())
        spec = np.empty((nzm // 2 + 1, nym, nxm), dtype=np.complex128)
        X.check(lib.x3d_host_fft3d_r2c(C.c_void_p(f.ctypes.data), C.c_void_p(spec.ctypes.data), X.iref(nxm),
                                       X.iref(nym), X.iref(nzm)), "host_fft3d_r2c")
        assert np.array_equal(spec, X.fft3d_r2c(torch.tensor(f, device="cuda")).cpu().numpy())
        back = np.empty_like(f)
        X.check(lib.x3d_host_fft3d_c2r(C.c_void_p(spec.ctypes.data), C.c_void_p(back.ctypes.data), X.iref(nxm),
                                       X.iref(nym), X.iref(nzm)), "host_fft3d_c2r")
        assert np.abs(back - f.size * f).max() <= 1e-12 * f.size * np.abs(f).max()
        u = seeded_field((nz, ny, nx), 22)
        for nm in ("transpose_x_to_y", "transpose_y_to_z", "transpose_z_to_y", "transpose_y_to_x"):
            v = np.empty_like(u)
            X.check(getattr(lib, f"x3d_host_{nm}")(C.c_void_p(u.ctypes.data), C.c_void_p(v.ctypes.data),
                                                   X.iref(nx), X.iref(ny), X.iref(nz)), nm)
            assert np.array_equal(u, v)
    finally:
        X.finalize()
