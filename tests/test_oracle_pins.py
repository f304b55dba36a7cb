"""Pins of the CPU oracle (CPU only).

Every check here compares the oracle with something it was not written from:
closed-form eigenvalues of the circulant compact operators, the compact
systems assembled densely and solved with numpy.linalg, polynomial
exactness, numpy/scipy FFT and DCT closed forms of the Poisson solves, and
the reference's published TGV DNS statistics (tests/golden/tgv_reference.json,
extracted from examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import x3d_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TWOPI = 2.0 * math.acos(-1.0)


def _line_field(p, ax, f):
    """A field that varies only along axis `ax` (values f(coords))."""
    n = {"x": p.nx, "y": p.ny, "z": p.nz}[ax]
    d = {"x": p.dx, "y": p.dy, "z": p.dz}[ax]
    c = np.arange(n) * d
    shape = [1, 1, 1]
    shape[O.AXIS[ax]] = n
    return np.broadcast_to(f(c).reshape(shape), (p.nz, p.ny, p.nx)).copy(), c, shape


# ------------------------------------------------------------------ periodic closed forms
@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_periodic_derivatives_are_circulant_eigen(ax):
    p = O.Params(32, 24, 16)
    S = O.Schemes(p)
    n = {"x": p.nx, "y": p.ny, "z": p.nz}[ax]
    h = {"x": p.dx, "y": p.dy, "z": p.dz}[ax]
    a, b, al = 7.0 / 9.0, 1.0 / 36.0, 1.0 / 3.0            # 6th-order compact (Lele)
    a2, b2, al2 = 12.0 / 11.0, 3.0 / 44.0, 2.0 / 11.0
    for k in (1, 2, n // 2 - 1):
        w = k * h
        u, c, shape = _line_field(p, ax, lambda x: np.sin(k * x))
        kp = (2 * a * math.sin(w) + 2 * b * math.sin(2 * w)) / (1 + 2 * al * math.cos(w)) / h
        assert np.allclose(O.der1(S, u, ax, 0), kp * np.cos(k * c).reshape(shape), rtol=0, atol=1e-12 * kp)
        k2 = (2 * a2 * (1 - math.cos(w)) + 2 * b2 * (1 - math.cos(2 * w))) / (1 + 2 * al2 * math.cos(w)) / h ** 2
        assert np.allclose(O.der2(S, u, ax, 1), -k2 * np.sin(k * c).reshape(shape), rtol=0, atol=1e-12 * k2)


@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_periodic_staggered_closed_forms(ax):
    p = O.Params(32, 24, 16)
    S = O.Schemes(p)
    n = {"x": p.nx, "y": p.ny, "z": p.nz}[ax]
    h = {"x": p.dx, "y": p.dy, "z": p.dz}[ax]
    ac, bc, alc = 63.0 / 62.0, 17.0 / 62.0 / 3.0, 9.0 / 62.0           # staggered derivative
    ali = 0.49                                                          # ipinter=3
    ai, bi, ci = (75 + 70 * ali) / 128.0, (126 * ali - 25) / 256.0, (-10 * ali + 3) / 256.0
    for k in (1, 3, n // 2 - 1):
        w = k * h
        ks = (2 * ac * math.sin(w / 2) + 2 * bc * math.sin(1.5 * w)) / (1 + 2 * alc * math.cos(w)) / h
        T = (2 * ai * math.cos(w / 2) + 2 * bi * math.cos(1.5 * w) + 2 * ci * math.cos(2.5 * w)) / (
            1 + 2 * ali * math.cos(w))
        u, c, shape = _line_field(p, ax, lambda x: np.sin(k * x))
        half = (c + h / 2).reshape(shape)
        assert np.allclose(O.dervp(S, u, ax, 0), ks * np.cos(k * half), atol=1e-12 * ks)
        assert np.allclose(O.intervp(S, u, ax, 1), T * np.sin(k * half), atol=1e-12)
        up, _, _ = _line_field(p, ax, lambda x: np.sin(k * (x + h / 2)))
        assert np.allclose(O.derpv(S, up, ax, 1), ks * np.cos(k * c).reshape(shape), atol=1e-12 * ks)
        assert np.allclose(O.interpv(S, up, ax, 1), T * np.sin(k * c).reshape(shape), atol=1e-12)


# ------------------------------------------------------------------ free-slip = periodic on the doubled box
def test_freeslip_equals_periodic_extension():
    """With ncl=1 an even (odd) field is differentiated as its even (odd)
    periodic extension: cos(k pi x/L) and sin(k pi x/L) have the periodic
    modified wavenumbers of the doubled grid."""
    n = 33
    L = math.pi
    p = O.Params(n, n, n, xlx=L, yly=L, zlz=L, nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)
    S = O.Schemes(p)
    h = p.dx
    a, b, al = 7.0 / 9.0, 1.0 / 36.0, 1.0 / 3.0
    a2, b2, al2 = 12.0 / 11.0, 3.0 / 44.0, 2.0 / 11.0
    for ax in "xyz":
        for k in (1, 4, 11):
            w = k * h
            kp = (2 * a * math.sin(w) + 2 * b * math.sin(2 * w)) / (1 + 2 * al * math.cos(w)) / h
            k2 = (2 * a2 * (1 - math.cos(w)) + 2 * b2 * (1 - math.cos(2 * w))) / (1 + 2 * al2 * math.cos(w)) / h ** 2
            ue, c, shape = _line_field(p, ax, lambda x: np.cos(k * x))
            uo, _, _ = _line_field(p, ax, lambda x: np.sin(k * x))
            cc = c.reshape(shape)
            assert np.allclose(O.der1(S, ue, ax, 1), -kp * np.sin(k * cc), atol=1e-11 * kp)
            assert np.allclose(O.der1(S, uo, ax, 0), kp * np.cos(k * cc), atol=1e-11 * kp)
            assert np.allclose(O.der2(S, ue, ax, 1), -k2 * np.cos(k * cc), atol=1e-11 * k2)
            assert np.allclose(O.der2(S, uo, ax, 0), -k2 * np.sin(k * cc), atol=1e-11 * k2)


# ------------------------------------------------------------------ dense restatement of the wall schemes
def _dense_first_derivative_dirichlet(n, h):
    """The compact system of the Dirichlet closures (Lele 1992 boundary
    schemes: 3rd order at the wall, 4th order next to it, 6th inside)."""
    A = np.zeros((n, n))
    B = np.zeros((n, n))
    for i in range(2, n - 2):
        A[i, i - 1:i + 2] = [1 / 3, 1, 1 / 3]
        B[i, i - 2:i + 3] = np.array([-1 / 36, -7 / 9, 0, 7 / 9, 1 / 36]) / h
    A[0, :2] = [1, 2]
    B[0, :3] = np.array([-5 / 2, 2, 1 / 2]) / h
    A[1, :3] = [1 / 4, 1, 1 / 4]
    B[1, [0, 2]] = np.array([-3 / 4, 3 / 4]) / h
    A[n - 2, n - 3:] = [1 / 4, 1, 1 / 4]
    B[n - 2, [n - 3, n - 1]] = np.array([-3 / 4, 3 / 4]) / h
    A[n - 1, n - 2:] = [2, 1]
    B[n - 1, n - 3:] = np.array([-1 / 2, -2, 5 / 2]) / h
    return A, B


def test_dirichlet_first_derivative_dense_and_polynomials():
    n = 17
    p = O.Params(16, n, 8, ncly1=2, nclyn=2, yly=2.0)
    S = O.Schemes(p)
    rng = np.random.default_rng(0)
    u = rng.standard_normal((p.nz, p.ny, p.nx))
    A, B = _dense_first_derivative_dirichlet(n, p.dy)
    ref = np.einsum("ij,kjl->kil", np.linalg.solve(A, B), u)
    assert np.allclose(O.der1(S, u, "y", 0), ref, rtol=0, atol=1e-11 * np.abs(ref).max())
    # exact for polynomials of degree <= 2 (third-order wall closure)
    y = np.arange(n) * p.dy
    for deg in (0, 1, 2):
        f = np.broadcast_to((y ** deg)[None, :, None], (p.nz, p.ny, p.nx)).copy()
        df = deg * y ** (deg - 1) if deg else 0 * y
        assert np.allclose(O.der1(S, f, "y", 0), np.broadcast_to(df[None, :, None], f.shape), atol=1e-11)


def test_dirichlet_second_derivative_polynomials():
    n = 17
    p = O.Params(16, n, 8, ncly1=2, nclyn=2, yly=2.0)
    S = O.Schemes(p)
    y = np.arange(n) * p.dy
    for deg in (0, 1, 2, 3):
        f = np.broadcast_to((y ** deg)[None, :, None], (p.nz, p.ny, p.nx)).copy()
        d2 = deg * (deg - 1) * y ** (deg - 2) if deg >= 2 else 0 * y
        assert np.allclose(O.der2(S, f, "y", 0), np.broadcast_to(d2[None, :, None], f.shape), atol=1e-9)


def test_thomas_factors_solve_the_compact_system():
    """prepare() + the two sweeps solve A x = r for the assembled matrix."""
    n = 19
    c = O.first_derivative(0.1, n, 2, 2)
    A, _ = _dense_first_derivative_dirichlet(n, 0.1)
    r = np.random.default_rng(1).standard_normal((n, 3))
    x = r.copy()
    O._thomas(x, c["ff"], c["fs"], c["fw"])
    assert np.allclose(A @ x, r, atol=1e-13)


# ------------------------------------------------------------------ Poisson
def test_poisson_000_closed_form_and_known_answer():
    p = O.Params(32, 16, 24)
    S = O.Schemes(p)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((p.nz, p.ny, p.nx))
    f -= f.mean()
    a = O.poisson_000(S, f)
    b = O.poisson_000_closed_form(S, f)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
    pe = rng.standard_normal((p.nz, p.ny, p.nx))
    px, py, pz = O.gradp(S, pe)
    pr = O.poisson(S, O.divergence(S, px, py, pz))
    K = O.waves_kxyz(S)
    ph = np.fft.fftn(pe)
    ph[K < 1e-16] = 0.0
    assert np.abs(pr - np.fft.ifftn(ph).real).max() <= 1e-12 * np.abs(pe).max()


def test_poisson_010_is_dct_solve():
    import scipy.fft as sf
    p = O.Params(16, 17, 8, ncly1=2, nclyn=2)
    S = O.Schemes(p)
    f = np.random.default_rng(3).standard_normal((p.nz, p.nym, p.nx))
    a = O.poisson(S, f)
    F = sf.dct(np.fft.fft(np.fft.fft(f, axis=0), axis=2), type=2, axis=1)
    K = O.waves_kxyz(S)
    Z = K < 1e-16
    Z[p.nz // 2, :, p.nx // 2] = True
    P = np.where(Z, 0.0, F / (-np.where(Z, 1.0, K)))
    b = np.fft.ifft(np.fft.ifft(sf.idct(P, type=2, axis=1), axis=0), axis=2).real
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


FREESLIP = dict(nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)


@pytest.mark.parametrize("zfree", [True, False])
def test_poisson_11x_is_cosine_solve(zfree):
    """poisson_11x restated step by step (reorders, r2c along z, pairings,
    real/imaginary split of K) equals a direct cosine-transform solve with K
    at modes w = pi m / nm (scipy DCT-II), z free-slip or periodic."""
    import scipy.fft as sf
    kw = dict(FREESLIP) if zfree else dict(nclx1=1, nclxn=1, ncly1=1, nclyn=1)
    p = O.Params(17, 9, 13 if zfree else 12, **kw)
    S = O.Schemes(p)
    nz = p.nzm if zfree else p.nz
    f = np.random.default_rng(5).standard_normal((nz, p.nym, p.nxm))
    a = O.poisson(S, f)
    kx2, tx = O._wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, 1)
    ky2, ty = O._wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, 1)
    kz2, tz = O._wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, 1 if zfree else 0)
    K = (kx2[None, None, :] * (ty[None, :, None] * tz[:, None, None]) ** 2
         + ky2[None, :, None] * (tx[None, None, :] * tz[:, None, None]) ** 2
         + kz2[:, None, None] * (tx[None, None, :] * ty[None, :, None]) ** 2)
    F = sf.dctn(f, type=2, axes=(0, 1, 2)) if zfree else np.fft.fft(sf.dctn(f, type=2, axes=(1, 2)), axis=0)
    Z = K < 1e-16
    P = np.where(Z, 0.0, F / (-np.where(Z, 1.0, K)))
    if zfree:
        b = sf.idctn(P, type=2, axes=(0, 1, 2))
    else:
        b = sf.idctn(np.fft.ifft(P, axis=0), type=2, axes=(1, 2)).real
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_poisson_11x_inverts_freeslip_div_grad():
    """On a free-slip box the staggered operators' parities (navier.f90) make
    div(grad p) a cosine-diagonal operator: the solve returns p minus its mean."""
    p = O.Params(17, 9, 13, **FREESLIP)
    S = O.Schemes(p)
    pe = np.random.default_rng(7).standard_normal((p.nzm, p.nym, p.nxm))
    pr = O.poisson(S, O.divergence(S, *O.gradp(S, pe)))
    assert np.abs(pr - (pe - pe.mean())).max() <= 1e-13 * np.abs(pe).max()


@pytest.mark.parametrize("dims", [(17, 8, 12), (9, 16, 6), (33, 4, 8)])
def test_poisson_100_is_cosine_fourier_solve(dims):
    """poisson_100 restated step by step (x reorder, r2c along z, shift
    factors in z and y with the upper-half sign, x pairing, real K) equals
    a direct solve in the cosine (x) by Fourier (y, z) basis: scipy DCT-II
    along x, numpy FFTs along y and z, p_hat = -f_hat / K with K zero below
    1e-16.  A planted sign on the y shift, or a dropped upper-half sign,
    fails it."""
    import scipy.fft as sf
    p = O.Params(*dims, nclx1=1, nclxn=1)
    S = O.Schemes(p)
    f = np.random.default_rng(9).standard_normal((p.nz, p.ny, p.nxm))
    a = O.poisson(S, f)
    kx2, tx = O._wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, 1)
    ky2, ty = O._wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, 0)
    kz2, tz = O._wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, 0)
    K = (kx2[None, None, :] * (ty[None, :, None] * tz[:, None, None]) ** 2
         + ky2[None, :, None] * (tx[None, None, :] * tz[:, None, None]) ** 2
         + kz2[:, None, None] * (tx[None, None, :] * ty[None, :, None]) ** 2)
    F = np.fft.fftn(sf.dct(f, type=2, axis=2), axes=(0, 1))
    Z = K < 1e-16
    P = np.where(Z, 0.0, F / (-np.where(Z, 1.0, K)))
    b = sf.idct(np.fft.ifftn(P, axes=(0, 1)).real, type=2, axis=2)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_poisson_100_inverts_freeslip_x_div_grad():
    """Free-slip faces in x, periodic y and z: div(grad p) of the staggered
    operators is diagonal in the cosine-Fourier basis, and the solve returns
    p minus its mean.  The modes Nyquist in both y and z are in its null
    space (each term of div(grad) carries the mid-point interpolation's
    transfer function in y or z, zero at w = pi), so the field is drawn
    without them."""
    p = O.Params(17, 8, 12, nclx1=1, nclxn=1)
    S = O.Schemes(p)
    pe = np.random.default_rng(8).standard_normal((p.nz, p.ny, p.nxm))
    F = np.fft.fftn(pe, axes=(0, 1))
    F[p.nz // 2, p.ny // 2] = 0.0
    pe = np.fft.ifftn(F, axes=(0, 1)).real
    pr = O.poisson(S, O.divergence(S, *O.gradp(S, pe)))
    assert np.abs(pr - (pe - pe.mean())).max() <= 1e-13 * np.abs(pe).max()


# ------------------------------------------------------------------ stretched y (istret = 1, 2)
@pytest.mark.parametrize("istret", [1, 2, 3])
def test_stretching_metric_is_the_inverse_jacobian(istret):
    """ppy = yly * d(eta)/dy (twice that for istret = 3, the half channel
    whose eta spans half the range): checked against 4th-order differences
    of yp on a fine mesh; yp runs from 0 to yly."""
    p = O.Params(8, 2049, 8, ncly1=2, nclyn=2, istret=istret, beta=0.7)
    st = O.stretching(p)
    yp, deta = st["yp"], st["yeta"][2] - st["yeta"][1]
    j = np.arange(3, p.ny - 3)
    dyp = (-yp[j + 2] + 8.0 * yp[j + 1] - 8.0 * yp[j - 1] + yp[j - 2]) / (12.0 * deta)
    scale = 2.0 if istret == 3 else 1.0
    # eta = 1/2 (istret = 1) is the one branch switch of the mapping
    ok = np.abs(st["yeta"][j] - 0.5) > 3 * deta if istret == 1 else np.ones(j.size, bool)
    r = dyp * st["ppy"][j] / (p.yly * scale)
    assert np.abs(r[ok] - 1.0).max() < 1e-8
    assert yp[0] == 0.0 and abs(yp[-1] - p.yly) < 1e-12 * p.yly
    assert np.all(np.diff(yp) > 0)


def _stretched_operator_modes(S):
    """div(grad .) of the stretched staggered operators on each cosine mode
    of a y-only field, analysed with scipy's DCT-II: the dense spectral
    operator for kx = kz = 0."""
    import scipy.fft as sf
    p = S.p
    M = np.zeros((p.nym, p.nym))
    for m in range(p.nym):
        e = np.zeros(p.nym)
        e[m] = 1.0
        f = np.broadcast_to(sf.idct(e, type=2)[None, :, None], (p.nz, p.nym, p.nx)).copy()
        M[:, m] = sf.dct(O.divergence(S, *O.gradp(S, f))[0, :, 0], type=2)
    return M


@pytest.mark.parametrize("istret", [1, 2])
def test_matrice_refinement_is_the_stretched_operator(istret):
    """The pentadiagonal matrices of matrice_refinement equal, entry by entry,
    the even and odd cosine-mode blocks of div(grad .) built from the
    stretched derivative operators (ppy, ppyi metrics), which couple even and
    odd modes not at all; only the first row of the even block differs, where
    the reference puts the unit row of the mean mode."""
    p = O.Params(4, 17, 4, ncly1=1, nclyn=1, istret=istret, beta=0.7)
    S = O.Schemes(p)
    M = _stretched_operator_modes(S)
    assert max(np.abs(M[0::2, 1::2]).max(), np.abs(M[1::2, 0::2]).max()) < 1e-11
    a, a2 = O.matrice_refinement(S)
    nyh = p.nym // 2

    def dense(B):
        D = np.zeros((nyh, nyh))
        for j in range(nyh):
            for d in range(5):
                if 0 <= j + d - 2 < nyh:
                    D[j, j + d - 2] = B[j, d]
        return D

    Ae, Ao = dense(a[0, :, 0, :]), dense(a2[0, :, 0, :])
    Me, Mo = M[0::2, 0::2], M[1::2, 1::2]
    tol = 1e-11 * np.abs(M).max()
    assert np.abs(Ae[1:] - Me[1:]).max() < tol
    assert np.abs(Ao - Mo).max() < tol
    assert Ae[0, 0] == 1.0 and np.all(Ae[0, 1:] == 0.0)


def test_matrice_refinement_istret3_rows():
    """istret = 3 (one system over all modes): every row but the first
    equals the stretched div(grad .) in cosine modes (here at kx = 1, where
    the first row couples to other modes); the first row's off-diagonals are
    half of the operator's, as the reference writes them without the factor
    2 that its even system of istret 1, 2 carries (poisson.f90:2196-2198,
    :2205-2207 against :2034-2035, :2059)."""
    import scipy.fft as sf
    p = O.Params(8, 17, 4, ncly1=1, nclyn=1, istret=3, beta=0.7)
    S = O.Schemes(p)
    nym = p.nym
    x = np.arange(p.nx)
    cx = np.cos(2.0 * np.pi * x / p.nx)
    M = np.zeros((nym, nym))
    for m in range(nym):
        e = np.zeros(nym)
        e[m] = 1.0
        f = (cx[None, None, :] * sf.idct(e, type=2)[None, :, None]) * np.ones((p.nz, 1, 1))
        r = O.divergence(S, *O.gradp(S, f))
        M[:, m] = sf.dct((r[0] * cx[None, :]).sum(axis=1) / (p.nx / 2), type=2)
    (a,) = O.matrice_refinement(S)
    A = np.zeros((nym, nym))
    for j in range(nym):
        for d in range(5):
            if 0 <= j + d - 2 < nym:
                A[j, j + d - 2] = a[0, j, 1, d]
    tol = 1e-11 * np.abs(M).max()
    assert np.abs(A[1:] - M[1:]).max() < tol
    assert abs(A[0, 0] - M[0, 0]) < tol
    assert np.abs(A[0, 1:3] - 0.5 * M[0, 1:3]).max() < tol and np.abs(M[0, 1:3]).min() > 1e-3


def _no_nyquist(p, f):
    import scipy.fft as sf
    F = sf.dct(np.fft.fft(np.fft.fft(f, axis=0), axis=2), type=2, axis=1)
    F[p.nz // 2, :, :] = 0.0
    F[:, :, p.nx // 2] = 0.0
    return np.fft.ifft(np.fft.ifft(sf.idct(F, type=2, axis=1), axis=0), axis=2).real


@pytest.mark.parametrize("istret,ncl", [(1, 1), (2, 1), (2, 2)])
def test_poisson_010_stretched_inverts_div_grad(istret, ncl):
    """poisson_010 with istret != 0 (matrice_refinement + inversion5_v1)
    inverts div(grad p) of the stretched staggered operators up to a
    constant, for p without the Nyquist modes that the mid-point
    interpolation removes (poisson.f90:846-879)."""
    p = O.Params(16, 33, 8, ncly1=ncl, nclyn=ncl, istret=istret, beta=0.7)
    S = O.Schemes(p)
    pe = _no_nyquist(p, np.random.default_rng(7).standard_normal((p.nz, p.nym, p.nx)))
    pr = O.poisson(S, O.divergence(S, *O.gradp(S, pe)))
    assert np.abs((pr - pr.mean()) - (pe - pe.mean())).max() <= 1e-11 * np.abs(pe).max()


def test_inversion5_v1_solves_dense_pentadiagonal():
    """inversion5_v1 equals numpy.linalg.solve on random diagonally dominant
    pentadiagonal systems."""
    rng = np.random.default_rng(11)
    n = 12
    A = rng.standard_normal((2, n, 3, 5))
    A[..., 2] += 8.0
    A[:, 0, :, 0:2] = 0.0
    A[:, 1, :, 0] = 0.0
    A[:, n - 1, :, 3:5] = 0.0
    A[:, n - 2, :, 4] = 0.0
    e = rng.standard_normal((2, n, 3))
    x = O.inversion5_v1(A, e)
    for k in range(2):
        for i in range(3):
            D = np.zeros((n, n))
            for j in range(n):
                for d in range(5):
                    if 0 <= j + d - 2 < n:
                        D[j, j + d - 2] = A[k, j, i, d]
            assert np.abs(x[k, :, i] - np.linalg.solve(D, e[k, :, i])).max() < 1e-12


# ------------------------------------------------------------------ time integration and TGV physics
def test_ab3_is_exact_for_quadratic_rhs():
    dt = 0.1
    for f in (lambda t: 1.0, lambda t: t, lambda t: t * t):
        var = np.zeros(1)
        hist = [None, None, None]
        exact_start = None
        for it in range(1, 8):
            t = (it - 1) * dt
            hist[0] = np.array([f(t)])
            var = O.intt_ab3(var, hist, it, dt)
            if it == 3:
                exact_start = var.copy()
                t3 = 3 * dt
        # after step 3 AB3 integrates f exactly: var(7dt)-var(3dt) = int_{3dt}^{7dt} f
        import scipy.integrate as si
        val = si.quad(f, 3 * dt, 7 * dt)[0]
        # AB3 uses f at t_n, t_{n-1}, t_{n-2} to integrate over [t_n, t_{n+1}]
        assert abs((var - exact_start)[0] - val) < 1e-12, (var, exact_start, val)
        del t3


def _fixture():
    with open(os.path.join(HERE, "golden", "tgv_reference.json")) as f:
        return json.load(f)


def test_tgv_initial_statistics_match_reference_dns_row0():
    p = O.Params(32, 32, 32, re=1600.0)
    S = O.Schemes(p)
    st = O.tgv_state(p)
    eek, eps, eps2, enst = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    t, ek, eps_t, eps_ref, ens = _fixture()["TGV_Re1600_dns512"]["rows"][0]
    assert t == 0.0
    assert abs(eek - ek) < 1e-14
    assert abs(eps - eps_ref) / eps_ref < 1e-6       # modified-wavenumber error at 32^3
    assert abs(eps2 - eps_ref) / eps_ref < 1e-6
    assert abs(enst - ens) / ens < 1e-6


def test_tgv_32_to_t1_matches_reference_dns():
    """200 AB3 steps (dt=0.005) of the whole oracle step reproduce the
    reference's 512^3 DNS at t=1 (TGV_Re1600.dat) to 1e-5."""
    p = O.Params(32, 32, 32, re=1600.0, dt=0.005)
    S = O.Schemes(p)
    st = O.tgv_state(p)
    for _ in range(200):
        O.step(S, st)
    eek, eps, eps2, enst = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    row = [r for r in _fixture()["TGV_Re1600_dns512"]["rows"] if r[0] == 1.0][0]
    assert abs(eek - row[1]) / row[1] < 1e-5
    assert abs(eps - row[3]) / row[3] < 1e-5
    dmax, dmean = O.div_stats(O.divergence(S, st.ux, st.uy, st.uz))
    assert dmax < 1e-12


def test_freeslip_rk3_matches_reference_ctest():
    """The reference's own CTest (tests/TGV-Taylor-Green-vortex): TGV on a
    65^3 free-slip box of side pi, Re=1600, RK3 (itimescheme=5), dt=0.005,
    statistics every 10 steps written at (itime-1)*dt.  The first two rows
    of its reference_time_evol.dat (committed as
    tests/golden/tgv_freeslip_rk3_time_evol.dat) are matched at the CTest's
    tolerance (numpy assert_allclose default, rtol 1e-7)."""
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "tgv_freeslip_rk3_time_evol.dat"))
    p = O.Params(65, 65, 65, xlx=math.pi, yly=math.pi, zlz=math.pi, re=1600.0, dt=0.005, itimescheme=5,
                 **FREESLIP)
    S = O.Schemes(p)
    ux, uy, uz = O.init_tgv(p)
    st = O.State(ux, uy, uz, [ux.copy(), ux.copy()], [uy.copy(), uy.copy()], [uz.copy(), uz.copy()], 0)
    for it in range(1, 21):
        O.step_rk3(S, st)
        if it % 10 == 0:
            row = ref[it // 10 - 1]
            assert abs((it - 1) * p.dt - row[0]) < 1e-12
            np.testing.assert_allclose(O.tgv_diagnostics(S, st.ux, st.uy, st.uz), row[1:], rtol=1e-7)

