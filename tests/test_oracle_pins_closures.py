"""Pins of the oracle's non-periodic closures (CPU only), each against a
formulation the oracle was not written from.

Compact derivatives with mixed faces (derx_12 / derx_21, derive.f90:141-277;
derxx_12 / derxx_21, :1666-1974) and every combination of free-slip (ncl=1)
and Dirichlet (ncl=2) faces:
  * a dense restatement assembled from Lele's published scheme (J. Comput.
    Phys. 103, 1992): the 6th-order interior rows, the 3rd-order one-sided
    wall row and the 4th-order row next to it (first derivative), the
    alpha = 11 wall row and the 4th-order Pade row next to it (second
    derivative) at Dirichlet faces, and at free-slip faces the interior rows
    applied to the field mirrored with its parity (and the result mirrored
    with the derivative's parity), solved with numpy.linalg;
  * exactness on the polynomials those closures reproduce, with the parity
    a free-slip face imposes.

Staggered operators on non-periodic lines (derive.f90:3851-3899, 3989-4031,
4096-4116, 4209-4255): the reference extends the input past each face (even
reflection, or u(-1) = 2 u(0) - u(1) for derxvp with npaire = 0) and solves
a system whose face rows extend the output evenly (derpv: oddly).  On cosine
modes of the doubled box, and on sine modes for derxvp with npaire = 0, these
extensions are exact, so each operator returns the periodic closed form:
the staggered modified wavenumber and the interpolation transfer function
(the formulas of poisson.f90:1566-1596, :1716-1731).
"""
import math

import numpy as np
import pytest

from oracle import x3d_oracle as O

# Lele's 6th-order interior scheme (schemes.f90:477-480, :658-673)
AL1, A1, B1 = 1.0 / 3.0, 7.0 / 9.0, 1.0 / 36.0
AL2, A2, B2 = 2.0 / 11.0, 12.0 / 11.0, 3.0 / 44.0


def _dense(n, h, kind, bc1, bcn, npaire):
    """A, B with A t = B u for the compact derivative on a non-periodic line.
    kind 1: first derivative, 2: second derivative.  bc 1: free-slip face
    (mirror with parity), 2: Dirichlet face (one-sided closures)."""
    s = 1.0 if npaire == 1 else -1.0           # parity of u about a free-slip face
    sd = -s if kind == 1 else s                 # parity of its derivative
    A = np.zeros((n, n))
    B = np.zeros((n, n))

    def put(M, i, j, v, par):
        # column j of the extended line, folded back onto the line
        if j < 0:
            M[i, -j] += par * v
        elif j > n - 1:
            M[i, 2 * (n - 1) - j] += par * v
        else:
            M[i, j] += v

    for i in range(n):
        if kind == 1:
            al, st = AL1, {1: A1 / h, 2: B1 / h, -1: -A1 / h, -2: -B1 / h}
        else:
            al = AL2
            st = {0: -2.0 * (A2 + B2) / h ** 2, 1: A2 / h ** 2, -1: A2 / h ** 2, 2: B2 / h ** 2, -2: B2 / h ** 2}
        put(A, i, i, 1.0, 1.0)
        put(A, i, i - 1, al, sd)
        put(A, i, i + 1, al, sd)
        for o, v in st.items():
            put(B, i, i + o, v, s)
    # Dirichlet faces: Lele's boundary rows replace the mirrored ones
    for face, bc in ((0, bc1), (1, bcn)):
        if bc != 2:
            continue
        rows = (0, 1) if face == 0 else (n - 1, n - 2)
        sgn = 1.0 if face == 0 else -1.0         # direction into the line
        r0, r1 = rows
        inw = lambda k: r0 + int(sgn) * k        # noqa: E731  k-th point inward
        A[r0, :] = 0.0
        B[r0, :] = 0.0
        A[r1, :] = 0.0
        B[r1, :] = 0.0
        if kind == 1:
            A[r0, inw(0)], A[r0, inw(1)] = 1.0, 2.0
            for k, v in enumerate((-2.5, 2.0, 0.5)):
                B[r0, inw(k)] = sgn * v / h
            A[r1, inw(0)], A[r1, inw(1)], A[r1, inw(2)] = 0.25, 1.0, 0.25
            B[r1, inw(2)], B[r1, inw(0)] = sgn * 0.75 / h, -sgn * 0.75 / h
        else:
            A[r0, inw(0)], A[r0, inw(1)] = 1.0, 11.0
            for k, v in enumerate((13.0, -27.0, 15.0, -1.0)):
                B[r0, inw(k)] = v / h ** 2
            A[r1, inw(0)], A[r1, inw(1)], A[r1, inw(2)] = 0.1, 1.0, 0.1
            for k, v in enumerate((1.2, -2.4, 1.2)):
                B[r1, inw(k)] = v / h ** 2
    return A, B


MIXED = [(1, 2), (2, 1), (1, 1), (2, 2)]


@pytest.mark.parametrize("bc", MIXED, ids=lambda b: f"{b[0]}{b[1]}")
@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("npaire", [0, 1])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_wall_closures_equal_dense_restatement(bc, kind, npaire, ax):
    n = 19
    kw = {"x": dict(nclx1=bc[0], nclxn=bc[1]), "y": dict(ncly1=bc[0], nclyn=bc[1]),
          "z": dict(nclz1=bc[0], nclzn=bc[1])}[ax]
    dims = {"x": (n, 6, 5), "y": (6, n, 5), "z": (6, 5, n)}[ax]
    p = O.Params(*dims, xlx=1.3, yly=1.7, zlz=2.1, **kw)
    S = O.Schemes(p)
    h = {"x": p.dx, "y": p.dy, "z": p.dz}[ax]
    rng = np.random.default_rng(11 + kind + 2 * npaire)
    u = rng.standard_normal((p.nz, p.ny, p.nx))
    if npaire == 0:
        # an odd field vanishes on a free-slip face; the reference's rows
        # assume it (derxx_11 and the odd set sf(1) = 0 put 0 in the wall row)
        lines = np.moveaxis(u, O.AXIS[ax], 0)
        if bc[0] == 1:
            lines[0] = 0.0
        if bc[1] == 1:
            lines[-1] = 0.0
    got = O.der1(S, u, ax, npaire) if kind == 1 else O.der2(S, u, ax, npaire)
    A, B = _dense(n, h, kind, bc[0], bc[1], npaire)
    M = np.linalg.solve(A, B)
    axis = O.AXIS[ax]
    ref = np.moveaxis(np.tensordot(M, np.moveaxis(u, axis, 0), axes=(1, 0)), 0, axis)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (bc, kind, npaire, ax)


def test_dense_restatement_is_exact_on_polynomials():
    """The dense builder itself: a wrong mirror sign or closure coefficient
    breaks the polynomial exactness these published closures have."""
    n = 21
    h = 0.1
    x = np.arange(n) * h
    L = x[-1]
    for bc1, bcn in MIXED:
        # a free-slip face imposes a parity about itself: use polynomials in
        # the distance to that face (both faces free-slip: the line's centre
        # is no symmetry point, so only even/odd about face 0 is tried when
        # face n is Dirichlet, and vice versa)
        if bc1 == 1 and bcn == 1:
            continue
        c = 0.0 if bc1 == 1 else (L if bcn == 1 else 0.37)
        y = x - c
        cases = [(1, 1, y ** 2, 2 * y), (1, 2, y ** 2, 2 + 0 * y),
                 (0, 1, y, 1 + 0 * y), (0, 1, y ** 3, 3 * y ** 2), (0, 2, y ** 3, 6 * y)]
        if bc1 == 2 and bcn == 2:
            cases += [(0, 1, y ** 2, 2 * y), (0, 2, y ** 2, 2 + 0 * y)]
        for npaire, kind, f, df in cases:
            A, B = _dense(n, h, kind, bc1, bcn, npaire)
            t = np.linalg.solve(A, B @ f)
            assert np.allclose(t, df, atol=1e-9 * max(1.0, np.abs(df).max())), (bc1, bcn, npaire, kind)


@pytest.mark.parametrize("bc", [(1, 2), (2, 1)], ids=lambda b: f"{b[0]}{b[1]}")
def test_mixed_closures_exact_on_polynomials(bc):
    """The oracle's _12 / _21 rows on polynomials in the distance to the
    free-slip face, even (npaire = 1) or odd (npaire = 0) about it."""
    n = 17
    p = O.Params(5, n, 4, yly=1.6, ncly1=bc[0], nclyn=bc[1])
    S = O.Schemes(p)
    yv = np.arange(n) * p.dy
    y = yv - (0.0 if bc[0] == 1 else yv[-1])

    def field(v):
        return np.broadcast_to(v[None, :, None], (p.nz, p.ny, p.nx)).copy()

    for npaire, f, d1, d2 in ((1, y ** 2, 2 * y, 2 + 0 * y), (0, y, 1 + 0 * y, 0 * y),
                              (0, y ** 3, 3 * y ** 2, 6 * y)):
        assert np.allclose(O.der1(S, field(f), "y", npaire), field(d1), atol=1e-11), (bc, npaire)
        assert np.allclose(O.der2(S, field(f), "y", npaire), field(d2), atol=1e-9), (bc, npaire)


# ------------------------------------------------------------------ staggered, non-periodic
ALC, AC, BC = 9.0 / 62.0, 63.0 / 62.0, 17.0 / 62.0 / 3.0      # staggered derivative
ALI = 0.49                                                      # ipinter = 3 interpolation
AI, BI, CI = (75 + 70 * ALI) / 128.0, (126 * ALI - 25) / 256.0, (-10 * ALI + 3) / 256.0


def _kprime(w, h):
    return (2 * AC * math.sin(w / 2) + 2 * BC * math.sin(1.5 * w)) / (1 + 2 * ALC * math.cos(w)) / h


def _transfer(w):
    return (2 * AI * math.cos(w / 2) + 2 * BI * math.cos(1.5 * w) + 2 * CI * math.cos(2.5 * w)) / (
        1 + 2 * ALI * math.cos(w))


@pytest.mark.parametrize("bc", [(1, 1), (2, 2), (1, 2), (2, 1)], ids=lambda b: f"{b[0]}{b[1]}")
@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_nonperiodic_staggered_closed_forms(bc, ax):
    n = 33
    kw = {"x": dict(nclx1=bc[0], nclxn=bc[1]), "y": dict(ncly1=bc[0], nclyn=bc[1]),
          "z": dict(nclz1=bc[0], nclzn=bc[1])}[ax]
    dims = {"x": (n, 4, 3), "y": (4, n, 3), "z": (4, 3, n)}[ax]
    L = 1.9
    p = O.Params(*dims, xlx=L, yly=L, zlz=L, **kw)
    S = O.Schemes(p)
    h = L / (n - 1)
    axis = O.AXIS[ax]
    xv = np.arange(n) * h                    # velocity points, faces at 0 and L
    xp = (np.arange(n - 1) + 0.5) * h        # pressure points

    def field(v, m):
        shape = [1, 1, 1]
        shape[axis] = m
        full = [p.nz, p.ny, p.nx]
        full[axis] = m
        return np.broadcast_to(v.reshape(shape), full).copy()

    for k in (1, 5, 13, 27):
        q = k * math.pi / L                  # a mode of the doubled box 2L
        w = q * h
        kp, T = _kprime(w, h), _transfer(w)
        got = O.intervp(S, field(np.cos(q * xv), n), ax, 1)
        assert np.allclose(got, field(T * np.cos(q * xp), n - 1), atol=1e-13), ("intervp", k)
        got = O.interpv(S, field(np.cos(q * xp), n - 1), ax, 1)
        assert np.allclose(got, field(T * np.cos(q * xv), n), atol=1e-13), ("interpv", k)
        got = O.derpv(S, field(np.cos(q * xp), n - 1), ax, 1)
        assert np.allclose(got, field(-kp * np.sin(q * xv), n), atol=1e-12 * kp), ("derpv", k)
        got = O.dervp(S, field(np.sin(q * xv), n), ax, 0)
        assert np.allclose(got, field(kp * np.cos(q * xp), n - 1), atol=1e-12 * kp), ("dervp", k)


def test_hyperviscous_second_derivative_design_conditions():
    """isecondder = 5 (schemes.f90:674-688, the channel configuration's
    scheme) on periodic lines, pinned by the conditions it is designed to
    meet rather than its coefficient formulas: the modified wavenumber of a
    cosine mode, -lambda h^2, equals pi^2 (1 + nu0/nu) at the grid cutoff
    w = pi and (2 pi/3)^2 (1 + c_nu nu0/nu) at w = 2 pi/3, and approaches
    w^2 at small w with at least fourth-order error."""
    n, L = 48, 2.0 * math.pi
    nu0nu, cnu = 4.0, 0.44
    p = O.Params(n, 8, 8, xlx=L, isecondder=5, nu0nu=nu0nu, cnu=cnu)
    S = O.Schemes(p)
    h = p.dx
    x = np.arange(n) * h

    def kpp(m):
        u = np.broadcast_to(np.cos(m * x)[None, None, :], (8, 8, n)).copy()
        d2 = O.der2(S, u, "x", 0)
        lam = d2[0, 0] / u[0, 0]
        ok = np.abs(u[0, 0]) > 0.5
        return float(np.mean(-lam[ok] * h * h)), float(np.ptp(lam[ok]) * h * h)

    k_pi, spread = kpp(n // 2)               # w = pi
    assert spread < 1e-9 and abs(k_pi - math.pi ** 2 * (1 + nu0nu)) < 1e-10
    k_23, spread = kpp(n // 3)               # w = 2 pi / 3
    assert spread < 1e-9 and abs(k_23 - (2 * math.pi / 3) ** 2 * (1 + cnu * nu0nu)) < 1e-10
    errs = []
    for m in (1, 2):
        w = m * h
        k, _ = kpp(m)
        errs.append(abs(k - w * w) / (w * w))
    assert errs[0] < 1e-4 and errs[1] / errs[0] > 2 ** 4 * 0.9, errs


def _hyperviscous_interior(nu0nu, cnu):
    """Coefficients (alpha, c1..c4) of the isecondder = 5 interior row
    alpha t(i-1) + t(i) + alpha t(i+1) = sum_m c_m (u(i+m) - 2 u(i) + u(i-m)) / h^2,
    solved from its five design conditions (Lamballais, Fortune & Laizet,
    J. Comput. Phys. 230, 2011): sixth-order accuracy (the w^2, w^4 and w^6
    terms of the modified wavenumber), pi^2 (1 + nu0/nu) at w = pi and
    (2 pi/3)^2 (1 + c_nu nu0/nu) at w = 2 pi/3.  The unknowns enter linearly
    once the wavenumber relation is multiplied by 1 + 2 alpha cos w."""
    m = np.arange(1, 5, dtype=float)
    rows, rhs = [], []
    # Taylor: sum c_m m^2 = 1 + 2 alpha, sum c_m m^4 / 12 = alpha, sum c_m m^6 / 360 = alpha / 12
    rows.append(np.r_[-2.0, m ** 2]); rhs.append(1.0)
    rows.append(np.r_[-1.0, m ** 4 / 12.0]); rhs.append(0.0)
    rows.append(np.r_[-1.0 / 12.0, m ** 6 / 360.0]); rhs.append(0.0)
    for w, k in ((math.pi, math.pi ** 2 * (1 + nu0nu)),
                 (2 * math.pi / 3, (2 * math.pi / 3) ** 2 * (1 + cnu * nu0nu))):
        # sum 2 c_m (1 - cos m w) = k (1 + 2 alpha cos w)
        rows.append(np.r_[-2.0 * k * math.cos(w), 2.0 * (1.0 - np.cos(m * w))]); rhs.append(k)
    sol = np.linalg.solve(np.array(rows), np.array(rhs))
    return sol[0], sol[1:]


def _dense_hyperviscous(n, h, bc1, bcn, npaire, al, c):
    """A, B of the isecondder = 5 second derivative on a non-periodic line
    (schemes.f90:694-856, derive.f90:1500-1976): the hyperviscous interior
    row everywhere, mirrored with the field's parity at free-slip faces; at
    a Dirichlet face Lele's alpha = 11 wall row, the fourth-order Pade row
    next to it and Lele's sixth-order interior row on the third and fourth
    points."""
    s = 1.0 if npaire == 1 else -1.0
    A = np.zeros((n, n))
    B = np.zeros((n, n))

    def put(M, i, j, v, par):
        if j < 0:
            M[i, -j] += par * v
        elif j > n - 1:
            M[i, 2 * (n - 1) - j] += par * v
        else:
            M[i, j] += v

    for i in range(n):
        put(A, i, i, 1.0, 1.0)
        put(A, i, i - 1, al, s)
        put(A, i, i + 1, al, s)
        for k, cm in enumerate(c, start=1):
            put(B, i, i, -2.0 * cm / h ** 2, 1.0)
            put(B, i, i + k, cm / h ** 2, s)
            put(B, i, i - k, cm / h ** 2, s)
    for face, bc in ((0, bc1), (1, bcn)):
        if bc != 2:
            continue
        sgn = 1 if face == 0 else -1
        r0 = 0 if face == 0 else n - 1
        inw = lambda k: r0 + sgn * k         # noqa: E731
        for k in range(4):
            A[inw(k), :] = 0.0
            B[inw(k), :] = 0.0
        A[inw(0), inw(0)], A[inw(0), inw(1)] = 1.0, 11.0
        for k, v in enumerate((13.0, -27.0, 15.0, -1.0)):
            B[inw(0), inw(k)] = v / h ** 2
        A[inw(1), inw(0)], A[inw(1), inw(1)], A[inw(1), inw(2)] = 0.1, 1.0, 0.1
        for k, v in enumerate((1.2, -2.4, 1.2)):
            B[inw(1), inw(k)] = v / h ** 2
        for r in (2, 3):
            A[inw(r), inw(r - 1)], A[inw(r), inw(r)], A[inw(r), inw(r + 1)] = AL2, 1.0, AL2
            for o, v in ((0, -2.0 * (A2 + B2)), (1, A2), (-1, A2), (2, B2), (-2, B2)):
                B[inw(r), inw(r + o)] += v / h ** 2
    return A, B


def test_hyperviscous_interior_from_design_conditions_matches_oracle_periodic():
    """The solved coefficients reproduce the oracle's periodic isecondder = 5
    operator on every Fourier mode of a 24-point line."""
    nu0nu, cnu = 4.0, 0.44
    al, c = _hyperviscous_interior(nu0nu, cnu)
    n = 24
    p = O.Params(n, 4, 4, xlx=2 * math.pi, isecondder=5, nu0nu=nu0nu, cnu=cnu)
    S = O.Schemes(p)
    h = p.dx
    x = np.arange(n) * h
    for q in range(n // 2 + 1):
        w = q * h
        kpp = sum(2 * cm * (1 - math.cos(k * w)) for k, cm in enumerate(c, start=1)) / (1 + 2 * al * math.cos(w))
        u = np.broadcast_to(np.cos(q * x)[None, None, :], (4, 4, n)).copy()
        got = O.der2(S, u, "x", 0)
        assert np.abs(got + kpp / h ** 2 * u).max() <= 1e-11 * max(1.0, kpp / h ** 2), q


@pytest.mark.parametrize("bc", MIXED, ids=lambda b: f"{b[0]}{b[1]}")
@pytest.mark.parametrize("npaire", [0, 1])
@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_hyperviscous_wall_closures_equal_dense_restatement(bc, npaire, ax):
    """isecondder = 5 (the channel's scheme) with free-slip and Dirichlet
    faces: the oracle's deriv-2 rows equal the dense restatement above."""
    n, nu0nu, cnu = 21, 4.0, 0.44
    kw = {"x": dict(nclx1=bc[0], nclxn=bc[1]), "y": dict(ncly1=bc[0], nclyn=bc[1]),
          "z": dict(nclz1=bc[0], nclzn=bc[1])}[ax]
    dims = {"x": (n, 5, 4), "y": (5, n, 4), "z": (5, 4, n)}[ax]
    p = O.Params(*dims, xlx=1.3, yly=1.7, zlz=2.1, isecondder=5, nu0nu=nu0nu, cnu=cnu, **kw)
    S = O.Schemes(p)
    h = {"x": p.dx, "y": p.dy, "z": p.dz}[ax]
    rng = np.random.default_rng(31 + npaire)
    u = rng.standard_normal((p.nz, p.ny, p.nx))
    if npaire == 0:
        lines = np.moveaxis(u, O.AXIS[ax], 0)
        if bc[0] == 1:
            lines[0] = 0.0
        if bc[1] == 1:
            lines[-1] = 0.0
    al, c = _hyperviscous_interior(nu0nu, cnu)
    A, B = _dense_hyperviscous(n, h, bc[0], bc[1], npaire, al, c)
    M = np.linalg.solve(A, B)
    axis = O.AXIS[ax]
    ref = np.moveaxis(np.tensordot(M, np.moveaxis(u, axis, 0), axes=(1, 0)), 0, axis)
    got = O.der2(S, u, ax, npaire)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max(), (bc, npaire, ax)


def test_hyperviscous_dense_restatement_planted_mistake_fails():
    """A wrong sign on the Dirichlet face's third-row stencil, planted in the
    restatement, is caught against the oracle."""
    n = 21
    p = O.Params(5, n, 4, yly=1.7, ncly1=2, nclyn=2, isecondder=5)
    S = O.Schemes(p)
    al, c = _hyperviscous_interior(p.nu0nu, p.cnu)
    A, B = _dense_hyperviscous(n, p.dy, 2, 2, 1, al, c)
    B[2, 0] = -B[2, 0]
    u = np.random.default_rng(5).standard_normal((p.nz, p.ny, p.nx))
    ref = np.moveaxis(np.tensordot(np.linalg.solve(A, B), np.moveaxis(u, 1, 0), axes=(1, 0)), 0, 1)
    assert np.abs(O.der2(S, u, "y", 1) - ref).max() > 1e-6 * np.abs(ref).max()
