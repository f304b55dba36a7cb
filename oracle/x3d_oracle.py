"""Numpy FP64 restatement of the Xcompact3d hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may use this module, and only as the checker.  The HIP library never
calls it.

Array convention: a field is a numpy array of shape ``(nz, ny, nx)`` in C
order, which is byte-for-byte the reference's Fortran ``(nx, ny, nz)``
column-major array (i fastest).  Axis 2 is x, axis 1 is y, axis 0 is z.

Every line operator restates the reference's per-line arithmetic in the same
order (right-hand-side stencil, forward elimination with ``fs``, backward
substitution with ``ff``/``fw``, Sherman-Morrison correction for periodic
lines).  The loop over the line index is sequential, and it is vectorised over
all lines at once, exactly like ``dery``'s i-innermost loops.

Parity pins (each checked in ``tests/test_oracle_*.py``):

* coefficients (``prepare``/``first_derivative``/``second_derivative``/
  ``interpolation``): pinned by the LU identity of the Thomas factors (the
  factors reproduce the dense tridiagonal matrix) and by the closed-form
  modified wavenumbers / transfer functions of the periodic operators;
* periodic derivative and staggered operators: pinned by the eigenfunction
  (single Fourier mode) closed forms to ~1e-14 and by dense-matrix solves;
* wall/free-slip closures, every 1/2 face pairing (``_11 _12 _21 _22``):
  pinned by dense solves of the banded system built independently from
  Lele's published scheme and by polynomial exactness; the staggered
  operators on non-periodic lines by their periodic closed forms on cosine
  (and sine) modes of the doubled box;
* the hyperviscous second derivative (isecondder = 5): its interior row by
  the scheme's five design conditions (coefficients solved independently),
  its free-slip and Dirichlet rows by the dense restatement built on them;
* channel (``step_channel``): the laminar Poiseuille profile is a fixed
  point, and the constant-flow-rate shift holds the bulk velocity;
* ``waves`` and ``poisson_000``: pinned by the algebraic identity
  p_hat = -f_hat/K (zero when K < 1e-16) computed with numpy FFTs, and by the
  known answer p = poisson(div(grad p_exact)) -> p_exact - mean;
* ``poisson_010``: pinned by the DCT-II closed form (scipy.fft) and by the
  known answer above;
* ``stretching``: the metric ppy equals yly d(eta)/dy of the mapping yp
  (4th-order differences on a fine mesh);
* ``matrice_refinement``: equal, entry by entry, to the even and odd
  cosine-mode blocks of div(grad .) assembled from the stretched staggered
  operators; ``inversion5_v1``: equal to dense numpy solves; the stretched
  ``poisson_010``: inverts div(grad p) up to a constant;
* TGV diagnostics: pinned by the closed forms at t=0 (KE = 1/8,
  eps = 3 nu / 4, as in examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat row 1)
  and by the analytic enstrophy/eps2 of the initial field;
* AB3: pinned by exactness on polynomial-in-time right-hand sides.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = math.acos(-1.0)          # module_param.f90:534
TWOPI = 2.0 * math.acos(-1.0)  # module_param.f90:535


# ---------------------------------------------------------------------------
# Parameters (parameters.f90:273-304, defaults :678-850)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    nx: int
    ny: int
    nz: int
    xlx: float = TWOPI
    yly: float = TWOPI
    zlz: float = TWOPI
    nclx1: int = 0
    nclxn: int = 0
    ncly1: int = 0
    nclyn: int = 0
    nclz1: int = 0
    nclzn: int = 0
    ifirstder: int = 4
    isecondder: int = 4
    ipinter: int = 3
    re: float = 1600.0
    dt: float = 0.005
    itimescheme: int = 3
    istret: int = 0
    beta: float = 0.259
    nu0nu: float = 4.0
    cnu: float = 0.44

    def __post_init__(self):
        # parameters.f90:272-298
        self.nclx = self.nclx1 == 0 and self.nclxn == 0
        self.ncly = self.ncly1 == 0 and self.nclyn == 0
        self.nclz = self.nclz1 == 0 and self.nclzn == 0
        self.nxm = self.nx if self.nclx else self.nx - 1
        self.nym = self.ny if self.ncly else self.ny - 1
        self.nzm = self.nz if self.nclz else self.nz - 1
        self.dx = self.xlx / float(self.nxm)
        self.dy = self.yly / float(self.nym)
        self.dz = self.zlz / float(self.nzm)
        self.dx2 = self.dx * self.dx
        self.dy2 = self.dy * self.dy
        self.dz2 = self.dz * self.dz
        self.xnu = 1.0 / self.re


# ---------------------------------------------------------------------------
# schemes.f90
# ---------------------------------------------------------------------------
def prepare(b, c, f):
    """Thomas factors (schemes.f90:413-439): s = multipliers, w = 1/pivots."""
    n = len(c)
    w = np.array(c, dtype=np.float64)
    s = np.zeros(n)
    for i in range(1, n):
        s[i] = b[i - 1] / w[i - 1]
        w[i] = w[i] - f[i - 1] * s[i]
    w = 1.0 / w
    return s, w


def first_derivative(d, n, ncl1, ncln, ifirstder=4):
    """schemes.f90:443-599.  Returns scalars and (ff,fs,fw,ffp,fsp,fwp)."""
    o = {}
    if ifirstder == 1:
        alfai, afi, bfi = 0.0, 1.0 / (2.0 * d), 0.0
    elif ifirstder == 4:
        alfai, afi, bfi = 1.0 / 3.0, (7.0 / 9.0) / d, (1.0 / 36.0) / d
    else:
        raise ValueError("ifirstder must be 1 or 4")
    o.update(alfai=alfai, afi=afi, bfi=bfi)
    if ifirstder == 1:
        o.update(alfa1=0.0, af1=0.0, bf1=0.0, cf1=0.0, df1=0.0, alfa2=0.0, af2=0.0,
                 alfam=0.0, afm=0.0, alfan=0.0, afn=0.0, bfn=0.0, cfn=0.0, dfn=0.0)
    else:
        o.update(alfa1=2.0, af1=-(5.0 / 2.0) / d, bf1=2.0 / d, cf1=0.5 / d, df1=0.0,
                 alfa2=1.0 / 4.0, af2=(3.0 / 4.0) / d,
                 alfan=2.0, afn=-(5.0 / 2.0) / d, bfn=2.0 / d, cfn=0.5 / d, dfn=0.0,
                 alfam=1.0 / 4.0, afm=(3.0 / 4.0) / d)
    ff = np.zeros(n); fb = np.zeros(n); fc = np.zeros(n)
    if ncl1 == 0:
        ff[0] = alfai; ff[1] = alfai; fc[0] = 2.0; fc[1] = 1.0; fb[0] = alfai; fb[1] = alfai
    elif ncl1 == 1:
        ff[0] = alfai + alfai; ff[1] = alfai; fc[0] = 1.0; fc[1] = 1.0; fb[0] = alfai; fb[1] = alfai
    elif ncl1 == 2:
        ff[0] = o["alfa1"]; ff[1] = o["alfa2"]; fc[0] = 1.0; fc[1] = 1.0
        fb[0] = o["alfa2"]; fb[1] = alfai
    if ncln == 0:
        ff[n - 3] = alfai; ff[n - 2] = alfai; ff[n - 1] = 0.0
        fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0 + alfai * alfai
        fb[n - 3] = alfai; fb[n - 2] = alfai; fb[n - 1] = 0.0
    elif ncln == 1:
        ff[n - 3] = alfai; ff[n - 2] = alfai; ff[n - 1] = 0.0
        fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0
        fb[n - 3] = alfai; fb[n - 2] = alfai + alfai; fb[n - 1] = 0.0
    elif ncln == 2:
        ff[n - 3] = alfai; ff[n - 2] = o["alfam"]; ff[n - 1] = 0.0
        fc[n - 3] = 1.0; fc[n - 2] = 1.0; fc[n - 1] = 1.0
        fb[n - 3] = o["alfam"]; fb[n - 2] = o["alfan"]; fb[n - 1] = 0.0
    for i in range(2, n - 3):
        ff[i] = alfai; fc[i] = 1.0; fb[i] = alfai
    ffp = ff.copy()
    fs, fw = prepare(fb, fc, ff)
    if ncl1 == 1:
        ffp[0] = 0.0
    if ncln == 1:
        fb[n - 2] = 0.0
    fsp, fwp = prepare(fb, fc, ffp)
    o.update(ff=ff, fs=fs, fw=fw, ffp=ffp, fsp=fsp, fwp=fwp)
    return o


def second_derivative(d2, n, ncl1, ncln, isecondder=4, nu0nu=4.0, cnu=0.44):
    """schemes.f90:602-856.  Returns scalars and (sf,ss,sw,sfp,ssp,swp)."""
    o = {}
    if isecondder == 1:
        alsai, asi, bsi, csi, dsi = 0.0, 1.0 / d2, 0.0, 0.0, 0.0
    elif isecondder == 4:
        alsai = 2.0 / 11.0
        asi = (12.0 / 11.0) / d2
        bsi = (3.0 / 44.0) / d2
        csi = 0.0
        dsi = 0.0
    elif isecondder == 5:
        dpis3 = 2.0 * PI / 3.0
        kppkc = PI * PI * (1.0 + nu0nu)
        kppkm = dpis3 * dpis3 * (1.0 + cnu * nu0nu)
        xnpi2, xmpi2 = kppkc, kppkm
        den = 405.0 * xnpi2 - 640.0 * xmpi2 + 144.0
        alsai = 0.5 - (320.0 * xmpi2 - 1296.0) / den
        asi = -(4329.0 * xnpi2 / 8.0 - 32.0 * xmpi2 - 140.0 * xnpi2 * xmpi2 + 286.0) / den / d2
        bsi = (2115.0 * xnpi2 - 1792.0 * xmpi2 - 280.0 * xnpi2 * xmpi2 + 1328.0) / den / (4.0 * d2)
        csi = -(7695.0 * xnpi2 / 8.0 + 288.0 * xmpi2 - 180.0 * xnpi2 * xmpi2 - 2574.0) / den / (9.0 * d2)
        dsi = (198.0 * xnpi2 + 128.0 * xmpi2 - 40.0 * xnpi2 * xmpi2 - 736.0) / den / (4.0 ** 2 * d2)
    else:
        raise ValueError("isecondder must be 1, 4 or 5")
    o.update(alsai=alsai, asi=asi, bsi=bsi, csi=csi, dsi=dsi)
    o.update(alsa1=11.0, as1=13.0 / d2, bs1=-27.0 / d2, cs1=15.0 / d2, ds1=-1.0 / d2)
    if isecondder == 1:
        o.update(alsa2=0.0, as2=1.0 / d2, alsam=0.0, asm=1.0 / d2)
    else:
        o.update(alsa2=0.1, as2=(6.0 / 5.0) / d2, alsam=0.1, asm=(6.0 / 5.0) / d2)
    o.update(alsa3=2.0 / 11.0, as3=(12.0 / 11.0) / d2, bs3=(3.0 / 44.0) / d2)
    o.update(alsa4=2.0 / 11.0, as4=(12.0 / 11.0) / d2, bs4=(3.0 / 44.0) / d2, cs4=0.0)
    o.update(alsan=11.0, asn=13.0 / d2, bsn=-27.0 / d2, csn=15.0 / d2, dsn=-1.0 / d2)
    o.update(alsat=2.0 / 11.0, ast=(12.0 / 11.0) / d2, bst=(3.0 / 44.0) / d2)
    o.update(alsatt=2.0 / 11.0, astt=(12.0 / 11.0) / d2, bstt=(3.0 / 44.0) / d2, cstt=0.0)
    sf = np.zeros(n); sb = np.zeros(n); sc = np.zeros(n)
    if ncl1 == 0:
        sf[0:4] = alsai; sc[0] = 2.0; sc[1:4] = 1.0; sb[0:4] = alsai
    elif ncl1 == 1:
        sf[0] = alsai + alsai; sf[1:4] = alsai; sc[0:4] = 1.0; sb[0:4] = alsai
    elif ncl1 == 2:
        sf[0] = o["alsa1"]; sf[1] = o["alsa2"]; sf[2] = o["alsa3"]; sf[3] = o["alsa4"]
        sc[0:4] = 1.0
        sb[0] = o["alsa2"]; sb[1] = o["alsa3"]; sb[2] = o["alsa4"]; sb[3] = alsai
    if ncln == 0:
        sf[n - 5:n - 1] = alsai; sf[n - 1] = 0.0
        sc[n - 5:n - 1] = 1.0; sc[n - 1] = 1.0 + alsai * alsai
        sb[n - 5:n - 1] = alsai; sb[n - 1] = 0.0
    elif ncln == 1:
        sf[n - 5:n - 1] = alsai; sf[n - 1] = 0.0
        sc[n - 5:n] = 1.0
        sb[n - 5:n - 2] = alsai; sb[n - 2] = alsai + alsai; sb[n - 1] = 0.0
    elif ncln == 2:
        sf[n - 5] = alsai; sf[n - 4] = o["alsatt"]; sf[n - 3] = o["alsat"]; sf[n - 2] = o["alsam"]
        sf[n - 1] = 0.0
        sc[n - 5:n] = 1.0
        sb[n - 5] = o["alsatt"]; sb[n - 4] = o["alsat"]; sb[n - 3] = o["alsam"]; sb[n - 2] = o["alsan"]
        sb[n - 1] = 0.0
    for i in range(4, n - 5):
        sf[i] = alsai; sc[i] = 1.0; sb[i] = alsai
    sfp = sf.copy()
    if ncl1 == 1:
        sf[0] = 0.0
    ss, sw = prepare(sb, sc, sf)
    ssp, swp = prepare(sb, sc, sfp)
    if ncln == 1:
        sb[n - 2] = 0.0
        ss, sw = prepare(sb, sc, sf)
    o.update(sf=sf, ss=ss, sw=sw, sfp=sfp, ssp=ssp, swp=swp)
    return o


def interpolation(dx, nxm, nx, ncl1, ncln, ifirstder=4, ipinter=3):
    """schemes.f90:860-1066 (staggered derivative + mid-point interpolation)."""
    o = {}
    if ifirstder == 1:
        alcai, aci, bci = 0.0, 1.0 / dx, 0.0
    else:
        alcai = 9.0 / 62.0
        aci = (63.0 / 62.0) / dx
        bci = (17.0 / 62.0) / 3.0 / dx
    if ifirstder == 1:
        ailcai, aici, bici, cici, dici = 0.0, 0.5, 0.0, 0.0, 0.0
    elif ipinter == 1:
        ailcai, aici, bici, cici, dici = 3.0 / 10.0, 3.0 / 4.0, 1.0 / (2.0 * 10.0), 0.0, 0.0
    elif ipinter == 2:
        ailcai = 0.461658
        dici = 0.00293016
        aici = 1.0 / 64.0 * (75.0 + 70.0 * ailcai - 320.0 * dici)
        bici = 1.0 / 128.0 * (126.0 * ailcai - 25.0 + 1152.0 * dici)
        cici = 1.0 / 128.0 * (-10.0 * ailcai + 3.0 - 640.0 * dici)
        aici /= 2.0; bici /= 2.0; cici /= 2.0; dici /= 2.0
    elif ipinter == 3:
        ailcai = 0.49
        aici = 1.0 / 128.0 * (75.0 + 70.0 * ailcai)
        bici = 1.0 / 256.0 * (126.0 * ailcai - 25.0)
        cici = 1.0 / 256.0 * (-10.0 * ailcai + 3.0)
        dici = 0.0
    else:
        raise ValueError("ipinter must be 1, 2 or 3")
    o.update(alcai=alcai, aci=aci, bci=bci, ailcai=ailcai, aici=aici, bici=bici,
             cici=cici, dici=dici)

    def tri_m(alpha):
        f = np.zeros(nxm); c = np.zeros(nxm); b = np.zeros(nxm)
        f[0] = alpha; f[1] = alpha; f[nxm - 3] = alpha; f[nxm - 2] = alpha; f[nxm - 1] = 0.0
        c[0] = 2.0 if ncl1 == 0 else 1.0 + alpha
        c[1] = 1.0; c[nxm - 3] = 1.0; c[nxm - 2] = 1.0
        c[nxm - 1] = 1.0 + alpha * alpha if ncln == 0 else 1.0 + alpha
        b[0] = alpha; b[1] = alpha; b[nxm - 3] = alpha; b[nxm - 2] = alpha; b[nxm - 1] = 0.0
        for i in range(2, nxm - 3):
            f[i] = alpha; c[i] = 1.0; b[i] = alpha
        return f, c, b

    def tri_n(alpha):
        f = np.zeros(nx); c = np.zeros(nx); b = np.zeros(nx)
        f[0] = alpha + alpha; f[1] = alpha; f[nx - 3] = alpha; f[nx - 2] = alpha; f[nx - 1] = 0.0
        c[0] = 1.0; c[1] = 1.0; c[nx - 3] = 1.0; c[nx - 2] = 1.0; c[nx - 1] = 1.0
        b[0] = alpha; b[1] = alpha; b[nx - 3] = alpha; b[nx - 2] = alpha + alpha; b[nx - 1] = 0.0
        for i in range(2, nx - 3):
            f[i] = alpha; c[i] = 1.0; b[i] = alpha
        return f, c, b

    cfx6, ccx6, cbx6 = tri_m(alcai)
    cfi6, cci6, cbi6 = tri_n(alcai)
    cifx6, cicx6, cibx6 = tri_m(ailcai)
    cifi6, cici6, cibi6 = tri_n(ailcai)
    cfxp6 = cfx6.copy(); cifxp6 = cifx6.copy()
    cifip6 = cifi6.copy(); cfip6 = cfi6.copy()
    cfxp6[0] = 0.0
    cfip6[0] = 0.0
    csx6, cwx6 = prepare(cbx6, ccx6, cfx6)
    csxp6, cwxp6 = prepare(cbx6, ccx6, cfxp6)
    cisx6, ciwx6 = prepare(cibx6, cicx6, cifx6)
    cisxp6, ciwxp6 = prepare(cibx6, cicx6, cifxp6)
    csi6, cwi6 = prepare(cbi6, cci6, cfi6)
    csip6, cwip6 = prepare(cbi6, cci6, cfip6)
    cisi6, ciwi6 = prepare(cibi6, cici6, cifi6)
    cisip6, ciwip6 = prepare(cibi6, cici6, cifip6)
    if ncln in (1, 2):
        cbx6[nxm - 2] = 0.0
        cibx6[nxm - 1] = 0.0
        cbi6[nx - 2] = 0.0
        cibi6[nx - 1] = 0.0
        csxp6, cwxp6 = prepare(cbx6, ccx6, cfxp6)
        cisxp6, ciwxp6 = prepare(cibx6, cicx6, cifxp6)
        csip6, cwip6 = prepare(cbi6, cci6, cfip6)
        cisip6, ciwip6 = prepare(cibi6, cici6, cifip6)
    o.update(cfx6=cfx6, csx6=csx6, cwx6=cwx6, cfxp6=cfxp6, csxp6=csxp6, cwxp6=cwxp6,
             cifx6=cifx6, cisx6=cisx6, ciwx6=ciwx6, cifxp6=cifxp6, cisxp6=cisxp6,
             ciwxp6=ciwxp6, cfi6=cfi6, csi6=csi6, cwi6=cwi6, cfip6=cfip6, csip6=csip6,
             cwip6=cwip6, cifi6=cifi6, cisi6=cisi6, ciwi6=ciwi6, cifip6=cifip6,
             cisip6=cisip6, ciwip6=ciwip6)
    return o


class Schemes:
    """The coefficient state that schemes() leaves in derivX/Y/Z (schemes.f90:7-409)."""

    def __init__(self, p: Params):
        self.p = p
        self.d1 = {}
        self.d2 = {}
        self.st = {}
        for ax, d, d2, n, nm, c1, cn in (
                ("x", p.dx, p.dx2, p.nx, p.nxm, p.nclx1, p.nclxn),
                ("y", p.dy, p.dy2, p.ny, p.nym, p.ncly1, p.nclyn),
                ("z", p.dz, p.dz2, p.nz, p.nzm, p.nclz1, p.nclzn)):
            self.d1[ax] = first_derivative(d, n, c1, cn, p.ifirstder)
            self.d2[ax] = second_derivative(d2, n, c1, cn, p.isecondder, p.nu0nu, p.cnu)
            self.st[ax] = interpolation(d, nm, n, c1, cn, p.ifirstder, p.ipinter)
            self.st[ax]["n"] = n
            self.st[ax]["nm"] = nm
        self.bc = {"x": (p.nclx1, p.nclxn), "y": (p.ncly1, p.nclyn), "z": (p.nclz1, p.nclzn)}
        self.n = {"x": p.nx, "y": p.ny, "z": p.nz}
        self.nm = {"x": p.nxm, "y": p.nym, "z": p.nzm}
        # mesh stretching in y (xcompact3d init: stretching before schemes)
        self.stret = stretching(p) if p.istret != 0 else None


# ---------------------------------------------------------------------------
# stretching.f90: the y mapping and its metrics
# ---------------------------------------------------------------------------
def stretching(p: Params):
    """yp, ypi, ppy, pp2y, pp4y, ppyi, pp2yi, pp4yi and alpha
    (stretching.f90:92-318).  alpha = |xnum/den| is never zero for beta > 0,
    so only the alpha != 0 branch (:143-226, :247-274) is restated."""
    ny, yly, beta, ist = p.ny, p.yly, p.beta, p.istret
    nym = p.nym
    yinf = -yly / 2.0
    den = 2.0 * beta * yinf
    xnum = -yinf - math.sqrt(PI * PI * beta * beta + yinf * yinf)
    alpha = abs(xnum / den)
    if alpha == 0.0:
        raise NotImplementedError("stretching: alpha == 0 branch")

    def mapped(eta):
        den1 = math.sqrt(alpha * beta + 1.0)
        xn = den1 / math.sqrt(alpha / PI) / math.sqrt(beta) / math.sqrt(PI)
        dn = 2.0 * math.sqrt(alpha / PI) * math.sqrt(beta) * PI * math.sqrt(PI)
        den3 = ((math.sin(PI * eta)) * (math.sin(PI * eta)) / beta / PI) + alpha / PI
        den4 = 2.0 * alpha * beta - math.cos(2.0 * PI * eta) + 1.0
        xnum1 = (math.atan(xn * math.tan(PI * eta))) * den4 / den1 / den3 / dn
        cst = math.sqrt(beta) * PI / (2.0 * math.sqrt(alpha) * math.sqrt(alpha * beta + 1.0))
        if eta < 0.5:
            v = xnum1 - cst
        elif eta == 0.5:
            v = 0.0
        else:
            v = xnum1 + cst
        if ist == 1:
            return v - yinf
        if ist == 2:
            return v + yly
        return (v + yly) * 2.0

    yeta = np.zeros(ny)
    yp = np.zeros(ny)
    yeta[0] = 0.0 if ist == 1 else -0.5
    for j in range(1, ny):          # Fortran j = 2..ny
        if ist == 1:
            yeta[j] = float(j) * (1.0 / nym)
        elif ist == 2:
            yeta[j] = float(j) * (1.0 / nym) - 0.5
        else:
            yeta[j] = float(j) * (0.5 / nym) - 0.5
        yp[j] = mapped(yeta[j])
    yetai = np.zeros(ny)
    ypi = np.zeros(ny)
    for j in range(ny):             # Fortran j = 1..ny
        if ist == 1:
            yetai[j] = (float(j + 1) - 0.5) * (1.0 / nym)
        elif ist == 2:
            yetai[j] = (float(j + 1) - 0.5) * (1.0 / nym) - 0.5
        else:
            yetai[j] = (float(j + 1) - 0.5) * (0.5 / nym) - 0.5
        ypi[j] = mapped(yetai[j])

    def metrics(eta):
        s_, c_ = np.sin(PI * eta), np.cos(PI * eta)
        pp = yly * (alpha / PI + (1.0 / PI / beta) * s_ * s_)
        p4 = (-2.0 / beta * c_ * s_)
        if ist == 3:
            p4 = p4 / 2.0
        return pp, pp * pp, p4

    ppy, pp2y, pp4y = metrics(yeta)
    ppyi, pp2yi, pp4yi = metrics(yetai)
    return dict(alpha=alpha, yeta=yeta, yetai=yetai, yp=yp, ypi=ypi, ppy=ppy, pp2y=pp2y, pp4y=pp4y,
                ppyi=ppyi, pp2yi=pp2yi, pp4yi=pp4yi)


AXIS = {"x": 2, "y": 1, "z": 0}


# ---------------------------------------------------------------------------
# line machinery
# ---------------------------------------------------------------------------
def _lines(u, axis):
    """(n, m) contiguous copy with the line index first."""
    v = np.moveaxis(u, axis, 0)
    return np.ascontiguousarray(v).reshape(v.shape[0], -1), v.shape


def _unlines(t, shape, axis):
    return np.ascontiguousarray(np.moveaxis(t.reshape(shape), 0, axis))


def _thomas(t, ff, fs, fw):
    """Forward elimination + back substitution, derive.f90:45-54 (and every variant)."""
    n = t.shape[0]
    for i in range(1, n):
        t[i] = t[i] - t[i - 1] * fs[i]
    t[n - 1] = t[n - 1] * fw[n - 1]
    for i in range(n - 2, -1, -1):
        t[i] = (t[i] - ff[i] * t[i + 1]) * fw[i]


def _thomas_cyclic(t, ff, fs, fw, alfa):
    """Periodic solve with the Sherman-Morrison vector r (derive.f90:18-59).

    r starts as (-1, 0, ..., 0, alfa) for every line; it is data independent,
    so it is carried as one vector with the same scalar recurrences.
    """
    n = t.shape[0]
    r = np.zeros(n)
    r[0] = -1.0
    r[n - 1] = alfa
    for i in range(1, n):
        t[i] = t[i] - t[i - 1] * fs[i]
        r[i] = r[i] - r[i - 1] * fs[i]
    t[n - 1] = t[n - 1] * fw[n - 1]
    r[n - 1] = r[n - 1] * fw[n - 1]
    for i in range(n - 2, -1, -1):
        t[i] = (t[i] - ff[i] * t[i + 1]) * fw[i]
        r[i] = (r[i] - ff[i] * r[i + 1]) * fw[i]
    s = (t[0] - alfa * t[n - 1]) / (1.0 + r[0] - alfa * r[n - 1])
    for i in range(n):
        t[i] = t[i] - s * r[i]


def sm_vector(ff, fs, fw, alfa):
    """The Sherman-Morrison solution vector r of _thomas_cyclic."""
    t = np.zeros((len(ff), 1))
    n = len(ff)
    r = np.zeros(n)
    r[0] = -1.0
    r[n - 1] = alfa
    for i in range(1, n):
        r[i] = r[i] - r[i - 1] * fs[i]
    r[n - 1] = r[n - 1] * fw[n - 1]
    for i in range(n - 2, -1, -1):
        r[i] = (r[i] - ff[i] * r[i + 1]) * fw[i]
    del t
    return r


# ---------------------------------------------------------------------------
# right-hand-side stencils (derive.f90)
# ---------------------------------------------------------------------------
def _rhs_der1(U, c, bc1, bcn, npaire):
    """First-derivative explicit stencils: derx_00/_11/_12/_21/_22 (derive.f90:7-321)."""
    a, b = c["afi"], c["bfi"]
    n = U.shape[0]
    T = np.empty_like(U)
    if bc1 == 0:
        for i in range(n):
            T[i] = a * (U[(i + 1) % n] - U[(i - 1) % n]) + b * (U[(i + 2) % n] - U[(i - 2) % n])
        return T
    for i in range(2, n - 2):
        T[i] = a * (U[i + 1] - U[i - 1]) + b * (U[i + 2] - U[i - 2])
    if bc1 == 1:
        if npaire == 1:
            T[0] = 0.0
            T[1] = a * (U[2] - U[0]) + b * (U[3] - U[1])
        else:
            T[0] = a * (U[1] + U[1]) + b * (U[2] + U[2])
            T[1] = a * (U[2] - U[0]) + b * (U[3] + U[1])
    else:
        T[0] = c["af1"] * U[0] + c["bf1"] * U[1] + c["cf1"] * U[2]
        T[1] = c["af2"] * (U[2] - U[0])
    if bcn == 1:
        if npaire == 1:
            T[n - 2] = a * (U[n - 1] - U[n - 3]) + b * (U[n - 2] - U[n - 4])
            T[n - 1] = 0.0
        else:
            T[n - 2] = a * (U[n - 1] - U[n - 3]) + b * ((-U[n - 2]) - U[n - 4])
            T[n - 1] = a * ((-U[n - 2]) - U[n - 2]) + b * ((-U[n - 3]) - U[n - 3])
    else:
        T[n - 2] = c["afm"] * (U[n - 1] - U[n - 3])
        T[n - 1] = (-c["afn"] * U[n - 1]) - c["bfn"] * U[n - 2] - c["cfn"] * U[n - 3]
    return T


def _d2(a, up, u0, um):
    # asix*(ux(i+1)-ux(i)-ux(i)+ux(i-1)) evaluation order
    return a * (up - u0 - u0 + um)


def _rhs_der2(U, c, bc1, bcn, npaire):
    """Second-derivative stencils: derxx_00/_11/_12/_21/_22 (derive.f90:1354-2048)."""
    A, B, C, D = c["asi"], c["bsi"], c["csi"], c["dsi"]
    n = U.shape[0]
    T = np.empty_like(U)

    def interior(i, g):
        return (_d2(A, g(i + 1), U[i], g(i - 1)) + _d2(B, g(i + 2), U[i], g(i - 2))
                + _d2(C, g(i + 3), U[i], g(i - 3)) + _d2(D, g(i + 4), U[i], g(i - 4)))

    if bc1 == 0:
        def gp(j):
            return U[j % n]
        for i in range(n):
            T[i] = interior(i, gp)
        return T

    def g(j):
        return U[j]
    for i in range(4, n - 4):
        T[i] = interior(i, g)
    u = U
    # ---- left closure
    if bc1 == 1 and npaire == 1:
        def ge(j):  # even reflection about index 0 (derxx_11 npaire=1 rows 1-4)
            return U[-j] if j < 0 else U[j]
        for i in range(4):
            T[i] = interior(i, ge)
    elif bc1 == 1:
        T[0] = 0.0
        T[1] = (_d2(A, u[2], u[1], u[0]) + B * (u[3] - u[1] - u[1] - u[1])
                + C * (u[4] - u[1] - u[1] - u[2]) + D * (u[5] - u[1] - u[1] - u[3]))
        T[2] = (_d2(A, u[3], u[2], u[1]) + _d2(B, u[4], u[2], u[0])
                + C * (u[5] - u[2] - u[2] - u[1]) + D * (u[6] - u[2] - u[2] - u[2]))
        T[3] = (_d2(A, u[4], u[3], u[2]) + _d2(B, u[5], u[3], u[1])
                + C * (u[6] - u[3] - u[3] - u[0]) + D * (u[7] - u[3] - u[3] - u[1]))
    else:
        T[0] = c["as1"] * u[0] + c["bs1"] * u[1] + c["cs1"] * u[2] + c["ds1"] * u[3]
        T[1] = _d2(c["as2"], u[2], u[1], u[0])
        T[2] = _d2(c["as3"], u[3], u[2], u[1]) + _d2(c["bs3"], u[4], u[2], u[0])
        T[3] = (_d2(c["as4"], u[4], u[3], u[2]) + _d2(c["bs4"], u[5], u[3], u[1])
                + _d2(c["cs4"], u[6], u[3], u[0]))
    # ---- right closure
    if bcn == 1 and npaire == 1:
        def gr(j):  # even reflection about index n-1
            return U[2 * (n - 1) - j] if j > n - 1 else U[j]
        for i in range(n - 4, n):
            T[i] = interior(i, gr)
    elif bcn == 1:
        T[n - 4] = (_d2(A, u[n - 3], u[n - 4], u[n - 5]) + _d2(B, u[n - 2], u[n - 4], u[n - 6])
                    + C * (-u[n - 1] - u[n - 4] - u[n - 4] + u[n - 7])
                    + D * (-u[n - 2] - u[n - 4] - u[n - 4] + u[n - 8]))
        T[n - 3] = (_d2(A, u[n - 2], u[n - 3], u[n - 4]) + _d2(B, u[n - 1], u[n - 3], u[n - 5])
                    + C * (-u[n - 2] - u[n - 3] - u[n - 3] + u[n - 6])
                    + D * (-u[n - 3] - u[n - 3] - u[n - 3] + u[n - 7]))
        T[n - 2] = (_d2(A, u[n - 1], u[n - 2], u[n - 3])
                    + B * (-u[n - 2] - u[n - 2] - u[n - 2] + u[n - 4])
                    + C * (-u[n - 3] - u[n - 2] - u[n - 2] + u[n - 5])
                    + D * (-u[n - 4] - u[n - 2] - u[n - 2] + u[n - 6]))
        T[n - 1] = 0.0
    else:
        T[n - 4] = (_d2(c["astt"], u[n - 3], u[n - 4], u[n - 5])
                    + _d2(c["bstt"], u[n - 2], u[n - 4], u[n - 6])
                    + _d2(c["cstt"], u[n - 1], u[n - 4], u[n - 7]))
        T[n - 3] = (_d2(c["ast"], u[n - 2], u[n - 3], u[n - 4])
                    + _d2(c["bst"], u[n - 1], u[n - 3], u[n - 5]))
        T[n - 2] = _d2(c["asm"], u[n - 1], u[n - 2], u[n - 3])
        T[n - 1] = c["asn"] * u[n - 1] + c["bsn"] * u[n - 2] + c["csn"] * u[n - 3] + c["dsn"] * u[n - 4]
    return T


def _rhs_dervp(U, c, periodic, npaire, nm):
    """derxvp stencil (derive.f90:3796-3907): velocity grid -> pressure grid."""
    a, b = c["aci"], c["bci"]
    n = U.shape[0]
    T = np.empty((nm,) + U.shape[1:])
    if periodic:
        for i in range(n):
            T[i] = a * (U[(i + 1) % n] - U[i]) + b * (U[(i + 2) % n] - U[(i - 1) % n])
        return T
    u = U
    for i in range(2, nm - 2):
        T[i] = a * (u[i + 1] - u[i]) + b * (u[i + 2] - u[i - 1])
    T[1] = a * (u[2] - u[1]) + b * (u[3] - u[0])
    T[nm - 2] = a * (u[nm - 1] - u[nm - 2]) + b * (u[n - 1] - u[nm - 3])
    if npaire == 1:
        T[0] = a * (u[1] - u[0]) + b * (u[2] - u[1])
        T[nm - 1] = a * (u[n - 1] - u[nm - 1]) + b * (u[nm - 1] - u[nm - 2])
    else:
        T[0] = a * (u[1] - u[0]) + b * (u[2] - 2.0 * u[0] + u[1])
        T[nm - 1] = a * (u[n - 1] - u[nm - 1]) + b * (2.0 * u[n - 1] - u[nm - 1] - u[nm - 2])
    return T


def _rhs_intervp(U, c, periodic, npaire, nm):
    """interxvp stencil (derive.f90:3911-4037)."""
    a, b, cc, d = c["aici"], c["bici"], c["cici"], c["dici"]
    n = U.shape[0]
    T = np.empty((nm,) + U.shape[1:])
    if periodic:
        w = lambda j: U[j % n]  # noqa: E731
        for i in range(n):
            T[i] = (a * (w(i + 1) + w(i)) + b * (w(i + 2) + w(i - 1))
                    + cc * (w(i + 3) + w(i - 2)) + d * (w(i + 4) + w(i - 3)))
        return T
    if npaire != 1:
        raise ValueError("interxvp: the reference has no odd non-periodic branch")

    def w(j):  # even reflection about velocity points 0 and n-1 (rows 1-3, nxm-2..nxm)
        if j < 0:
            return U[-j]
        if j > n - 1:
            return U[2 * (n - 1) - j]
        return U[j]
    for i in range(nm):
        T[i] = (a * (w(i + 1) + w(i)) + b * (w(i + 2) + w(i - 1))
                + cc * (w(i + 3) + w(i - 2)) + d * (w(i + 4) + w(i - 3)))
    return T


def _rhs_derpv(U, c, periodic, npaire, n):
    """derxpv stencil (derive.f90:4041-4122): pressure grid -> velocity grid."""
    a, b = c["aci"], c["bci"]
    nm = U.shape[0]
    T = np.empty((n,) + U.shape[1:])
    if periodic:
        for i in range(n):
            T[i] = a * (U[i] - U[(i - 1) % n]) + b * (U[(i + 1) % n] - U[(i - 2) % n])
        return T
    if npaire != 1:
        raise ValueError("derxpv: the reference has no odd non-periodic branch")
    u = U
    T[0] = 0.0
    T[1] = a * (u[1] - u[0]) + b * (u[2] - u[0])
    for i in range(2, n - 2):
        T[i] = a * (u[i] - u[i - 1]) + b * (u[i + 1] - u[i - 2])
    T[n - 2] = a * (u[n - 2] - u[n - 3]) + b * (u[n - 2] - u[n - 4])
    T[n - 1] = 0.0
    return T


def _rhs_interpv(U, c, periodic, npaire, n):
    """interxpv stencil (derive.f90:4126-4261)."""
    a, b, cc, d = c["aici"], c["bici"], c["cici"], c["dici"]
    nm = U.shape[0]
    T = np.empty((n,) + U.shape[1:])
    if periodic:
        w = lambda j: U[j % n]  # noqa: E731
        for i in range(n):
            T[i] = (a * (w(i) + w(i - 1)) + b * (w(i + 1) + w(i - 2))
                    + cc * (w(i + 2) + w(i - 3)) + d * (w(i + 3) + w(i - 4)))
        return T
    if npaire != 1:
        raise ValueError("interxpv: the reference has no odd non-periodic branch")

    def w(j):  # even reflection about pressure half-points -1/2 and nm-1/2
        if j < 0:
            return U[-1 - j]
        if j > nm - 1:
            return U[2 * nm - 1 - j]
        return U[j]
    for i in range(n):
        T[i] = (a * (w(i) + w(i - 1)) + b * (w(i + 1) + w(i - 2))
                + cc * (w(i + 2) + w(i - 3)) + d * (w(i + 3) + w(i - 4)))
    return T


# ---------------------------------------------------------------------------
# public operators mirroring derive.f90 (axis-generic)
# ---------------------------------------------------------------------------
def der1(S: Schemes, u, ax, npaire):
    """der{x,y,z} with the coefficient set the reference callers pass:
    (ff,fs,fw)p for npaire=1, (ff,fs,fw) for npaire=0 (transeq.f90:125-130)."""
    c = S.d1[ax]
    bc1, bcn = S.bc[ax]
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_der1(U, c, bc1, bcn, npaire)
    if npaire == 1:
        ff, fs, fw = c["ffp"], c["fsp"], c["fwp"]
    else:
        ff, fs, fw = c["ff"], c["fs"], c["fw"]
    if bc1 == 0:
        _thomas_cyclic(T, ff, fs, fw, c["alfai"])
    else:
        _thomas(T, ff, fs, fw)
    if ax == "y" and S.stret is not None:
        T = T * S.stret["ppy"][:, None]                 # derive.f90:409-417
    return _unlines(T, shape, AXIS[ax])


def der2(S: Schemes, u, ax, npaire):
    """der{xx,yy,zz}: (sf,ss,sw)p for npaire=1, (sf,ss,sw) for npaire=0
    (transeq.f90:301-303,336-362,442-444)."""
    c = S.d2[ax]
    bc1, bcn = S.bc[ax]
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_der2(U, c, bc1, bcn, npaire)
    if npaire == 1:
        ff, fs, fw = c["sfp"], c["ssp"], c["swp"]
    else:
        ff, fs, fw = c["sf"], c["ss"], c["sw"]
    if bc1 == 0:
        _thomas_cyclic(T, ff, fs, fw, c["alsai"])
    else:
        _thomas(T, ff, fs, fw)
    return _unlines(T, shape, AXIS[ax])


def dervp(S: Schemes, u, ax, npaire=0):
    """der{x,y,z}vp with (cf6,cs6,cw6) as in divergence (navier.f90:297,322,334)."""
    c = S.st[ax]
    periodic = S.bc[ax][0] == 0
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_dervp(U, c, periodic, npaire, S.nm[ax])
    if periodic:
        _thomas_cyclic(T, c["cfx6"], c["csx6"], c["cwx6"], c["alcai"])
    else:
        _thomas(T, c["cfx6"], c["csx6"], c["cwx6"])
    if ax == "y" and S.stret is not None:
        T = T * S.stret["ppyi"][:S.nm[ax], None]        # derive.f90:4572-4578
    return _unlines(T, (S.nm[ax],) + shape[1:], AXIS[ax])


def intervp(S: Schemes, u, ax, npaire=1):
    """inter{x,y,z}vp with (cif6,cis6,ciw6)p as in divergence (navier.f90:313-314,321,333)."""
    c = S.st[ax]
    periodic = S.bc[ax][0] == 0
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_intervp(U, c, periodic, npaire, S.nm[ax])
    if periodic:
        _thomas_cyclic(T, c["cifxp6"], c["cisxp6"], c["ciwxp6"], c["ailcai"])
    else:
        _thomas(T, c["cifxp6"], c["cisxp6"], c["ciwxp6"])
    return _unlines(T, (S.nm[ax],) + shape[1:], AXIS[ax])


def derpv(S: Schemes, u, ax, npaire=1):
    """der{x,y,z}pv as in gradp (navier.f90:405,415,426): periodic lines use
    (cf6,cs6,cw6), non-periodic (cfip6,csip6,cwip6) (derive.f90:4062-4118)."""
    c = S.st[ax]
    periodic = S.bc[ax][0] == 0
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_derpv(U, c, periodic, npaire, S.n[ax])
    if periodic:
        _thomas_cyclic(T, c["cfx6"], c["csx6"], c["cwx6"], c["alcai"])
    else:
        _thomas(T, c["cfip6"], c["csip6"], c["cwip6"])
    if ax == "y" and S.stret is not None:
        T = T * S.stret["ppy"][:, None]                 # derive.f90:4905-4911
    return _unlines(T, (S.n[ax],) + shape[1:], AXIS[ax])


def interpv(S: Schemes, u, ax, npaire=1):
    """inter{x,y,z}pv as in gradp (navier.f90:403,413,417,428,430): periodic lines
    use (cif6,cis6,ciw6), non-periodic (cifip6,cisip6,ciwip6)."""
    c = S.st[ax]
    periodic = S.bc[ax][0] == 0
    U, shape = _lines(u, AXIS[ax])
    T = _rhs_interpv(U, c, periodic, npaire, S.n[ax])
    if periodic:
        _thomas_cyclic(T, c["cifx6"], c["cisx6"], c["ciwx6"], c["ailcai"])
    else:
        _thomas(T, c["cifip6"], c["cisip6"], c["ciwip6"])
    return _unlines(T, (S.n[ax],) + shape[1:], AXIS[ax])


# ---------------------------------------------------------------------------
# poisson.f90: abxyz, waves, poisson_000, poisson_010
# ---------------------------------------------------------------------------
def abxyz(n, bc):
    """Shift factors (poisson.f90:1469-1526)."""
    i = np.arange(n, dtype=np.float64)
    if bc == 0:
        return np.sin(i * PI / n), np.cos(i * PI / n)
    return np.sin(i * PI * 0.5 / n), np.cos(i * PI * 0.5 / n)


def _wave_1d(c, d, n_glob, nm, L, bc):
    """Per-direction modified wavenumber squared and transfer function
    (poisson.f90:1566-1700 for bc=0/1 in x and y).  Returns (k2, T) over the
    n entries of the (x or y) spectral index."""
    if bc == 0:
        n = n_glob
        k2 = np.zeros(n)
        es = np.zeros(n)
        for i in range(n // 2 + 1):
            w = TWOPI * i / n
            wp = c["aci"] * 2.0 * d * math.sin(w * 0.5) + c["bci"] * 2.0 * d * math.sin(1.5 * w)
            wp = wp / (1.0 + 2.0 * c["alcai"] * math.cos(w))
            es[i] = n * w / L
            k2[i] = (n * wp / L) ** 2
        for i in range(n // 2 + 1, n):
            k2[i] = k2[n - i]
            es[i] = es[n - i]
    else:
        n = nm
        k2 = np.zeros(n)
        es = np.zeros(n)
        for i in range(n):
            w = TWOPI * 0.5 * i / nm
            wp = c["aci"] * 2.0 * d * math.sin(w * 0.5) + (c["bci"] * 2.0 * d) * math.sin(1.5 * w)
            wp = wp / (1.0 + 2.0 * c["alcai"] * math.cos(w))
            es[i] = nm * w / L
            k2[i] = (nm * wp / L) ** 2
        k2[0] = 0.0
        es[0] = 0.0
    # transfer function of the mid-point interpolation at w = es*d
    T = np.empty(n)
    for i in range(n):
        r = es[i] * d
        tt = 2.0 * (c["bici"] * math.cos(r * 1.5) + c["cici"] * math.cos(r * 2.5)
                    + c["dici"] * math.cos(r * 3.5))
        tt1 = 2.0 * c["aici"] * math.cos(r * 0.5)
        t1 = 1.0 + 2.0 * c["ailcai"] * math.cos(r)
        T[i] = (tt1 + tt) / t1
    return k2, T


def waves_kxyz(S: Schemes):
    """Real part of kxyz for bcx=bcz=0 (poisson.f90:1706-1783); the imaginary
    part is identical (cx_one_one factors).  Shape (nz, ny, nx) on the
    pressure grid, full spectral index range in every direction."""
    p = S.p
    bcx = 0 if p.nclx else 1
    bcy = 0 if p.ncly else 1
    bcz = 0 if p.nclz else 1
    if bcx != 0 or bcz != 0:
        raise NotImplementedError("only x/z periodic wave numbers are restated")
    kx2, Tx = _wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, bcx)
    ky2, Ty = _wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, bcy)
    kz2, Tz = _wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, bcz)
    X2 = kx2[None, None, :]; Y2 = ky2[None, :, None]; Z2 = kz2[:, None, None]
    TX = Tx[None, None, :]; TY = Ty[None, :, None]; TZ = Tz[:, None, None]
    xt2 = X2 * ((TY * TZ) ** 2)
    yt2 = Y2 * ((TX * TZ) ** 2)
    zt2 = Z2 * ((TX * TY) ** 2)
    return xt2 + yt2 + zt2


EPS_POISSON = 1.0e-16  # poisson.f90:24-28 (DOUBLE_PREC)


def poisson_000(S: Schemes, rhs):
    """Periodic Poisson (poisson.f90:298-410), restated step by step:
    r2c (z real axis, as PHYSICAL_IN_Z) -> 1/N -> per-mode shift
    post-processing -> divide by -K (zero if K<eps) -> inverse shifts -> c2r.
    numpy.fft stands in for 2DECOMP&FFT v2.0.4 (absent from the reference)."""
    p = S.p
    nx, ny, nz = p.nx, p.ny, p.nz
    cw = np.fft.rfft(rhs, axis=0)          # z (real) -> nz/2+1
    cw = np.fft.fft(cw, axis=1)            # y
    cw = np.fft.fft(cw, axis=2)            # x  -> Fortran cw1(nx, ny, nz/2+1)
    cw = cw / float(nx) / float(ny) / float(nz)
    ax_, bx_ = abxyz(nx, 0)
    ay_, by_ = abxyz(ny, 0)
    az_, bz_ = abxyz(nz, 0)
    nzh = nz // 2 + 1
    K = waves_kxyz(S)[:nzh]
    AZ = az_[:nzh, None, None]; BZ = bz_[:nzh, None, None]
    AY = ay_[None, :, None]; BY = by_[None, :, None]
    AX = ax_[None, None, :]; BX = bx_[None, None, :]
    sy = np.where(np.arange(ny) > ny // 2, -1.0, 1.0)[None, :, None]   # j > ny/2+1 (1-based)
    sx = np.where(np.arange(nx) > nx // 2, -1.0, 1.0)[None, None, :]
    t1, t2 = cw.real, cw.imag
    t1, t2 = t1 * BZ + t2 * AZ, t2 * BZ - t1 * AZ
    t1, t2 = t1 * BY + t2 * AY, t2 * BY - t1 * AY
    t1, t2 = t1 * sy, t2 * sy
    t1, t2 = t1 * BX + t2 * AX, t2 * BX - t1 * AX
    t1, t2 = t1 * sx, t2 * sx
    zero = K < EPS_POISSON
    Ks = np.where(zero, 1.0, K)
    t1 = np.where(zero, 0.0, t1 / (-Ks))
    t2 = np.where(zero, 0.0, t2 / (-Ks))
    t1, t2 = t1 * BZ - t2 * AZ, -t2 * BZ - t1 * AZ
    t1, t2 = t1 * BY + t2 * AY, t2 * BY - t1 * AY
    t1, t2 = t1 * sy, t2 * sy
    t1, t2 = t1 * BX + t2 * AX, -t2 * BX + t1 * AX
    t1, t2 = t1 * sx, t2 * sx
    cw = t1 + 1j * t2
    cw = np.fft.ifft(cw, axis=2) * nx
    cw = np.fft.ifft(cw, axis=1) * ny
    return np.fft.irfft(cw, n=nz, axis=0) * nz


def poisson_000_closed_form(S: Schemes, rhs):
    """p_hat = -f_hat/K with K<1e-16 zeroed (the algebra of poisson_000)."""
    K = waves_kxyz(S)
    f = np.fft.fftn(rhs)
    zero = K < EPS_POISSON
    ph = np.where(zero, 0.0, f / (-np.where(zero, 1.0, K)))
    return np.fft.ifftn(ph).real


def _refinement_1d(S: Schemes):
    """The 1-D factors matrice_refinement builds (poisson.f90:1845-1925) for
    bcx = bcz = 0, bcy = 1: the mid-point interpolation transfer functions
    transx_rl, transy_rl, transz_rl (at exs*dx, eys*dy, ezs*dz), xk2, zk2, and
    yky = nym*wp without the 1/yly of the unstretched case (poisson.f90:1614).
    Real and imaginary parts are equal (cx_one_one), so one real array each."""
    p = S.p
    kx2, tx = _wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, 0)
    _, ty = _wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, 1)
    kz2, tz = _wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, 0)
    c = S.st["y"]
    yky = np.empty(p.nym)
    for j in range(p.nym):
        w = TWOPI * 0.5 * j / p.nym
        wp = c["aci"] * 2.0 * p.dy * math.sin(w * 0.5) + (c["bci"] * 2.0 * p.dy) * math.sin(1.5 * w)
        wp = wp / (1.0 + 2.0 * c["alcai"] * math.cos(w))
        yky[j] = p.nym * wp
    yky[0] = 0.0
    nzh = p.nz // 2 + 1
    return tx, ty, tz[:nzh], kx2, kz2[:nzh], yky


def matrice_refinement(S: Schemes):
    """Pentadiagonal systems of the stretched-y Poisson solve
    (poisson.f90:1814-2038, istret = 1 or 2, bcz = 0): a for the even cosine
    modes 0, 2, 4, ... and a2 for the odd modes 1, 3, 5, ..., shape
    (nz/2+1, nym/2, nx, 5) over (kz, row, kx, diagonal -2..+2).  Entries are
    real: the reference's real and imaginary parts are equal here."""
    p = S.p
    if p.istret == 3:
        return (_matrice_refinement_3(S),)
    alpha, beta = S.stret["alpha"], p.beta
    tx, ty, tz, kx2, kz2, yky = _refinement_1d(S)
    nyh = p.ny // 2                      # the velocity grid's ny/2 = nym/2
    xa0 = alpha / PI + 0.5 / beta / PI
    xa1 = (1.0 / 4.0 / beta / PI) if p.istret == 1 else -(1.0 / 4.0 / beta / PI)
    xa0_2, xa1_2, xa01, xa0p1_2 = xa0 * xa0, xa1 * xa1, xa0 * xa1, (xa0 + xa1) * (xa0 + xa1)
    TX = tx[None, :]; TZ = tz[:, None]
    TX2 = TX ** 2; TZ2 = TZ ** 2
    XK2 = kx2[None, :]; ZK2 = kz2[:, None]
    ty2 = ty ** 2
    # cw22(j) = transx * yky(2j-1) * transz, cw2(j) = transx * yky(2j) * transz (1-based j)
    cw22 = [TX * (yky[2 * j] * TZ) for j in range(nyh)]
    cw2 = [TX * (yky[2 * j + 1] * TZ) for j in range(nyh)]
    shape = (tz.size, nyh, tx.size, 5)
    a = np.zeros(shape)
    a2 = np.zeros(shape)

    def kxz(t2):
        return XK2 * t2 * TZ2 + ZK2 * t2 * TX2

    # main diagonal (:1954-2016)
    for j in range(1, nyh - 1):
        a[:, j, :, 2] = -(kxz(ty2[2 * j]) + xa0_2 * cw22[j] ** 2 + xa1_2 * cw22[j] * (cw22[j - 1] + cw22[j + 1]))
        a2[:, j, :, 2] = -(kxz(ty2[2 * j + 1]) + xa0_2 * cw2[j] ** 2 + xa1_2 * cw2[j] * (cw2[j - 1] + cw2[j + 1]))
    n = nyh - 1
    a[:, 0, :, 2] = -(kxz(ty2[0]) + xa0_2 * cw22[0] ** 2 + xa1_2 * cw22[0] * cw22[1])
    a[:, n, :, 2] = -(kxz(ty2[p.ny - 3]) + xa0_2 * cw22[n] ** 2 + xa1_2 * cw22[n] * cw22[n - 1])
    a2[:, 0, :, 2] = -(kxz(ty2[1]) + (xa0_2 - xa1_2) * cw2[0] ** 2 + xa1_2 * cw2[0] * cw2[1])
    a2[:, n, :, 2] = -(kxz(ty2[p.ny - 2]) + xa0p1_2 * cw2[n] ** 2 + xa1_2 * cw2[n] * cw2[n - 1])
    # super-diagonal +1 (:2020-2047)
    for j in range(1, nyh - 1):
        a[:, j, :, 3] = xa01 * (cw22[j + 1] * (cw22[j] + cw22[j + 1]))
        a2[:, j, :, 3] = xa01 * (cw2[j + 1] * (cw2[j] + cw2[j + 1]))
    a[:, 0, :, 3] = 2.0 * xa01 * (cw22[0] * cw22[1] + cw22[1] * cw22[1])
    a2[:, 0, :, 3] = (xa0 - xa1) * xa1 * (cw2[0] * cw2[1]) + xa0 * xa1 * (cw2[1] * cw2[1])
    a2[:, n - 1, :, 3] = xa0 * xa1 * cw2[n - 1] * cw2[n] + (xa0 + xa1) * xa1 * (cw2[n] ** 2)
    a2[:, n, :, 3] = 0.0
    # super-diagonal +2 (:2050-2069)
    for j in range(0, nyh - 2):
        a[:, j, :, 4] = xa1_2 * (-cw22[j + 1] * cw22[j + 2])
        a2[:, j, :, 4] = xa1_2 * (-cw2[j + 1] * cw2[j + 2])
    a[:, 0, :, 4] = 2.0 * a[:, 0, :, 4]
    a[:, n - 1, :, 4] = 0.0
    a[:, n, :, 4] = 0.0
    a2[:, n - 1, :, 4] = 0.0
    a2[:, n, :, 4] = 0.0
    # sub-diagonal -1 (:2072-2099)
    for j in range(1, nyh):
        a[:, j, :, 1] = xa01 * (cw22[j - 1] * (cw22[j] + cw22[j - 1]))
        a2[:, j, :, 1] = xa01 * (cw2[j - 1] * (cw2[j] + cw2[j - 1]))
    a[:, 0, :, 1] = 0.0
    a2[:, 0, :, 1] = 0.0
    a2[:, 1, :, 1] = xa0 * xa1 * (cw2[1] * cw2[0]) + (xa0 + xa1) * xa1 * (cw2[0] * cw2[0])
    a2[:, n, :, 1] = (xa0 + xa1) * xa1 * (cw2[n] * cw2[n - 1]) + xa0 * xa1 * (cw2[n - 1] * cw2[n - 1])
    # sub-diagonal -2 (:2101-2114)
    for j in range(2, nyh):
        a[:, j, :, 0] = xa1_2 * (-cw22[j - 1] * cw22[j - 2])
        a2[:, j, :, 0] = xa1_2 * (-cw2[j - 1] * cw2[j - 2])
    a[:, 0:2, :, 0] = 0.0
    a2[:, 0:2, :, 0] = 0.0
    # the mean mode (kx = kz = 0): unit first row (:2116-2124)
    sing = (XK2 == 0.0) & (ZK2 == 0.0)
    a[:, 0, :, 2] = np.where(sing, 1.0, a[:, 0, :, 2])
    a[:, 0, :, 3] = np.where(sing, 0.0, a[:, 0, :, 3])
    a[:, 0, :, 4] = np.where(sing, 0.0, a[:, 0, :, 4])
    return a, a2


def _matrice_refinement_3(S: Schemes):
    """istret = 3 (poisson.f90:2126-2247): one pentadiagonal system over all
    nym cosine modes, cw22(j) = transx * yky(j) * transz, shape
    (nz/2+1, nym, nx, 5)."""
    p = S.p
    alpha, beta = S.stret["alpha"], p.beta
    tx, ty, tz, kx2, kz2, yky = _refinement_1d(S)
    nym = p.nym
    xa0 = alpha / PI + 0.5 / beta / PI
    xa1 = -(1.0 / 4.0 / beta / PI)
    xa0_2, xa1_2, xa01 = xa0 * xa0, xa1 * xa1, xa0 * xa1
    TX = tx[None, :]; TZ = tz[:, None]
    TX2 = TX ** 2; TZ2 = TZ ** 2
    XK2 = kx2[None, :]; ZK2 = kz2[:, None]
    ty2 = ty ** 2
    cw = [TX * (yky[j] * TZ) for j in range(nym)]
    a = np.zeros((tz.size, nym, tx.size, 5))

    def kxz(t2):
        return XK2 * t2 * TZ2 + ZK2 * t2 * TX2

    n = nym - 1
    for j in range(1, n):
        a[:, j, :, 2] = -(kxz(ty2[j]) + xa0_2 * cw[j] ** 2 + xa1_2 * cw[j] * (cw[j - 1] + cw[j + 1]))
    a[:, 0, :, 2] = -(kxz(ty2[0]) + xa0_2 * cw[0] ** 2 + xa1_2 * cw[0] * cw[1])
    a[:, n, :, 2] = -(kxz(ty2[n]) + xa0_2 * cw[n] ** 2 + xa1_2 * cw[n] * cw[n - 1])
    for j in range(0, n):
        a[:, j, :, 3] = xa01 * (cw[j + 1] * (cw[j] + cw[j + 1]))
    for j in range(0, nym - 2):
        a[:, j, :, 4] = -xa1_2 * (cw[j + 1] * cw[j + 2])
    for j in range(1, nym):
        a[:, j, :, 1] = xa01 * (cw[j - 1] * (cw[j] + cw[j - 1]))
    for j in range(2, nym):
        a[:, j, :, 0] = -xa1_2 * (cw[j - 1] * cw[j - 2])
    # the mean mode (kx = kz = 0): unit first row (:2240-2245)
    sing = (XK2 == 0.0) & (ZK2 == 0.0)
    a[:, 0, :, 2] = np.where(sing, 1.0, a[:, 0, :, 2])
    a[:, 0, :, 3] = np.where(sing, 0.0, a[:, 0, :, 3])
    a[:, 0, :, 4] = np.where(sing, 0.0, a[:, 0, :, 4])
    return a


EPS_INV5 = 1.0e-16   # tools.f90:1241-1245 (DOUBLE_PREC)


def inversion5_v1(aaa_in, eee):
    """Pentadiagonal solve of tools.f90:1225-1362 on real systems: aaa
    (nz', m, nx', 5) over (kz, row, kx, diagonal -2..+2), eee (nz', m, nx').
    Elimination without pivoting, then the last 2x2 block and the back
    substitution with every division by a pivot of magnitude <= 1e-16 taken
    as zero.  Where the reference finds an exactly zero pivot during the
    elimination it keeps the previous line's multiplier; such a line's
    matrix is zero, its back substitution returns zero either way, and here
    the multiplier is taken as zero."""
    aaa = aaa_in.copy()
    e = eee.copy()
    m_ = aaa.shape[1]
    nz = m_ - 1
    for m in range(m_ - 2):
        for i in (1, 2):
            mi = m + i
            piv = aaa[:, m, :, 2]
            sr = np.where(piv != 0.0, aaa[:, mi, :, 2 - i] / np.where(piv != 0.0, piv, 1.0), 0.0)
            e[:, mi, :] = e[:, mi, :] - sr * e[:, m, :]
            for jc in ((3, 4) if i == 1 else (2, 3)):
                aaa[:, mi, :, jc - 1] = aaa[:, mi, :, jc - 1] - sr * aaa[:, m, :, jc - 1 + i]

    def guarded(num, den):
        ok = np.abs(den) > EPS_INV5
        return np.where(ok, num / np.where(ok, den, 1.0), 0.0), ok

    sr, _ = guarded(aaa[:, nz, :, 1], aaa[:, nz - 1, :, 2])
    b1 = aaa[:, nz, :, 2] - sr * aaa[:, nz - 1, :, 3]
    ok = np.abs(b1) > EPS_INV5
    b1s = np.where(ok, b1, 1.0)
    t1 = np.where(ok, sr / b1s, 0.0)
    t3 = np.where(ok, e[:, nz, :] / b1s - t1 * e[:, nz - 1, :], 0.0)
    e[:, nz, :] = t3
    rp, _ = guarded(1.0, aaa[:, nz - 1, :, 2])
    a1 = aaa[:, nz - 1, :, 3] * rp
    e[:, nz - 1, :] = e[:, nz - 1, :] * rp - a1 * e[:, nz, :]
    for i in range(m_ - 3, -1, -1):
        rp, _ = guarded(1.0, aaa[:, i, :, 2])
        a1 = aaa[:, i, :, 3] * rp
        b1 = aaa[:, i, :, 4] * rp
        e[:, i, :] = e[:, i, :] * rp - a1 * e[:, i + 1, :] - b1 * e[:, i + 2, :]
    return e


def poisson_010(S: Schemes, rhs):
    """Poisson with non-periodic y (poisson.f90:665-1013, istret=0 branch),
    restated step by step on the pressure grid (nz, nym, nx)."""
    p = S.p
    nx, ny, nz = p.nx, p.nym, p.nz
    assert rhs.shape == (nz, ny, nx)
    # even/odd reorder in y (poisson.f90:688-697)
    rb = np.empty_like(rhs)
    for j in range(ny // 2):
        rb[:, j, :] = rhs[:, 2 * j, :]
    for j in range(ny // 2, ny):
        rb[:, j, :] = rhs[:, 2 * ny - 2 * (j + 1) + 1, :]
    cw = np.fft.rfft(rb, axis=0)
    cw = np.fft.fft(cw, axis=1)
    cw = np.fft.fft(cw, axis=2)
    cw = cw / float(nx) / float(ny) / float(nz)
    nzh = nz // 2 + 1
    ax_, bx_ = abxyz(nx, 0)
    ay_, by_ = abxyz(ny, 1)
    az_, bz_ = abxyz(nz, 0)
    AZ = az_[:nzh, None, None]; BZ = bz_[:nzh, None, None]
    AX = ax_[None, None, :]; BX = bx_[None, None, :]
    sx = np.where(np.arange(nx) > nx // 2, -1.0, 1.0)[None, None, :]
    t1, t2 = cw.real, cw.imag
    t1, t2 = t1 * BZ + t2 * AZ, t2 * BZ - t1 * AZ
    t1, t2 = t1 * BX + t2 * AX, t2 * BX - t1 * AX
    t1, t2 = t1 * sx, t2 * sx
    # y post-processing (poisson.f90:761-781)
    c1, c2 = t1.copy(), t2.copy()
    jj = np.arange(1, ny)
    mir = ny - jj          # 0-based index of ny-j+2 (1-based) for j=jj+1
    BYv = by_[jj][None, :, None]; AYv = ay_[jj][None, :, None]
    tmp1, tmp2 = t1[:, jj, :], t2[:, jj, :]
    tmp3, tmp4 = t1[:, mir, :], t2[:, mir, :]
    xx1 = tmp1 * BYv; xx2 = tmp1 * AYv; xx3 = tmp2 * BYv; xx4 = tmp2 * AYv
    xx5 = tmp3 * BYv; xx6 = tmp3 * AYv; xx7 = tmp4 * BYv; xx8 = tmp4 * AYv
    c1[:, jj, :] = 0.5 * (xx1 + xx4 + xx5 - xx8)
    c2[:, jj, :] = 0.5 * (-xx2 + xx3 + xx6 + xx7)
    if p.istret == 0:
        # divide (poisson.f90:795-819); K real and imag parts are equal
        K = waves_kxyz(S)[:nzh]
        small = np.abs(K) < EPS_POISSON
        Ks = np.where(small, 1.0, K)
        c1 = np.where(small, 0.0, c1 / (-Ks))
        c2 = np.where(small, 0.0, c2 / (-Ks))
    else:
        # stretched y: even and odd cosine modes each solve a pentadiagonal
        # system, real and imaginary parts separately (poisson.f90:846-879)
        mats = matrice_refinement(S)
        if len(mats) == 1:
            # istret = 3: one system over all modes, inversion5_v2
            # (tools.f90:1365-1497, the same elimination on nym rows)
            c1 = inversion5_v1(mats[0], c1)
            c2 = inversion5_v1(mats[0], c2)
        else:
            for par, A in ((0, mats[0]), (1, mats[1])):
                c1[:, par::2, :] = inversion5_v1(A, c1[:, par::2, :])
                c2[:, par::2, :] = inversion5_v1(A, c2[:, par::2, :])
    # zero the (nx/2+1, :, nz/2+1) line (poisson.f90:902-908)
    if nz // 2 < nzh:
        c1[nz // 2, :, nx // 2] = 0.0
        c2[nz // 2, :, nx // 2] = 0.0
    # backward y post-processing (poisson.f90:924-943)
    d1, d2 = c1.copy(), c2.copy()
    tmp1, tmp2 = c1[:, jj, :], c2[:, jj, :]
    tmp3, tmp4 = c1[:, mir, :], c2[:, mir, :]
    xx1 = tmp1 * BYv; xx2 = tmp1 * AYv; xx3 = tmp2 * BYv; xx4 = tmp2 * AYv
    xx5 = tmp3 * BYv; xx6 = tmp3 * AYv; xx7 = tmp4 * BYv; xx8 = tmp4 * AYv
    d1[:, jj, :] = xx1 - xx4 + xx6 + xx7
    d2[:, jj, :] = -(-xx2 - xx3 + xx5 - xx8)
    t1, t2 = d1, d2
    t1, t2 = t1 * BX - t2 * AX, t2 * BX + t1 * AX
    t1, t2 = t1 * sx, t2 * sx
    t1, t2 = t1 * BZ - t2 * AZ, t2 * BZ + t1 * AZ
    cw = t1 + 1j * t2
    cw = np.fft.ifft(cw, axis=2) * nx
    cw = np.fft.ifft(cw, axis=1) * ny
    r = np.fft.irfft(cw, n=nz, axis=0) * nz
    out = np.empty_like(r)
    for j in range(ny // 2):
        out[:, 2 * j, :] = r[:, j, :]
        out[:, 2 * j + 1, :] = r[:, ny - 1 - j, :]
    return out


def _mode_k2_t(c, d, nm, L, w):
    """k'^2 and interpolation transfer at angle w of a non-periodic direction
    (the bc=1 formulas of waves(), poisson.f90:1583-1596 / 1646-1655)."""
    wp = c["aci"] * 2.0 * d * math.sin(w * 0.5) + (c["bci"] * 2.0 * d) * math.sin(1.5 * w)
    wp = wp / (1.0 + 2.0 * c["alcai"] * math.cos(w))
    r = (nm * w / L) * d
    tt = 2.0 * (c["bici"] * math.cos(r * 1.5) + c["cici"] * math.cos(r * 2.5) + c["dici"] * math.cos(r * 3.5))
    tt1 = 2.0 * c["aici"] * math.cos(r * 0.5)
    t1 = 1.0 + 2.0 * c["ailcai"] * math.cos(r)
    return (nm * wp / L) ** 2, (tt1 + tt) / t1


def waves_kxyz_11x(S: Schemes):
    """kxyz for bcx = bcy = 1 (poisson.f90:1740-1800) on the half spectrum in
    z: (K_re, K_im), shape (nz/2+1, ny, nx) of the pressure grid.  With bcz=1
    the real part of z index k is cosine mode k and the imaginary part mode
    nzm-k (ezs, zk2 of poisson.f90:1646-1655); with bcz=0 both are mode k."""
    p = S.p
    bcz = 0 if p.nclz else 1
    kx2, Tx = _wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, 1)
    ky2, Ty = _wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, 1)
    nz = p.nzm if bcz else p.nz
    nzh = nz // 2 + 1
    if bcz:
        kre = np.zeros(nzh); tre = np.zeros(nzh); kim = np.zeros(nzh); tim = np.zeros(nzh)
        for k in range(nzh):
            kre[k], tre[k] = _mode_k2_t(S.st["z"], p.dz, p.nzm, p.zlz, PI * k / p.nzm)
            kim[k], tim[k] = _mode_k2_t(S.st["z"], p.dz, p.nzm, p.zlz, PI * (p.nzm - k) / p.nzm)
    else:
        kz2, Tz = _wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, 0)
        kre, tre = kz2[:nzh], Tz[:nzh]
        kim, tim = kre, tre
    X2 = kx2[None, None, :]; Y2 = ky2[None, :, None]
    TX = Tx[None, None, :]; TY = Ty[None, :, None]
    out = []
    for kz2, Tz in ((kre, tre), (kim, tim)):
        Z2 = kz2[:, None, None]; TZ = Tz[:, None, None]
        out.append(X2 * (TY * TZ) ** 2 + Y2 * (TX * TZ) ** 2 + Z2 * (TX * TY) ** 2)
    return out[0], out[1]


def _reorder(a, axis):
    """Even/odd reorder of poisson.f90:1037-1095: entry j < n/2 takes point
    2j, entry j >= n/2 takes point 2n-2j-1 (0-based)."""
    n = a.shape[axis]
    idx = np.array([2 * j if j < n // 2 else 2 * n - 2 * j - 1 for j in range(n)])
    return np.take(a, idx, axis=axis)


def _unreorder(a, axis):
    """Inverse of _reorder (poisson.f90:1414-1460)."""
    n = a.shape[axis]
    idx = np.array([2 * j if j < n // 2 else 2 * n - 2 * j - 1 for j in range(n)])
    out = np.empty_like(a)
    sl = [slice(None)] * a.ndim
    for j in range(n):
        src = [slice(None)] * a.ndim
        src[axis] = j
        sl[axis] = idx[j]
        out[tuple(sl)] = a[tuple(src)]
    return out


def _pair_fwd(t1, t2, A, B, axis):
    """Forward post-processing of a cosine direction (poisson.f90:1137-1160):
    entry j >= 1 combined with entry n-j, times 1/2; entry 0 unchanged."""
    n = t1.shape[axis]
    j = np.arange(1, n)
    m = n - j
    sh = [1] * t1.ndim
    sh[axis] = n - 1
    Bv = B[j].reshape(sh); Av = A[j].reshape(sh)
    tmp1, tmp2 = np.take(t1, j, axis=axis), np.take(t2, j, axis=axis)
    tmp3, tmp4 = np.take(t1, m, axis=axis), np.take(t2, m, axis=axis)
    xx1 = tmp1 * Bv * 0.5; xx2 = tmp1 * Av * 0.5; xx3 = tmp2 * Bv * 0.5; xx4 = tmp2 * Av * 0.5
    xx5 = tmp3 * Bv * 0.5; xx6 = tmp3 * Av * 0.5; xx7 = tmp4 * Bv * 0.5; xx8 = tmp4 * Av * 0.5
    o1, o2 = t1.copy(), t2.copy()
    idx = [slice(None)] * t1.ndim
    idx[axis] = slice(1, n)
    o1[tuple(idx)] = xx1 + xx4 + xx5 - xx8
    o2[tuple(idx)] = -xx2 + xx3 + xx6 + xx7
    return o1, o2


def _pair_bwd(t1, t2, A, B, axis):
    """Backward post-processing of a cosine direction (poisson.f90:1327-1351)."""
    n = t1.shape[axis]
    j = np.arange(1, n)
    m = n - j
    sh = [1] * t1.ndim
    sh[axis] = n - 1
    Bv = B[j].reshape(sh); Av = A[j].reshape(sh)
    tmp1, tmp2 = np.take(t1, j, axis=axis), np.take(t2, j, axis=axis)
    tmp3, tmp4 = np.take(t1, m, axis=axis), np.take(t2, m, axis=axis)
    xx1 = tmp1 * Bv; xx2 = tmp1 * Av; xx3 = tmp2 * Bv; xx4 = tmp2 * Av
    xx5 = tmp3 * Bv; xx6 = tmp3 * Av; xx7 = tmp4 * Bv; xx8 = tmp4 * Av
    o1, o2 = t1.copy(), t2.copy()
    idx = [slice(None)] * t1.ndim
    idx[axis] = slice(1, n)
    o1[tuple(idx)] = xx1 - xx4 + xx6 + xx7
    o2[tuple(idx)] = -(-xx2 - xx3 + xx5 - xx8)
    return o1, o2


def poisson_11x(S: Schemes, rhs):
    """Poisson with non-periodic x and y (poisson.f90:1019-1465, istret=0),
    z non-periodic (bcz=1) or periodic, restated step by step on the pressure
    grid (nz', nym, nxm): even/odd reorders, r2c along z, post-processing in
    z, y, x, divide real and imaginary parts by their own K, inverse."""
    p = S.p
    bcz = 0 if p.nclz else 1
    nx, ny = p.nxm, p.nym
    nz = p.nzm if bcz else p.nz
    assert rhs.shape == (nz, ny, nx)
    r = _reorder(rhs, 0) if bcz else rhs
    r = _reorder(r, 1)
    r = _reorder(r, 2)
    cw = np.fft.rfft(r, axis=0)
    cw = np.fft.fft(cw, axis=1)
    cw = np.fft.fft(cw, axis=2)
    cw = cw / float(nx) / float(ny) / float(nz)
    nzh = nz // 2 + 1
    ax_, bx_ = abxyz(nx, 1)
    ay_, by_ = abxyz(ny, 1)
    az_, bz_ = abxyz(nz, bcz)
    AZ = az_[:nzh, None, None]; BZ = bz_[:nzh, None, None]
    t1, t2 = cw.real, cw.imag
    t1, t2 = t1 * BZ + t2 * AZ, t2 * BZ - t1 * AZ          # poisson.f90:1112-1121
    t1, t2 = _pair_fwd(t1, t2, ay_, by_, 1)               # :1129-1152
    t1, t2 = _pair_fwd(t1, t2, ax_, bx_, 2)               # :1165-1186
    Kr, Ki = waves_kxyz_11x(S)
    c1 = np.where(np.abs(Kr) < EPS_POISSON, 0.0, t1 / (-np.where(np.abs(Kr) < EPS_POISSON, 1.0, Kr)))
    c2 = np.where(np.abs(Ki) < EPS_POISSON, 0.0, t2 / (-np.where(np.abs(Ki) < EPS_POISSON, 1.0, Ki)))
    t1, t2 = _pair_bwd(c1, c2, ax_, bx_, 2)               # :1306-1325
    t1, t2 = _pair_bwd(t1, t2, ay_, by_, 1)               # :1333-1351
    t1, t2 = t1 * BZ - t2 * AZ, t2 * BZ + t1 * AZ          # :1364-1373
    cw = t1 + 1j * t2
    cw = np.fft.ifft(cw, axis=2) * nx
    cw = np.fft.ifft(cw, axis=1) * ny
    out = np.fft.irfft(cw, n=nz, axis=0) * nz
    if bcz:
        out = _unreorder(out, 0)
    out = _unreorder(out, 1)
    return _unreorder(out, 2)


def waves_kxyz_100(S: Schemes):
    """kxyz for bcx = 1, bcy = bcz = 0 (the bcz == 0 branch of
    poisson.f90:1702-1740): real, on the half spectrum in z, shape
    (nz/2+1, ny, nxm); x index i is cosine mode i (w = pi i / nxm)."""
    p = S.p
    kx2, Tx = _wave_1d(S.st["x"], p.dx, p.nx, p.nxm, p.xlx, 1)
    ky2, Ty = _wave_1d(S.st["y"], p.dy, p.ny, p.nym, p.yly, 0)
    kz2, Tz = _wave_1d(S.st["z"], p.dz, p.nz, p.nzm, p.zlz, 0)
    nzh = p.nz // 2 + 1
    X2 = kx2[None, None, :]; Y2 = ky2[None, :, None]; Z2 = kz2[:nzh, None, None]
    TX = Tx[None, None, :]; TY = Ty[None, :, None]; TZ = Tz[:nzh, None, None]
    return X2 * (TY * TZ) ** 2 + Y2 * (TX * TZ) ** 2 + Z2 * (TX * TY) ** 2


def poisson_100(S: Schemes, rhs):
    """Poisson with non-periodic x, periodic y and z (poisson.f90:413-661),
    restated step by step on the pressure grid (nz, ny, nxm): even/odd
    reorder along x, r2c along z and complex FFTs along y and x, the shift
    factors in z and y (with the sign of the upper half in y), the x pairing
    of entries i and nx-i, the divide by kxyz (real and imaginary parts,
    zero where |K| < epsilon), the inverse post-processing, c2r and the
    inverse reorder."""
    p = S.p
    nx, ny, nz = p.nxm, p.ny, p.nz
    assert rhs.shape == (nz, ny, nx)
    r = _reorder(rhs, 2)                                   # :431-443
    cw = np.fft.rfft(r, axis=0)
    cw = np.fft.fft(cw, axis=1)
    cw = np.fft.fft(cw, axis=2)
    cw = cw / float(nx) / float(ny) / float(nz)            # :455-456
    nzh = nz // 2 + 1
    ax_, bx_ = abxyz(nx, 1)
    ay_, by_ = abxyz(ny, 0)
    az_, bz_ = abxyz(nz, 0)
    AZ = az_[:nzh, None, None]; BZ = bz_[:nzh, None, None]
    AY = ay_[None, :, None]; BY = by_[None, :, None]
    upper = (np.arange(ny) > ny // 2)[None, :, None]       # j > ny/2+1, 1-based
    t1, t2 = cw.real, cw.imag
    t1, t2 = t1 * BZ + t2 * AZ, t2 * BZ - t1 * AZ          # :473-486
    t1, t2 = t1 * BY + t2 * AY, t2 * BY - t1 * AY          # :488-503
    t1, t2 = np.where(upper, -t1, t1), np.where(upper, -t2, t2)
    t1, t2 = _pair_fwd(t1, t2, ax_, bx_, 2)               # :505-527
    K = waves_kxyz_100(S)
    small = np.abs(K) < EPS_POISSON
    c1 = np.where(small, 0.0, t1 / (-np.where(small, 1.0, K)))   # :541-563
    c2 = np.where(small, 0.0, t2 / (-np.where(small, 1.0, K)))
    t1, t2 = _pair_bwd(c1, c2, ax_, bx_, 2)               # :567-587
    t1, t2 = t1 * BY - t2 * AY, t2 * BY + t1 * AY          # :601-614
    t1, t2 = np.where(upper, -t1, t1), np.where(upper, -t2, t2)
    t1, t2 = t1 * BZ - t2 * AZ, t2 * BZ + t1 * AZ          # :617-628
    cw = t1 + 1j * t2
    cw = np.fft.ifft(cw, axis=2) * nx
    cw = np.fft.ifft(cw, axis=1) * ny
    out = np.fft.irfft(cw, n=nz, axis=0) * nz
    return _unreorder(out, 2)                              # :640-650


def poisson(S: Schemes, rhs):
    """Dispatch as decomp_2d_poisson_init (poisson.f90:98-108)."""
    p = S.p
    if p.nclx and p.ncly and p.nclz:
        return poisson_000(S, rhs)
    if p.nclx and (not p.ncly) and p.nclz:
        return poisson_010(S, rhs)
    if (not p.nclx) and p.ncly and p.nclz:
        return poisson_100(S, rhs)
    if (not p.nclx) and (not p.ncly):
        return poisson_11x(S, rhs)
    raise NotImplementedError("boundary combination not supported (poisson.f90:98-108)")


# ---------------------------------------------------------------------------
# navier.f90: divergence, gradp, cor_vel
# ---------------------------------------------------------------------------
def divergence(S: Schemes, ux, uy, uz):
    """div u on the pressure grid (navier.f90:257-372, nlock=1, no IBM/LMN)."""
    pp1 = dervp(S, ux, "x", 0)
    pgy1 = intervp(S, uy, "x", 1)
    pgz1 = intervp(S, uz, "x", 1)
    upi2 = intervp(S, pp1, "y", 1)
    duydypi2 = dervp(S, pgy1, "y", 0)
    duydypi2 = duydypi2 + upi2
    upi2 = intervp(S, pgz1, "y", 1)
    pp3 = intervp(S, duydypi2, "z", 1)
    po3 = dervp(S, upi2, "z", 0)
    return pp3 + po3


def gradp(S: Schemes, pp3):
    """Pressure gradient on the velocity grid (navier.f90:386-499)."""
    ppi3 = interpv(S, pp3, "z", 1)
    pgz3 = derpv(S, pp3, "z", 1)
    ppi2 = interpv(S, ppi3, "y", 1)
    pgy2 = derpv(S, ppi3, "y", 1)
    pgzi2 = interpv(S, pgz3, "y", 1)
    px1 = derpv(S, ppi2, "x", 1)
    py1 = interpv(S, pgy2, "x", 1)
    pz1 = interpv(S, pgzi2, "x", 1)
    return px1, py1, pz1


def div_stats(pp3):
    """tmax and tmoy of navier.f90:346-361."""
    return float(pp3.max()), float(np.abs(pp3).sum() / pp3.size)


# ---------------------------------------------------------------------------
# transeq.f90: momentum_rhs_eq (no LMN, no IBM, iimplicit=0)
# ---------------------------------------------------------------------------
def momentum_rhs(S: Schemes, ux, uy, uz):
    """Skew-symmetric convection + nu*Laplacian (transeq.f90:73-470); on a
    stretched y mesh the y viscous term carries the metrics pp2y, pp4y."""
    xnu = S.p.xnu
    d1, d2 = der1, der2
    # X pencil (transeq.f90:114-146)
    td1 = d1(S, ux * ux, "x", 1)
    te1 = d1(S, ux * uy, "x", 0)
    tf1 = d1(S, ux * uz, "x", 0)
    ta1 = d1(S, ux, "x", 0)
    tb1 = d1(S, uy, "x", 1)
    tc1 = d1(S, uz, "x", 1)
    tg1 = td1 + ux * ta1
    th1 = te1 + ux * tb1
    ti1 = tf1 + ux * tc1
    # Y pencil (transeq.f90:188-214)
    tg2 = d1(S, ux * uy, "y", 0)
    th2 = d1(S, uy * uy, "y", 1)
    ti2 = d1(S, uz * uy, "y", 0)
    td2 = d1(S, ux, "y", 1)
    te2 = d1(S, uy, "y", 0)
    tf2 = d1(S, uz, "y", 1)
    tg2 = tg2 + uy * td2
    th2 = th2 + uy * te2
    ti2 = ti2 + uy * tf2
    # Z pencil (transeq.f90:249-314)
    tg3 = d1(S, ux * uz, "z", 0)
    th3 = d1(S, uy * uz, "z", 0)
    ti3 = d1(S, uz * uz, "z", 1)
    td3 = d1(S, ux, "z", 1)
    te3 = d1(S, uy, "z", 1)
    tf3 = d1(S, uz, "z", 0)
    ta3 = tg3 + uz * td3
    tb3 = th3 + uz * te3
    tc3 = ti3 + uz * tf3
    td3, te3, tf3 = ta3, tb3, tc3
    ta3 = d2(S, ux, "z", 1)
    tb3 = d2(S, uy, "z", 1)
    tc3 = d2(S, uz, "z", 0)
    td3 = xnu * ta3 - 0.5 * td3
    te3 = xnu * tb3 - 0.5 * te3
    tf3 = xnu * tc3 - 0.5 * tf3
    # back in Y (transeq.f90:323-433)
    tg2 = td3 - 0.5 * tg2
    th2 = te3 - 0.5 * th2
    ti2 = tf3 - 0.5 * ti2
    td2 = d2(S, ux, "y", 1)
    te2 = d2(S, uy, "y", 0)
    tf2 = d2(S, uz, "y", 1)
    if S.stret is not None:
        # stretched y (transeq.f90:337-370): d2/dy2 = pp2y D2 - pp4y (ppy D1),
        # the first derivative taken with the metric as dery applies it
        pp2y = S.stret["pp2y"][:, None]
        pp4y = S.stret["pp4y"][:, None]
        td2 = td2 * pp2y - pp4y * d1(S, ux, "y", 1)
        te2 = te2 * pp2y - pp4y * d1(S, uy, "y", 0)
        tf2 = tf2 * pp2y - pp4y * d1(S, uz, "y", 1)
    ta2 = xnu * td2 + tg2
    tb2 = xnu * te2 + th2
    tc2 = xnu * tf2 + ti2
    # back in X (transeq.f90:442-470)
    td1 = xnu * d2(S, ux, "x", 0)
    te1 = xnu * d2(S, uy, "x", 1)
    tf1 = xnu * d2(S, uz, "x", 1)
    dux = ta2 - 0.5 * tg1 + td1
    duy = tb2 - 0.5 * th1 + te1
    duz = tc2 - 0.5 * ti1 + tf1
    return dux, duy, duz


# ---------------------------------------------------------------------------
# time_integrators.f90 (AB3 branch of intt) and the coefficients of
# variables.f90:1365-1375
# ---------------------------------------------------------------------------
def ab3_coefficients(dt):
    adt = (23.0 / 12.0) * dt
    bdt = -(16.0 / 12.0) * dt
    cdt = (5.0 / 12.0) * dt
    return adt, bdt, cdt


def intt_ab3(var, hist, itime, dt):
    """intt for itimescheme=3 (time_integrators.f90:85-100).  hist is a list
    [f1, f2, f3] (f1 newest) and is updated in place; returns new var."""
    adt, bdt, cdt = ab3_coefficients(dt)
    f1, f2, f3 = hist
    if itime == 1:
        var = dt * f1 + var
    elif itime == 2:
        var = 1.5 * dt * f1 - 0.5 * dt * f2 + var
        hist[2] = f2
    else:
        var = adt * f1 + bdt * f2 + cdt * f3 + var
        hist[2] = f2
    hist[1] = f1
    return var


def rk3_coefficients(dt):
    """adt, bdt, gdt of the low-storage RK3 (variables.f90:1388-1399), one
    entry per sub-step."""
    adt = [(8.0 / 15.0) * dt, (5.0 / 12.0) * dt, (3.0 / 4.0) * dt]
    bdt = [0.0, (-17.0 / 60.0) * dt, (-5.0 / 12.0) * dt]
    gdt = [adt[0], adt[1] + bdt[1], adt[2] + bdt[2]]
    return adt, bdt, gdt


def intt_rk3(var, hist, itr, dt):
    """intt for itimescheme=5 (time_integrators.f90:151-157): sub-step itr
    (1-based); hist = [f1, f2] with f1 the current right-hand side, updated
    in place (f2 <- f1); returns the new var."""
    adt, bdt, gdt = rk3_coefficients(dt)
    f1, f2 = hist[0], hist[1]
    if itr == 1:
        var = gdt[0] * f1 + var
    else:
        var = adt[itr - 1] * f1 + bdt[itr - 1] * f2 + var
    hist[1] = f1
    return var


def pre_correc(p: "Params", ux, uy, uz):
    """Free-slip faces of pre_correc (navier.f90:600-614, 686-704, 752-770):
    the normal velocity is set to zero on ncl=1 boundaries."""
    ux, uy, uz = ux.copy(), uy.copy(), uz.copy()
    if p.nclx1 == 1:
        ux[:, :, 0] = 0.0
    if p.nclxn == 1:
        ux[:, :, -1] = 0.0
    if p.ncly1 == 1:
        uy[:, 0, :] = 0.0
    if p.nclyn == 1:
        uy[:, -1, :] = 0.0
    if p.nclz1 == 1:
        uz[0, :, :] = 0.0
    if p.nclzn == 1:
        uz[-1, :, :] = 0.0
    return ux, uy, uz


# ---------------------------------------------------------------------------
# Case-TGV.f90
# ---------------------------------------------------------------------------
def init_tgv(p: Params):
    """TGV initial condition (Case-TGV.f90:64-96, iin=1, no MHD/scalar)."""
    x = np.arange(p.nx) * p.dx
    y = np.arange(p.ny) * p.dy
    z = np.arange(p.nz) * p.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ux = np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -np.cos(X) * np.sin(Y) * np.cos(Z)
    uz = np.zeros_like(ux)
    return ux, uy, uz


def tgv_diagnostics(S: Schemes, ux, uy, uz):
    """(eek, eps, eps2, enst) of postprocess_tgv (Case-TGV.f90:246-377),
    periodic box (no trimmed planes)."""
    p = S.p
    xnu = p.xnu
    ta1 = der1(S, ux, "x", 0); tb1 = der1(S, uy, "x", 1); tc1 = der1(S, uz, "x", 1)
    td1 = der1(S, ux, "y", 1); te1 = der1(S, uy, "y", 0); tf1 = der1(S, uz, "y", 1)
    tg1 = der1(S, ux, "z", 1); th1 = der1(S, uy, "z", 1); ti1 = der1(S, uz, "z", 0)
    sl = (slice(0, p.nz - 1 if p.nclz1 == 1 else p.nz),
          slice(0, p.ny - 1 if p.ncly1 == 1 else p.ny),
          slice(0, p.nx - 1 if p.nclx1 == 1 else p.nx))
    nxc = p.nxm if p.nclx1 == 1 else p.nx
    nyc = p.nym if p.ncly1 == 1 else p.ny
    nzc = p.nzm if p.nclz1 == 1 else p.nz
    ncell = float(nxc * nyc * nzc)
    enst = np.sum(0.5 * ((tf1 - th1) ** 2 + (tg1 - tc1) ** 2 + (tb1 - td1) ** 2)[sl]) / ncell
    eps = np.sum((0.5 * xnu * ((2.0 * ta1) ** 2 + (2.0 * te1) ** 2 + (2.0 * ti1) ** 2
                               + 2.0 * (td1 + tb1) ** 2 + 2.0 * (tg1 + tc1) ** 2
                               + 2.0 * (th1 + tf1) ** 2))[sl]) / ncell
    eek = np.sum(0.5 * (ux ** 2 + uy ** 2 + uz ** 2)[sl]) / ncell
    ta1 = der2(S, ux, "x", 0); tb1 = der2(S, uy, "x", 1); tc1 = der2(S, uz, "x", 1)
    td1 = der2(S, ux, "y", 1); te1 = der2(S, uy, "y", 0); tf1 = der2(S, uz, "y", 1)
    tg1 = der2(S, ux, "z", 1); th1 = der2(S, uy, "z", 1); ti1 = der2(S, uz, "z", 0)
    di = (-xnu) * (ux * (ta1 + td1 + tg1) + uy * (tb1 + te1 + th1) + uz * (tc1 + tf1 + ti1))
    eps2 = np.sum(di[sl]) / ncell
    return float(eek), float(eps), float(eps2), float(enst)


# ---------------------------------------------------------------------------
# one AB3 time step of the periodic incompressible solver (xcompact3d.f90:29-102)
# ---------------------------------------------------------------------------
@dataclass
class State:
    ux: np.ndarray
    uy: np.ndarray
    uz: np.ndarray
    hx: list = field(default_factory=list)
    hy: list = field(default_factory=list)
    hz: list = field(default_factory=list)
    itime: int = 0
    divu: tuple = None    # (max, mean |.|) of div u* of the last (sub-)step


def tgv_state(p: Params):
    ux, uy, uz = init_tgv(p)
    # case.f90:131-136: RHS history seeded with u
    return State(ux, uy, uz, [ux.copy()] * 3, [uy.copy()] * 3, [uz.copy()] * 3, 0)


def divu_stats(pp3):
    """The statistics divergence computes on every call (navier.f90:348-361):
    the maximum (starting from -1609) and the mean |.| of div u* over the
    pressure grid."""
    return max(-1609.0, float(pp3.max())), float(np.abs(pp3).mean())


def step_rk3(S: Schemes, st: State):
    """One RK3 time step (itimescheme=5, xcompact3d.f90:46-87): three
    sub-steps of calculate_transeq_rhs -> int_time -> pre_correc ->
    solve_poisson -> cor_vel.  st.hx/hy/hz hold [f1, f2]."""
    st.itime += 1
    dt = S.p.dt
    for itr in (1, 2, 3):
        dux, duy, duz = momentum_rhs(S, st.ux, st.uy, st.uz)
        st.hx[0], st.hy[0], st.hz[0] = dux, duy, duz
        st.ux = intt_rk3(st.ux, st.hx, itr, dt)
        st.uy = intt_rk3(st.uy, st.hy, itr, dt)
        st.uz = intt_rk3(st.uz, st.hz, itr, dt)
        st.ux, st.uy, st.uz = pre_correc(S.p, st.ux, st.uy, st.uz)
        pp3 = divergence(S, st.ux, st.uy, st.uz)
        st.divu = divu_stats(pp3)
        pp3 = poisson(S, pp3)
        px, py, pz = gradp(S, pp3)
        st.ux = st.ux - px
        st.uy = st.uy - py
        st.uz = st.uz - pz
    return st


def step(S: Schemes, st: State):
    """calculate_transeq_rhs -> int_time (AB3) -> solve_poisson -> cor_vel."""
    st.itime += 1
    dux, duy, duz = momentum_rhs(S, st.ux, st.uy, st.uz)
    st.hx[0], st.hy[0], st.hz[0] = dux, duy, duz
    dt = S.p.dt
    st.ux = intt_ab3(st.ux, st.hx, st.itime, dt)
    st.uy = intt_ab3(st.uy, st.hy, st.itime, dt)
    st.uz = intt_ab3(st.uz, st.hz, st.itime, dt)
    pp3 = divergence(S, st.ux, st.uy, st.uz)
    st.divu = divu_stats(pp3)
    pp3 = poisson(S, pp3)
    px, py, pz = gradp(S, pp3)
    st.ux = st.ux - px
    st.uy = st.uy - py
    st.uz = st.uz - pz
    return st


# ---------------------------------------------------------------------------
# Case-Channel.f90 and the wall (ncl=2) branches of navier.f90
# ---------------------------------------------------------------------------
MASK64 = (1 << 64) - 1


def hash_uniform(seed, comp, idx):
    """Uniform [0, 1) numbers of the channel initial noise: the SplitMix64
    output function (Steele, Lea & Flood 2014) of the counter
    seed * 2^40 + comp * 2^36 + idx (idx < 2^36), top 53 bits.  The
    reference draws its noise with the compiler's random_number
    (Case-Channel.f90:105-112), which no other program reproduces; this
    counter-based generator is the same on the host and the device."""
    idx = np.asarray(idx, dtype=np.uint64)
    z = (np.uint64((seed << 40) & MASK64) + np.uint64(comp << 36) + idx) * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def channel_y(S: "Schemes"):
    """y - yly/2 of the velocity-grid points (Case-Channel.f90:100-102):
    yp(j) on a stretched mesh, (j-1) dy otherwise."""
    p = S.p
    if S.stret is not None:
        return S.stret["yp"] - p.yly * 0.5
    return np.arange(p.ny, dtype=np.float64) * p.dy - p.yly * 0.5


def init_channel(S: "Schemes", init_noise, seed):
    """Laminar Poiseuille profile plus modulated noise, iin=1 of
    init_channel (Case-Channel.f90:96-128) with idir_stream=1:
    u = init_noise * um * (2 r - 1) + 1 - y^2, v, w = init_noise * um *
    (2 r - 1), um = exp(-0.2 y^2); r from hash_uniform of the point index
    i + nx (j + ny k) and the component."""
    p = S.p
    y = channel_y(S)
    um = np.exp(-0.2 * y * y)[None, :, None]
    idx = np.arange(p.nx * p.ny * p.nz, dtype=np.uint64).reshape(p.nz, p.ny, p.nx)
    r = [hash_uniform(seed, c, idx) for c in range(3)]
    ux = init_noise * um * (2.0 * r[0] - 1.0) + (1.0 - y * y)[None, :, None]
    uy = init_noise * um * (2.0 * r[1] - 1.0)
    uz = init_noise * um * (2.0 * r[2] - 1.0)
    return ux, uy, uz


def channel_cfr(S: "Schemes", ux, constant):
    """channel_cfr (Case-Channel.f90:220-261): the bulk velocity
    ub = dy / (yly nx nz) sum u / ppy(j) (ppy = 1 on a uniform mesh,
    variables.f90:1298) and u - can with can = -(constant - ub).  Returns
    (new u, ub)."""
    p = S.p
    ppy = S.stret["ppy"] if S.stret is not None else np.ones(p.ny)
    coeff = p.dy / (p.yly * float(p.nx * p.nz))
    ub = float(np.sum(ux / ppy[None, :, None])) * coeff
    can = -(constant - ub)
    return ux - can, ub


@dataclass
class WallGrad:
    """dpdxy1, dpdzy1, dpdxyn, dpdzyn of navier.f90 (the pressure gradient
    on the y walls over gdt), shape (nz, nx)."""
    x1: np.ndarray
    z1: np.ndarray
    xn: np.ndarray
    zn: np.ndarray


def wall_grad_zero(p: Params):
    z = np.zeros((p.nz, p.nx))
    return WallGrad(z.copy(), z.copy(), z.copy(), z.copy())


def pre_correc_walls(p: Params, ux, uy, uz, wg: WallGrad, gdt):
    """ncly=2 branches of pre_correc (navier.f90:616-686, no MHD) with the
    no-slip wall velocity of the channel (byx1 = byy1 = byz1 = 0,
    Case-Channel.f90:96): u(i,1,k) = dpdxy1 gdt, v = 0, w = dpdzy1 gdt,
    and the same on j = ny.  Free-slip faces as pre_correc above."""
    ux, uy, uz = pre_correc(p, ux, uy, uz)
    if p.ncly1 == 2:
        ux[:, 0, :] = 0.0 + wg.x1 * gdt
        uy[:, 0, :] = 0.0
        uz[:, 0, :] = 0.0 + wg.z1 * gdt
    if p.nclyn == 2:
        ux[:, -1, :] = 0.0 + wg.xn * gdt
        uy[:, -1, :] = 0.0
        uz[:, -1, :] = 0.0 + wg.zn * gdt
    return ux, uy, uz


def gradp_walls(p: Params, px, pz, wg: WallGrad, gdt):
    """The wall planes gradp stores for the next pre_correc
    (navier.f90:459-479): dpdxy1 = px(:,1,:)/gdt, dpdzy1 = pz(:,1,:)/gdt."""
    if p.ncly1 == 2:
        wg.x1 = px[:, 0, :] / gdt
        wg.z1 = pz[:, 0, :] / gdt
    if p.nclyn == 2:
        wg.xn = px[:, -1, :] / gdt
        wg.zn = pz[:, -1, :] / gdt
    return wg


def ab3_gdt(dt):
    """gdt(1) of AB3: adt + bdt + cdt (variables.f90:1368-1372)."""
    adt, bdt, cdt = ab3_coefficients(dt)
    return adt + bdt + cdt


def step_channel(S: Schemes, st: State, wg: WallGrad, cfr=2.0 / 3.0, fcpg=None, wrotation=0.0, spinup_time=0):
    """One time step of the channel (xcompact3d.f90:46-87 with
    itype_channel): per sub-step boundary_conditions (the constant flow
    rate of channel_cfr when cfr is not None), the momentum right-hand side
    with momentum_forcing_channel (Case-Channel.f90:396-420: the constant
    pressure gradient fcpg on dux when fcpg is not None, then, while
    itime < spinup_time, the rotation dux -= wrotation uy, duy += wrotation
    ux of the sub-step's velocity), intt (AB3, or RK3 with st.hx = [f1,
    f2]), pre_correc with the wall branches, divergence, Poisson, gradp
    storing the wall planes, cor_vel."""
    p = S.p
    st.itime += 1
    dt = p.dt
    if p.itimescheme == 5:
        _, _, gdts = rk3_coefficients(dt)
        subs = [(itr, gdts[itr - 1]) for itr in (1, 2, 3)]
    else:
        subs = [(1, ab3_gdt(dt))]
    for itr, gdt in subs:
        if cfr is not None:
            st.ux, _ = channel_cfr(S, st.ux, cfr)
        dux, duy, duz = momentum_rhs(S, st.ux, st.uy, st.uz)
        if fcpg is not None:
            dux = dux + fcpg
        if st.itime < spinup_time:
            dux = dux - wrotation * st.uy
            duy = duy + wrotation * st.ux
        st.hx[0], st.hy[0], st.hz[0] = dux, duy, duz
        if p.itimescheme == 5:
            st.ux = intt_rk3(st.ux, st.hx, itr, dt)
            st.uy = intt_rk3(st.uy, st.hy, itr, dt)
            st.uz = intt_rk3(st.uz, st.hz, itr, dt)
        else:
            st.ux = intt_ab3(st.ux, st.hx, st.itime, dt)
            st.uy = intt_ab3(st.uy, st.hy, st.itime, dt)
            st.uz = intt_ab3(st.uz, st.hz, st.itime, dt)
        st.ux, st.uy, st.uz = pre_correc_walls(p, st.ux, st.uy, st.uz, wg, gdt)
        pp3 = divergence(S, st.ux, st.uy, st.uz)
        st.divu = divu_stats(pp3)
        pp3 = poisson(S, pp3)
        px, py, pz = gradp(S, pp3)
        wg = gradp_walls(p, px, pz, wg, gdt)
        st.ux = st.ux - px
        st.uy = st.uy - py
        st.uz = st.uz - pz
    return st, wg
