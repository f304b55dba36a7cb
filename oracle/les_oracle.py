"""Numpy restatement of the LES subgrid-stress path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may use this module, and only as the checker.  The HIP library never
calls it.

Arrays follow ``x3d_oracle``: shape (nz, ny, nx), axis 2 = x.

Contents and pins (tests/test_oracle_les.py):

* ``smag`` -- les_models.f90:168-331 (jles=1, no ABL damping): the strain
  tensor from the nine gradients with the reference's parities, srt = S:S
  (off-diagonals twice), nut = (Cs*del)^2 * sqrt(2*srt), del = (dx*dy*dz)^(1/3)
  (init_explicit_les, les_models.f90:36-39).  Pinned by the closed form of a
  single shear mode (nut = (Cs del)^2 |k'| |cos|, k' the compact scheme's
  modified wavenumber).
* ``sgs_nonconservative`` -- les_models.f90:1091-1250 (istret=0):
  sgs_i = nut lap(u_i) + 2 S_ij d_j nut.  ``sgs_conservative`` --
  les_models.f90:1327-1456: sgs_i = d_j (2 nut S_ij).  Pinned against each
  other: for a smooth divergence-free field and a smooth nut the two forms are
  the same operator up to the schemes' truncation error, so a dropped term, a
  wrong factor or a wrong index fails the comparison.
* ``box_filter`` / ``sgs_stress_target`` -- top-hat filter over f^3 fine
  cells onto the coarse grid and T_ij = -(bar(u_i u_j) - bar(u_i) bar(u_j))
  (deviatoric part).  Pinned by invariants: the filter preserves the mean, is
  exact on box-constant fields (T = 0 there) and tau_ii >= 0.
* ``mlp_forward`` / ``mlp_backward`` -- the closure network 10-64-64-64-6
  with GELU (0.5 x (1 + erf(x/sqrt 2))), evaluated in float64 on the BF16
  roundings the device uses (inputs and hidden activations rounded to BF16
  before each product, weights held as BF16).  Pinned by central finite
  differences and by torch.autograd on CPU.
* ``adam_step`` -- Kingma & Ba (2015), Algorithm 1, with bias correction.
  Pinned by torch.optim.Adam on CPU.
* ``closure_stress`` -- the learned stress with the Smagorinsky fallback where
  the predicted stress would feed energy to the resolved scales beyond the
  bound: Pi = T_ij S_ij < -beta * 2 nut S:S; optionally projected on the
  strain (eddy-viscosity form, the solver's model 3).  No reference
  counterpart.  Pinned as a composition: a network whose last layer is zero
  returns its bias, so the stress, the fallback mask and the projection have
  closed forms written with full 3x3 tensors (T:S = trace(T S)); the feature
  order is pinned by velocity fields with a single nonzero gradient
  component; the network itself is pinned against torch.
* ``learned_sgs`` -- gradients, closure stress and the conservative
  divergence composed; pinned through the pieces above and
  ``stress_divergence`` (the conservative Smagorinsky pin).
"""
from __future__ import annotations

import math

import numpy as np

from . import x3d_oracle as O


# ---------------------------------------------------------------------------
# les_models.f90: Smagorinsky
# ---------------------------------------------------------------------------
def filter_width(p: O.Params):
    """del of init_explicit_les (les_models.f90:36-39), uniform y."""
    return (p.dx * p.dy * p.dz) ** (1.0 / 3.0)


def gradients(S: O.Schemes, ux, uy, uz):
    """The nine gradients of smag (les_models.f90:210-238) with the
    reference's parities: gxx=dux/dx (0), gyx=duy/dx (1), gzx=duz/dx (1),
    gxy=dux/dy (1), gyy (0), gzy (1), gxz=dux/dz (1), gyz (1), gzz (0)."""
    d1 = O.der1
    g = {
        "xx": d1(S, ux, "x", 0), "yx": d1(S, uy, "x", 1), "zx": d1(S, uz, "x", 1),
        "xy": d1(S, ux, "y", 1), "yy": d1(S, uy, "y", 0), "zy": d1(S, uz, "y", 1),
        "xz": d1(S, ux, "z", 1), "yz": d1(S, uy, "z", 1), "zz": d1(S, uz, "z", 0),
    }
    return g


def strain(g):
    """S_ij of smag (les_models.f90:214-269): sxx=gxx, sxy=(gxy+gyx)/2, ..."""
    return {
        "xx": g["xx"], "yy": g["yy"], "zz": g["zz"],
        "xy": 0.5 * (g["xy"] + g["yx"]),
        "xz": 0.5 * (g["zx"] + g["xz"]),
        "yz": 0.5 * (g["yz"] + g["zy"]),
    }


def srt(s):
    """srt_smag (les_models.f90:271): S:S with the off-diagonals counted twice."""
    return (s["xx"] * s["xx"] + s["yy"] * s["yy"] + s["zz"] * s["zz"] + 2.0 * s["xy"] * s["xy"]
            + 2.0 * s["xz"] * s["xz"] + 2.0 * s["yz"] * s["yz"])


def smag(S: O.Schemes, ux, uy, uz, cs):
    """nut and S_ij of smag (les_models.f90:168-331), istret=0, no ABL."""
    s = strain(gradients(S, ux, uy, uz))
    length = cs * filter_width(S.p)
    nut = (length ** 2) * np.sqrt(2.0 * srt(s))
    return nut, s


def sgs_nonconservative(S: O.Schemes, ux, uy, uz, nut, s):
    """sgs_mom_nonconservative (les_models.f90:1091-1250), istret=0:
    sgs_i = nut lap(u_i) + 2 S_ij d_j nut, accumulated x, y, z."""
    d1, d2 = O.der1, O.der2
    # x pencils (:1123-1131)
    ta = d1(S, nut, "x", 1)
    sx = d2(S, ux, "x", 0) * nut + 2.0 * s["xx"] * ta
    sy = d2(S, uy, "x", 1) * nut + 2.0 * s["xy"] * ta
    sz = d2(S, uz, "x", 1) * nut + 2.0 * s["xz"] * ta
    # y pencils (:1149-1196)
    ta = d1(S, nut, "y", 1)
    sx = sx + nut * d2(S, ux, "y", 1) + 2.0 * s["xy"] * ta
    sy = sy + nut * d2(S, uy, "y", 0) + 2.0 * s["yy"] * ta
    sz = sz + nut * d2(S, uz, "y", 1) + 2.0 * s["yz"] * ta
    # z pencils (:1210-1218)
    ta = d1(S, nut, "z", 1)
    sx = sx + nut * d2(S, ux, "z", 1) + 2.0 * s["xz"] * ta
    sy = sy + nut * d2(S, uy, "z", 1) + 2.0 * s["yz"] * ta
    sz = sz + nut * d2(S, uz, "z", 0) + 2.0 * s["zz"] * ta
    return sx, sy, sz


def stress_divergence(S: O.Schemes, t):
    """d_j T_ij with the parities of sgs_mom_conservative
    (les_models.f90:1384-1440): T given as a dict of the six components."""
    d1 = O.der1
    sx = d1(S, t["xx"], "x", 0)
    sy = d1(S, t["xy"], "x", 0)
    sz = d1(S, t["xz"], "x", 0)
    sx = sx + d1(S, t["xy"], "y", 0)
    sy = sy + d1(S, t["yy"], "y", 1)
    sz = sz + d1(S, t["yz"], "y", 0)
    sx = sx + d1(S, t["xz"], "z", 0)
    sy = sy + d1(S, t["yz"], "z", 0)
    sz = sz + d1(S, t["zz"], "z", 0)
    return sx, sy, sz


def sgs_conservative(S: O.Schemes, nut, s):
    """sgs_mom_conservative (les_models.f90:1327-1456): T_ij = 2 nut S_ij."""
    return stress_divergence(S, {k: 2.0 * nut * v for k, v in s.items()})


# ---------------------------------------------------------------------------
# one AB3 step with an SGS force added to the RHS (transeq.f90:498-503)
# ---------------------------------------------------------------------------
def step_les(S: O.Schemes, st: O.State, sgs_fn):
    st.itime += 1
    dux, duy, duz = O.momentum_rhs(S, st.ux, st.uy, st.uz)
    gx, gy, gz = sgs_fn(S, st.ux, st.uy, st.uz)
    st.hx[0], st.hy[0], st.hz[0] = dux + gx, duy + gy, duz + gz
    dt = S.p.dt
    st.ux = O.intt_ab3(st.ux, st.hx, st.itime, dt)
    st.uy = O.intt_ab3(st.uy, st.hy, st.itime, dt)
    st.uz = O.intt_ab3(st.uz, st.hz, st.itime, dt)
    pp3 = O.poisson(S, O.divergence(S, st.ux, st.uy, st.uz))
    px, py, pz = O.gradp(S, pp3)
    st.ux, st.uy, st.uz = st.ux - px, st.uy - py, st.uz - pz
    return st


def smag_sgs(cs, iconserv=0):
    def fn(S, ux, uy, uz):
        nut, s = smag(S, ux, uy, uz, cs)
        if iconserv == 0:
            return sgs_nonconservative(S, ux, uy, uz, nut, s)
        return sgs_conservative(S, nut, s)
    return fn


# ---------------------------------------------------------------------------
# box filter and the subgrid stress of the filtered DNS
# ---------------------------------------------------------------------------
COMPS = ("xx", "yy", "zz", "xy", "xz", "yz")
PAIRS = {"xx": (0, 0), "yy": (1, 1), "zz": (2, 2), "xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}


def box_filter(u, f):
    """Top-hat average over f^3 fine cells: (nz,ny,nx) -> (nz/f,ny/f,nx/f);
    coarse point I averages fine points f*I .. f*I+f-1 along each axis."""
    nz, ny, nx = u.shape
    return u.reshape(nz // f, f, ny // f, f, nx // f, f).mean(axis=(1, 3, 5))


def sgs_stress_target(ux, uy, uz, f):
    """Filtered velocity and T_ij = -dev(bar(u_i u_j) - bar(u_i) bar(u_j))."""
    u = (ux, uy, uz)
    ub = tuple(box_filter(c, f) for c in u)
    tau = {k: box_filter(u[i] * u[j], f) - ub[i] * ub[j] for k, (i, j) in PAIRS.items()}
    tr = (tau["xx"] + tau["yy"] + tau["zz"]) / 3.0
    t = {k: -(tau[k] - (tr if k in ("xx", "yy", "zz") else 0.0)) for k in COMPS}
    return ub, t, tau


# ---------------------------------------------------------------------------
# closure network (BF16 storage, float64 arithmetic here)
# ---------------------------------------------------------------------------
LAYERS = ((10, 64), (64, 64), (64, 64), (64, 6))
GRAD_ORDER = ("xx", "yx", "zx", "xy", "yy", "zy", "xz", "yz", "zz")


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    if np.any(nan):
        out[nan] = np.nan
    return out


# GELU of the closure network: the tanh form of Hendrycks & Gimpel (2016),
# 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))), which differs from the
# erf form z Phi(z) by at most 4.8e-4 (torch's gelu(approximate="tanh")).
GELU_K = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_K * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    th = np.tanh(GELU_K * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * GELU_K * (1.0 + 3.0 * GELU_A * z * z)


def init_params(seed=0):
    """He-style uniform initialisation (FP32 master weights)."""
    rng = np.random.default_rng(seed)
    ps = []
    for fan_in, fan_out in LAYERS:
        lim = math.sqrt(6.0 / fan_in) * (0.5 if fan_out == 6 else 1.0)
        w = rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)
        b = np.zeros(fan_out, dtype=np.float32)
        ps.append((w, b))
    return ps


def mlp_forward(params, x, keep=False):
    """x: (N, 10) features -> (N, 6).  Each product uses the BF16 rounding of
    its two operands (weights and the layer input), sums and biases in high
    precision; hidden layers apply GELU."""
    h = bf16(x)
    cache = []
    nl = len(params)
    for li, (w, b) in enumerate(params):
        z = h @ bf16(w).T + b.astype(np.float64)
        cache.append((h, z))
        if li < nl - 1:
            h = bf16(gelu(z))
        else:
            h = z
    return (h, cache) if keep else h


def mlp_backward(params, cache, dy, bf16_grads=False):
    """Gradients of sum(dy * y) w.r.t. every weight and bias (float64),
    given the forward cache; dy: (N, 6).  The BF16 roundings of the forward
    pass pass gradients straight through; the backward products use the BF16
    weights, as the device does.  bf16_grads=True also rounds each layer's
    output gradient dZ to BF16 before it is used, as the device stores it."""
    grads = [None] * len(params)
    g = dy.astype(np.float64)
    for li in range(len(params) - 1, -1, -1):
        w, _ = params[li]
        h, z = cache[li]
        if li < len(params) - 1:
            g = g * gelu_grad(z)
        if bf16_grads:
            g = bf16(g)
        grads[li] = (g.T @ h, g.sum(axis=0))
        if li > 0:
            g = g @ bf16(w)
    return grads


def flatten(params):
    """W1 b1 W2 b2 W3 b3 W4 b4 as one FP32 vector (the device layout)."""
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in params]).astype(np.float32)


def unflatten(flat):
    out, o = [], 0
    for fan_in, fan_out in LAYERS:
        w = np.asarray(flat[o:o + fan_in * fan_out], dtype=np.float32).reshape(fan_out, fan_in)
        o += fan_in * fan_out
        b = np.asarray(flat[o:o + fan_out], dtype=np.float32)
        o += fan_out
        out.append((w, b))
    return out


def mse_loss_grad(y, t):
    """L = mean((y - t)^2) over all N*6 entries; returns (L, dL/dy)."""
    d = y - t
    n = d.size
    return float(np.sum(d * d) / n), 2.0 * d / n


def adam_step(param, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), returns new arrays."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return param - lr * mh / (np.sqrt(vh) + eps), m, v


def features(g, delta, gscale, delta0):
    """(N,10) network input: the nine gradients over gscale, delta/delta0."""
    cols = [g[k].ravel() / gscale for k in GRAD_ORDER]
    cols.append(np.full(cols[0].shape, delta / delta0))
    return np.stack(cols, axis=1)


def closure_stress(params, g, delta, gscale, delta0, cs, beta, project=False):
    """Learned T_ij (scaled back by delta^2 gscale^2) with the Smagorinsky
    stress 2 nut S_ij where T_ij S_ij < -beta * 2 nut S:S.

    project=False is the solver's model 2 (the network's full stress).
    project=True is model 3: where the fallback does not apply, only the part
    of T along the strain is kept, T = 2 nu_t S with the eddy viscosity
    nu_t = max(T:S, 0) / (2 S:S) (0 where S:S = 0), so the closure dissipates
    like an eddy-viscosity model of the conservative form
    (les_models.f90:1327-1456) with a learned nu_t."""
    y = mlp_forward(params, features(g, delta, gscale, delta0))
    scale = delta * delta * gscale * gscale
    shape = g["xx"].shape
    t = {k: (y[:, i] * scale).reshape(shape) for i, k in enumerate(COMPS)}
    s = strain(g)
    ss = srt(s)
    nut = (cs * delta) ** 2 * np.sqrt(2.0 * ss)
    pi = sum(t[k] * s[k] * (1.0 if k in ("xx", "yy", "zz") else 2.0) for k in COMPS)
    bad = pi < -beta * 2.0 * nut * ss
    if project:
        ok = (ss > 0.0) & (pi > 0.0)
        nt = np.where(ok, pi / np.where(ok, 2.0 * ss, 1.0), 0.0)
        t = {k: 2.0 * nt * s[k] for k in COMPS}
    out = {k: np.where(bad, 2.0 * nut * s[k], t[k]) for k in COMPS}
    return out, bad


def learned_sgs(params, gscale, delta0, cs, beta, project=False):
    """SGS force of the solver's learned closure (models 2 and 3): the nine
    gradients of smag (les_models.f90:210-238), the closure stress above, and
    its divergence with the parities of sgs_mom_conservative
    (les_models.f90:1384-1440).  Returns fn(S, ux, uy, uz) for step_les."""
    def fn(S, ux, uy, uz):
        g = gradients(S, ux, uy, uz)
        t, _ = closure_stress(params, g, filter_width(S.p), gscale, delta0, cs, beta, project)
        return stress_divergence(S, t)
    return fn
