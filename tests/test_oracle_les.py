"""CPU pins of the LES oracle (oracle/les_oracle.py): BF16 rounding, the
closure network's backward pass and Adam against torch on the CPU (library
routines), box-filter invariants, the Smagorinsky closed form on a single
shear mode, and the equivalence of the nonconservative and conservative SGS
forms on smooth fields (a dropped term or wrong factor fails it)."""
import math

import numpy as np
import pytest

from oracle import les_oracle as L
from oracle import x3d_oracle as O


def test_bf16_matches_torch():
    import torch
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        # exact ties: the 16 dropped bits are 0x8000
                        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32).astype(np.float64)])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(L.bf16(x), ref)


def _torch_forward(params, x):
    import torch
    import torch.nn.functional as F

    def rnd(t):   # BF16 rounding with a straight-through gradient
        return t + (t.to(torch.float32).to(torch.bfloat16).to(torch.float64) - t).detach()

    ws = [torch.tensor(w.astype(np.float64), requires_grad=True) for w, _ in params]
    bs = [torch.tensor(b.astype(np.float64), requires_grad=True) for _, b in params]
    h = rnd(torch.tensor(x))
    for i in range(4):
        z = h @ rnd(ws[i]).T + bs[i]
        h = rnd(F.gelu(z, approximate="tanh")) if i < 3 else z
    return h, ws, bs


def test_mlp_backward_matches_autograd():
    import torch
    rng = np.random.default_rng(0)
    params = L.init_params(1)
    x = rng.standard_normal((257, 10))
    y, cache = L.mlp_forward(params, x, keep=True)
    yt, ws, bs = _torch_forward(params, x)
    assert np.max(np.abs(yt.detach().numpy() - y)) < 1e-12
    dy = rng.standard_normal(y.shape)
    grads = L.mlp_backward(params, cache, dy)
    (yt * torch.tensor(dy)).sum().backward()
    for i in range(4):
        gw, gb = grads[i]
        assert np.max(np.abs(gw - ws[i].grad.numpy())) <= 1e-10 * max(1.0, np.max(np.abs(gw)))
        assert np.max(np.abs(gb - bs[i].grad.numpy())) <= 1e-10 * max(1.0, np.max(np.abs(gb)))


def test_gelu_grad_finite_difference():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (L.gelu(z + h) - L.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - L.gelu_grad(z))) < 1e-8


def test_gelu_is_torch_tanh_gelu_and_close_to_erf_gelu():
    """The activation is torch's gelu(approximate="tanh") to roundoff, and
    within 4.8e-4 of the erf form z Phi(z) (a wrong constant in the tanh
    argument moves either bound by orders of magnitude)."""
    import torch
    import torch.nn.functional as F
    from scipy.special import erf
    z = np.linspace(-12, 12, 24001)
    ref = F.gelu(torch.tensor(z), approximate="tanh").numpy()
    assert np.max(np.abs(L.gelu(z) - ref)) < 1e-14
    exact = 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    dev = np.max(np.abs(L.gelu(z) - exact))
    assert 4.0e-4 < dev < 4.8e-4


def test_adam_matches_torch():
    import torch
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(4)]
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, start=1):
        opt.zero_grad()
        t.grad = torch.tensor(g)
        opt.step()
        p, m, v = L.adam_step(p, g, m, v, k, lr=3e-3)
    assert np.max(np.abs(p - t.detach().numpy())) < 1e-12


def test_flatten_roundtrip():
    ps = L.init_params(2)
    flat = L.flatten(ps)
    assert flat.size == 9414
    back = L.unflatten(flat)
    for (w, b), (w2, b2) in zip(ps, back):
        assert np.array_equal(w, w2) and np.array_equal(b, b2)


def test_box_filter_invariants():
    rng = np.random.default_rng(7)
    f = 4
    u = [rng.standard_normal((16, 12, 8)) for _ in range(3)]
    ub, t, tau = L.sgs_stress_target(*u, f)
    for c in range(3):
        assert abs(ub[c].mean() - u[c].mean()) < 1e-14
    # brute force on one coarse cell
    i, j, k = 1, 2, 3
    box = (slice(f * k, f * k + f), slice(f * j, f * j + f), slice(f * i, f * i + f))
    assert abs(ub[0][k, j, i] - u[0][box].mean()) < 1e-14
    assert abs(tau["xy"][k, j, i] - ((u[0][box] * u[1][box]).mean() - u[0][box].mean() * u[1][box].mean())) < 1e-14
    for c in ("xx", "yy", "zz"):
        assert np.all(tau[c] >= -1e-14)
    # deviatoric target
    assert np.max(np.abs(t["xx"] + t["yy"] + t["zz"])) < 1e-13
    # box-constant fields have no subfilter stress
    const = [np.kron(rng.standard_normal((4, 3, 2)), np.ones((f, f, f))) for _ in range(3)]
    _, t0, _ = L.sgs_stress_target(*const, f)
    assert max(np.max(np.abs(v)) for v in t0.values()) < 1e-13


def test_smag_single_shear_mode_closed_form():
    n = 16
    p = O.Params(n, n, n, re=1000.0, dt=0.001)
    S = O.Schemes(p)
    z = np.arange(n) * p.dz
    Z = np.broadcast_to(z[:, None, None], (n, n, n))
    ux = np.sin(Z).copy()
    zero = np.zeros_like(ux)
    cs = 0.17
    nut, s = L.smag(S, ux, zero, zero, cs)
    # modified wavenumber of the periodic 6th-order first derivative at k=1
    h = p.dz
    afi, bfi, alfa = (7.0 / 9.0) / h, (1.0 / 36.0) / h, 1.0 / 3.0
    kp = 2.0 * (afi * math.sin(h) + bfi * math.sin(2 * h)) / (1.0 + 2.0 * alfa * math.cos(h))
    delta = (p.dx * p.dy * p.dz) ** (1.0 / 3.0)
    ref = (cs * delta) ** 2 * kp * np.abs(np.cos(Z))
    assert np.max(np.abs(nut - ref)) < 1e-13 * np.max(ref) + 1e-18
    assert np.max(np.abs(s["xz"] - 0.5 * kp * np.cos(Z))) < 1e-13


def test_sgs_forms_agree_on_smooth_fields():
    n = 32
    p = O.Params(n, n, n, re=1000.0, dt=0.001)
    S = O.Schemes(p)
    ux, uy, uz = O.init_tgv(p)   # divergence free
    x = np.arange(n) * p.dx
    X = x[None, None, :]
    Y = x[None, :, None]
    Zc = x[:, None, None]
    nut = 1e-3 * (1.0 + 0.3 * np.sin(X) * np.cos(Y) * np.sin(Zc) + 0.2 * np.cos(2 * Zc))
    nut = np.broadcast_to(nut, ux.shape).copy()
    s = L.strain(L.gradients(S, ux, uy, uz))
    a = L.sgs_nonconservative(S, ux, uy, uz, nut, s)
    b = L.sgs_conservative(S, nut, s)
    scale = max(np.max(np.abs(c)) for c in a)
    for ca, cb in zip(a, b):
        assert np.max(np.abs(ca - cb)) < 1e-4 * scale
    # a wrong factor on one strain component is visible at this tolerance
    s2 = {k: (2.0 * v if k == "xz" else v) for k, v in s.items()}
    bad = L.sgs_conservative(S, nut, s2)
    assert max(np.max(np.abs(ca - cb)) for ca, cb in zip(a, bad)) > 1e-2 * scale


def test_closure_fallback_selects_smagorinsky_where_energy_injected():
    n = 8
    rng = np.random.default_rng(11)
    g = {k: rng.standard_normal((n, n, n)) for k in L.GRAD_ORDER}
    params = L.init_params(4)
    t, bad = L.closure_stress(params, g, 0.1, 1.0, 0.1, 0.17, 0.0)
    s = L.strain(g)
    ss = L.srt(s)
    nut = (0.17 * 0.1) ** 2 * np.sqrt(2 * ss)
    pi = sum(t[k] * s[k] * (1.0 if k in ("xx", "yy", "zz") else 2.0) for k in L.COMPS)
    assert np.all(pi[bad] >= 0.0)           # replaced points dissipate
    assert np.all(pi[~bad] >= -1e-12)       # kept points never inject (beta = 0)
    assert np.allclose(t["xy"][bad], 2 * nut[bad] * s["xy"][bad])


def _gradient_tensor(g):
    """Full 3x3 velocity-gradient tensor G[i][j] = d u_i / d x_j from the
    named gradients (name = component then direction)."""
    ax = "xyz"
    return [[g[ax[i] + ax[j]] for j in range(3)] for i in range(3)]


@pytest.mark.parametrize("beta", [0.0, 0.5])
@pytest.mark.parametrize("project", [False, True])
def test_closure_stress_composition_closed_form(beta, project):
    """A network whose last layer is zero returns its bias c, so the closure
    is T_ij = c_ij delta^2 gscale^2 whatever the hidden layers do.  The
    stress, the fallback mask and the strain projection then have closed
    forms, written here with full symmetric 3x3 tensors (T:S = sum_ij T_ij
    S_ij, no factor-2 bookkeeping)."""
    n = 8
    rng = np.random.default_rng(21)
    g = {k: rng.standard_normal((n, n, n)) for k in L.GRAD_ORDER}
    params = L.init_params(5)
    c = np.array([0.5, -0.25, 0.125, 0.375, -0.5, 0.0625], dtype=np.float32)   # BF16-exact
    params[3] = (np.zeros_like(params[3][0]), c)
    delta, gscale, delta0, cs = 0.1, 1.3, 0.07, 0.17
    got, bad = L.closure_stress(params, g, delta, gscale, delta0, cs, beta, project)
    scale = delta * delta * gscale * gscale
    idx = {"xx": 0, "yy": 1, "zz": 2, "xy": 3, "xz": 4, "yz": 5}
    cm = [[float(c[idx["".join(sorted("xyz"[i] + "xyz"[j]))]]) for j in range(3)] for i in range(3)]
    G = _gradient_tensor(g)
    Sm = [[0.5 * (G[i][j] + G[j][i]) for j in range(3)] for i in range(3)]
    Tm = [[np.full((n, n, n), cm[i][j] * scale) for j in range(3)] for i in range(3)]
    ts = sum(Tm[i][j] * Sm[i][j] for i in range(3) for j in range(3))
    ss = sum(Sm[i][j] * Sm[i][j] for i in range(3) for j in range(3))
    nut = (cs * delta) ** 2 * np.sqrt(2.0 * ss)
    want_bad = ts < -beta * 2.0 * nut * ss
    assert np.array_equal(bad, want_bad)
    assert 0 < want_bad.sum() < want_bad.size
    for name, (i, j) in L.PAIRS.items():
        keep = 2.0 * np.maximum(ts, 0.0) / (2.0 * ss) * Sm[i][j] if project else Tm[i][j]
        want = np.where(want_bad, 2.0 * nut * Sm[i][j], keep)
        assert np.max(np.abs(got[name] - want)) <= 1e-13 * np.max(np.abs(want)), name
    if project:
        # the projection keeps the dissipation T:S of the kept points
        tsg = sum(got[k] * L.strain(g)[k] * (1.0 if k in ("xx", "yy", "zz") else 2.0) for k in L.COMPS)
        assert np.allclose(tsg[~want_bad], np.maximum(ts, 0.0)[~want_bad], rtol=1e-12, atol=1e-15)


def test_closure_feature_order_single_gradient():
    """u_i = sin(x_j) has the single nonzero gradient d u_i / d x_j, which
    must land in feature 3*j + i (the documented device layout: g_xx g_yx
    g_zx g_xy ...); the last feature is delta/delta0."""
    n = 16
    p = O.Params(n, n, n)
    S = O.Schemes(p)
    coords = np.meshgrid(np.arange(n) * p.dz, np.arange(n) * p.dy, np.arange(n) * p.dx, indexing="ij")
    xyz = (coords[2], coords[1], coords[0])
    for i in range(3):
        for j in range(3):
            u = [np.zeros((n, n, n)) for _ in range(3)]
            u[i] = np.sin(xyz[j])
            f = L.features(L.gradients(S, *u), 0.3, 2.0, 0.6)
            nz = [k for k in range(9) if np.any(f[:, k] != 0.0)]
            assert nz == [3 * j + i], (i, j, nz)
            assert np.allclose(f[:, 3 * j + i], np.cos(xyz[j]).ravel() / 2.0, atol=1e-4)
            assert np.all(f[:, 9] == 0.5)


def test_learned_sgs_model3_is_eddy_viscosity_divergence():
    """Model 3's force equals sgs_mom_conservative (les_models.f90:1327-1456)
    with nu_t = max(T:S, 0)/(2 S:S) taken from the network's stress."""
    n = 16
    p = O.Params(n, n, n, re=1000.0)
    S = O.Schemes(p)
    from tests.helpers import perturbed_tgv
    u = perturbed_tgv(p, seed=3, amp=0.2)
    params = L.init_params(8)
    gscale, delta0, cs, beta = 1.0, L.filter_width(p), 0.17, 1e30
    force = L.learned_sgs(params, gscale, delta0, cs, beta, project=True)(S, *u)
    g = L.gradients(S, *u)
    t2, _ = L.closure_stress(params, g, L.filter_width(p), gscale, delta0, cs, beta, project=False)
    s = L.strain(g)
    ss = L.srt(s)
    ts = sum(t2[k] * s[k] * (1.0 if k in ("xx", "yy", "zz") else 2.0) for k in L.COMPS)
    nut = np.where((ss > 0) & (ts > 0), ts / (2.0 * np.where(ss > 0, ss, 1.0)), 0.0)
    want = L.sgs_conservative(S, nut, s)
    for a, b in zip(force, want):
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-300)
