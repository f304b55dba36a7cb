"""GPU parity of the LES path against the oracle (oracle/les_oracle.py).

* Smagorinsky SGS force, nonconservative and conservative forms
  (les_models.f90:168-331, 1091-1250, 1327-1456): FP64 within 1e-11 relative.
* LES time steps (SGS added before AB3, transeq.f90:498-503): FP64 within
  1e-10 relative after 4 steps on a 32^3 grid (per-derivative path; the fused
  64^3 case is in test_gpu_les_fused.py).
* Box filter: within 1e-13.
* Closure network on BF16 MFMA vs the oracle's float64 evaluation of the same
  BF16 roundings: outputs within 2e-2 of the output scale at worst and 2e-3
  on average (a hidden activation can round to the neighbouring BF16 value
  when the FP32 sums differ in the last bits); gradients within 1e-2 relative
  (Frobenius, per layer) of the oracle with BF16 gradient roundings; one
  Adam step within 1e-6.
"""
import numpy as np
import pytest

from oracle import les_oracle as L
from oracle import x3d_oracle as O
from tests.helpers import dev_tensor as _dev, les_steps_vs_oracle, perturbed_tgv as _perturbed_tgv, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def les32(torch_cuda):
    import incompact3d_amd as X
    kw = dict(nx=32, ny=32, nz=32, re=1000.0, dt=0.002)
    X.init(X.make_params(**kw))
    p = O.Params(**kw)
    yield X, p, O.Schemes(p)
    X.finalize()


@pytest.mark.parametrize("iconserv", [0, 1])
def test_smagorinsky_sgs_matches_oracle(les32, iconserv):
    X, p, S = les32
    print(f"smagorinsky sgs iconserv={iconserv}", flush=True)
    u = _perturbed_tgv(p)
    s = X.Solver(X.make_params(p.nx, p.ny, p.nz, re=p.re, dt=p.dt), 0)
    s.set_velocity(*[_dev(c) for c in u])
    s.set_les(1, cs=0.17, iconserv=iconserv)
    got = [t.cpu().numpy() for t in s.sgs()]
    ref = L.smag_sgs(0.17, iconserv)(S, *u)
    for g, r in zip(got, ref):
        assert rel_err(g, r) <= 1e-11
    s.close()


def test_smagorinsky_les_steps_32(les32):
    X, p, _ = les32
    print("smagorinsky LES steps 32^3", flush=True)
    les_steps_vs_oracle(X, dict(nx=32, ny=32, nz=32, re=1000.0, dt=0.002), 4, 0)


def test_box_filter_matches_oracle(les32):
    X, p, _ = les32
    rng = np.random.default_rng(4)
    u = [rng.standard_normal((32, 32, 32)) for _ in range(3)]
    ub, tau = X.box_filter(*[_dev(c) for c in u], 4)
    rub, _, rtau = L.sgs_stress_target(*u, 4)
    for g, r in zip(ub, rub):
        assert np.max(np.abs(g.cpu().numpy() - r)) < 1e-13
    for g, k in zip(tau, L.COMPS):
        assert np.max(np.abs(g.cpu().numpy() - rtau[k])) < 1e-13


def _mlp(X, seed):
    params = L.init_params(seed)
    m = X.Mlp(L.flatten(params))
    return m, L.unflatten(L.flatten(params))


def test_mlp_forward_matches_oracle(les32):
    import torch
    X, _, _ = les32
    print("mlp forward", flush=True)
    m, params = _mlp(X, 3)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((3001, 10)).astype(np.float32)
    y = m.forward(torch.tensor(x, device="cuda")).cpu().numpy()
    ref = L.mlp_forward(params, x.astype(np.float64))
    scale = np.max(np.abs(ref))
    err = np.abs(y - ref)
    assert np.max(err) <= 2e-2 * scale and np.mean(err) <= 2e-3 * scale, (np.max(err), np.mean(err), scale)
    assert np.allclose(m.get_params(), L.flatten(params))
    m.close()


def test_mlp_gradient_matches_oracle(les32):
    import torch
    X, _, _ = les32
    print("mlp gradient", flush=True)
    m, params = _mlp(X, 5)
    rng = np.random.default_rng(9)
    n = 4099
    x = rng.standard_normal((n, 10)).astype(np.float32)
    t = rng.standard_normal((n, 6)).astype(np.float32) * 0.5
    g, loss = m.gradient(torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda"))
    y, cache = L.mlp_forward(params, x.astype(np.float64), keep=True)
    rloss, dy = L.mse_loss_grad(y, t.astype(np.float64))
    assert abs(loss - rloss) <= 1e-2 * rloss
    ref = L.mlp_backward(params, cache, dy, bf16_grads=True)
    got = L.unflatten(g)
    for li in range(4):
        for a, b in zip(got[li], ref[li]):
            a = np.asarray(a, dtype=np.float64)
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), (li, np.linalg.norm(a - b), np.linalg.norm(b))
    m.close()


def test_mlp_adam_step(les32):
    import torch
    X, _, _ = les32
    m, params = _mlp(X, 6)
    rng = np.random.default_rng(10)
    n = 2048
    xd = torch.tensor(rng.standard_normal((n, 10)).astype(np.float32), device="cuda")
    td = torch.tensor(rng.standard_normal((n, 6)).astype(np.float32), device="cuda")
    p0 = m.get_params().astype(np.float64)
    g, _ = m.gradient(xd, td)
    m.train_step(xd, td, 1e-3)
    p1 = m.get_params().astype(np.float64)
    ref, _, _ = L.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, lr=1e-3)
    assert np.max(np.abs(p1 - ref)) <= 1e-6
    m.close()


def test_mlp_training_reduces_loss(les32):
    import torch
    X, _, _ = les32
    print("mlp training", flush=True)
    m, _ = _mlp(X, 7)
    rng = np.random.default_rng(12)
    n = 16384
    x = rng.standard_normal((n, 10)).astype(np.float32)
    A = rng.standard_normal((10, 6)) * 0.4
    t = np.tanh(x.astype(np.float64) @ A).astype(np.float32)
    xd, td = torch.tensor(x, device="cuda"), torch.tensor(t, device="cuda")
    first = m.train_step(xd, td, 2e-3)
    last = first
    for _ in range(300):
        last = m.train_step(xd, td, 2e-3)
    assert last < 0.2 * first, (first, last)
    m.close()


@pytest.mark.parametrize("project", [0, 1])
def test_closure_matches_oracle(les32, project):
    """project 0: the network's stress (model 2); 1: its eddy-viscosity
    projection (model 3), both with the Smagorinsky fallback."""
    import torch
    X, _, _ = les32
    print(f"closure kernel project={project}", flush=True)
    m, params = _mlp(X, 13)
    rng = np.random.default_rng(14)
    n = 16
    g = {k: rng.standard_normal((n, n, n)) for k in L.GRAD_ORDER}
    gd = [torch.tensor(g[k], device="cuda") for k in L.GRAD_ORDER]
    gscale, delta, delta0, cs, beta = 1.3, 0.1, 0.1, 0.17, 0.0
    t, nfb = m.closure(gd, gscale, delta, delta0, cs, beta, project)
    ref, bad = L.closure_stress(params, g, delta, gscale, delta0, cs, beta, project=bool(project))
    # points where the fallback decision agrees
    s = L.strain(g)
    nut = (cs * delta) ** 2 * np.sqrt(2 * L.srt(s))
    got = {k: t[i].cpu().numpy() for i, k in enumerate(L.COMPS)}
    smag = {k: 2 * nut * s[k] for k in L.COMPS}
    is_smag = np.all([np.abs(got[k] - smag[k]) <= 1e-12 * (1 + np.abs(smag[k])) for k in L.COMPS], axis=0)
    agree = is_smag == bad
    assert agree.mean() >= 0.99
    assert abs(nfb - int(bad.sum())) <= 0.01 * bad.size
    scale = max(np.max(np.abs(ref[k])) for k in L.COMPS)
    for k in L.COMPS:
        err = np.abs(got[k] - ref[k])[agree]
        assert np.max(err) <= 2e-2 * scale and np.mean(err) <= 2e-3 * scale, (k, np.max(err), scale)
    m.close()


@pytest.mark.parametrize("model", [2, 3])
def test_learned_sgs_and_steps_match_oracle(les32, model):
    """x3d_solver_sgs with the learned closure (model 2: the network's
    stress; model 3: its eddy-viscosity projection) against the oracle's
    learned_sgs, and four AB3 LES steps against its step_les.  The fallback
    bound beta is set so high that no point falls back (the fallback is
    checked at the closure level above); what remains is the network in
    BF16 on MFMA against the oracle's FP64 evaluation of the same BF16
    roundings, carried through the stress divergence."""
    X, p, S = les32
    print(f"learned sgs model={model}", flush=True)
    m, params = _mlp(X, 17)
    u = _perturbed_tgv(p, seed=3)
    gscale, cs, beta = 1.0, 0.17, 1e6
    delta0 = L.filter_width(p)
    s = X.Solver(X.make_params(p.nx, p.ny, p.nz, re=p.re, dt=p.dt), 0)
    s.set_velocity(*[_dev(c) for c in u])
    s.set_les(model, cs=cs, beta=beta, mlp=m, gscale=gscale, delta0=delta0)
    got = [t.cpu().numpy() for t in s.sgs()]
    assert s.les_fallbacks() == 0
    fn = L.learned_sgs(params, gscale, delta0, cs, beta, project=(model == 3))
    ref = fn(S, *u)
    scale = max(np.abs(r).max() for r in ref)
    for g, r in zip(got, ref):
        err = np.abs(g - r)
        print(f"  force max err {err.max() / scale:.2e}, mean {err.mean() / scale:.2e} of {scale:.3e}")
        assert err.max() <= 3e-2 * scale and err.mean() <= 3e-3 * scale
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    for _ in range(4):
        L.step_les(S, st, fn)
    s.set_velocity(*[_dev(c) for c in u])
    s.step(4)
    vel = [t.cpu().numpy() for t in s.velocity()]
    s.close()
    m.close()
    for g, r in zip(vel, (st.ux, st.uy, st.uz)):
        e = rel_err(g, r)
        print(f"  velocity after 4 steps rel err {e:.2e}")
        assert e <= 1e-5
