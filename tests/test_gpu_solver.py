"""GPU parity of the device time step (transeq + AB3 + divergence + Poisson
+ gradp/cor_vel) and of the TGV diagnostics against the oracle.

FP64: velocity after N steps within 1e-11 relative of the oracle's step and
diagnostics within 1e-11 relative.  FP32 bulk: the KE / dissipation history
stays within 1e-4 relative of the FP64 oracle on this short run.
"""
import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tgv32(torch_cuda):
    import incompact3d_amd as X
    kw = dict(nx=32, ny=32, nz=32, re=1600.0, dt=0.005)
    p = O.Params(**kw)
    X.init(X.make_params(**kw))
    yield p, O.Schemes(p), X
    X.finalize()


def test_tgv_initial_stats(tgv32):
    p, S, X = tgv32
    s = X.Solver(X.make_params(nx=p.nx, ny=p.ny, nz=p.nz, re=p.re, dt=p.dt), 0)
    s.init_tgv()
    st = O.tgv_state(p)
    ref = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    got = s.tgv_stats()
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-12 * max(abs(r), 1e-30), (got, ref)
    s.close()


def test_tgv_steps_fp64(tgv32):
    p, S, X = tgv32
    s = X.Solver(X.make_params(nx=p.nx, ny=p.ny, nz=p.nz, re=p.re, dt=p.dt), 0)
    s.init_tgv()
    st = O.tgv_state(p)
    for _ in range(12):
        O.step(S, st)
    s.step(12)
    ux, uy, uz = [t.cpu().numpy() for t in s.velocity()]
    for a, b in ((ux, st.ux), (uy, st.uy), (uz, st.uz)):
        assert rel_err(a, b) <= 1e-11
    ref = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    got = s.tgv_stats()
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-11 * abs(r), (got, ref)
    dmax, dmean = s.divergence_stats()
    assert dmax < 1e-10 and dmean < 1e-11
    s.close()


def test_tgv32_200_steps_match_oracle(tgv32):
    """200 FP64 steps on the GPU against 200 oracle steps (same 32^3 case)."""
    p, S, X = tgv32
    s = X.Solver(X.make_params(nx=p.nx, ny=p.ny, nz=p.nz, re=p.re, dt=p.dt), 0)
    s.init_tgv()
    s.step(200)
    st = O.tgv_state(p)
    for _ in range(200):
        O.step(S, st)
    got = s.tgv_stats()
    ref = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-10 * abs(r), (got, ref)
    ux = s.velocity()[0].cpu().numpy()
    assert rel_err(ux, st.ux) <= 1e-10
    s.close()


def test_tgv_steps_fp32_bulk(tgv32):
    p, S, X = tgv32
    s = X.Solver(X.make_params(nx=p.nx, ny=p.ny, nz=p.nz, re=p.re, dt=p.dt), 1)
    s.init_tgv()
    st = O.tgv_state(p)
    for _ in range(12):
        O.step(S, st)
    s.step(12)
    ref = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
    got = s.tgv_stats()
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-4 * abs(r), (got, ref)
    s.close()
