"""The OpenMP CPU restatement of the periodic step (oracle/cport, the
benchmark's CPU baseline) against the numpy oracle: five AB3 steps of a
perturbed Taylor-Green field at 32^3, 1e-12 relative; and a planted error
(a wrong viscosity) is caught."""
import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import perturbed_tgv, rel_err


@pytest.fixture(scope="module")
def cport():
    from oracle import cport_step
    cport_step.build()
    return cport_step


def _run(cport, p, u, nsteps):
    S = O.Schemes(p)
    c = cport.CportStep(S)
    v = [np.ascontiguousarray(a.copy()) for a in u]
    c.step(*v, nsteps=nsteps)
    c.close()
    return v


def test_cport_matches_oracle(cport):
    p = O.Params(32, 32, 32, re=1600.0, dt=0.005)
    u = perturbed_tgv(p, seed=4)
    S = O.Schemes(p)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    for _ in range(5):
        O.step(S, st)
    got = _run(cport, p, u, 5)
    for g, r in zip(got, (st.ux, st.uy, st.uz)):
        assert rel_err(g, r) <= 1e-12
    assert cport.CportStep.threads() >= 1


def test_cport_detects_a_planted_error(cport):
    p = O.Params(32, 32, 32, re=1600.0, dt=0.005)
    u = perturbed_tgv(p, seed=4)
    S = O.Schemes(p)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    for _ in range(3):
        O.step(S, st)
    wrong = O.Params(32, 32, 32, re=1500.0, dt=0.005)
    got = _run(cport, wrong, u, 3)
    assert rel_err(got[0], st.ux) > 1e-7
