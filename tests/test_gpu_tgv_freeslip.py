"""The reference's own CTest on the GPU (tests/TGV-Taylor-Green-vortex of
the reference): TGV on a 65^3 free-slip box of side pi, Re=1600, RK3
(itimescheme=5), dt=0.005, 500 steps (ilast=500), statistics (KE, eps, eps2,
enstrophy) every 10 steps written at (itime-1)*dt -- compared with all 50 rows of its
reference_time_evol.dat (tests/golden/tgv_freeslip_rk3_time_evol.dat) at the
CTest's tolerance, numpy assert_allclose's default rtol 1e-7.  The FP32-bulk
solver is held to 1e-5 of the same rows.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREESLIP = dict(nclx1=1, nclxn=1, ncly1=1, nclyn=1, nclz1=1, nclzn=1)


def _run(X, precision, nsteps):
    p = X.make_params(65, 65, 65, xlx=math.pi, yly=math.pi, zlz=math.pi, re=1600.0, dt=0.005, itimescheme=5,
                      **FREESLIP)
    X.init(p)
    try:
        s = X.Solver(p, precision)
        s.init_tgv()
        rows = []
        for it in range(1, nsteps + 1):
            s.step(1)
            if it % 10 == 0:
                rows.append([(it - 1) * 0.005] + list(s.tgv_stats()))
        s.close()
    finally:
        X.finalize()
    return np.array(rows)


@pytest.mark.parametrize("precision", [0, 1])
def test_freeslip_rk3_ctest(torch_cuda, precision):
    import incompact3d_amd as X
    print(f"free-slip RK3 CTest, precision {precision}", flush=True)
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "tgv_freeslip_rk3_time_evol.dat"))
    got = _run(X, precision, 500)
    assert got.shape == ref.shape
    rtol = 1e-7 if precision == 0 else 1e-5
    for col in range(ref.shape[1]):
        np.testing.assert_allclose(got[:, col], ref[:, col], rtol=rtol, err_msg=f"column {col + 1}")


@pytest.mark.parametrize("precision", [0, 1])
def test_freeslip_x_periodic_yz_steps_vs_oracle(torch_cuda, precision):
    """poisson_100 inside the time step: the TGV field on a box free-slip in
    x ([0, pi], 33 points) and periodic in y and z (2 pi, 16 points), RK3,
    5 steps against the oracle's step_rk3, whose pressure solve is the
    poisson_100 restatement: FP64 1e-11, FP32 bulk 2e-5."""
    import incompact3d_amd as X
    from oracle import x3d_oracle as O
    from tests.helpers import rel_err
    kw = dict(nx=33, ny=16, nz=16, xlx=math.pi, re=1600.0, dt=0.005, itimescheme=5, nclx1=1, nclxn=1)
    p = O.Params(**kw)
    S = O.Schemes(p)
    st = O.tgv_state(p)
    for _ in range(5):
        O.step_rk3(S, st)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), precision)
        s.init_tgv()
        s.step(5)
        got = [t.cpu().numpy() for t in s.velocity()]
        s.close()
    finally:
        X.finalize()
    tol = 1e-11 if precision == 0 else 2e-5
    for g, r in zip(got, (st.ux, st.uy, st.uz)):
        assert rel_err(g, r) <= tol
