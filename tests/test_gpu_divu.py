"""The statistics the reference's divergence computes on every call
(navier.f90:348-361): max and mean |.| of div u* over the pressure grid,
u* the velocity before the pressure correction, through
x3d_solver_divu_stats, against the oracle's steps (x3d_oracle.divu_stats).

* the fused periodic step (the z divergence sweep reduces them per block),
  64^3 AB3, FP64 and FP32 bulk;
* the per-operator step, 32^3 periodic AB3 and the 33 x 16 x 16 box free-slip
  in x (RK3, the last sub-step's values);
* the channel (walls in y, stretched, RK3).
Bars: FP64 1e-10 relative; FP32 bulk 2e-3 relative (div u* is a difference
of O(1) derivatives of the FP32 velocity)."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")


def _check(got, ref, precision):
    tol = 1e-10 if precision == 0 else 2e-3
    assert abs(got[0] - ref[0]) <= tol * abs(ref[0]), (got, ref)
    assert abs(got[1] - ref[1]) <= tol * abs(ref[1]), (got, ref)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["fused64", "per32", "freeslip_x_rk3"])
def test_divu_stats_match_oracle(case, precision):
    from oracle import x3d_oracle as O
    kw = {"fused64": dict(nx=64, ny=64, nz=64, re=1600.0, dt=0.005),
          "per32": dict(nx=32, ny=32, nz=32, re=1600.0, dt=0.005),
          "freeslip_x_rk3": dict(nx=33, ny=16, nz=16, xlx=math.pi, re=1600.0, dt=0.005, itimescheme=5,
                                 nclx1=1, nclxn=1)}[case]
    p = O.Params(**kw)
    S = O.Schemes(p)
    st = O.tgv_state(p)
    nsteps = 4
    for _ in range(nsteps):
        (O.step_rk3 if kw.get("itimescheme", 3) == 5 else O.step)(S, st)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), precision)
        s.init_tgv()
        with pytest.raises(X.X3DError):
            s.divu_stats()
        s.step(nsteps)
        got = s.divu_stats()
        s.close()
    finally:
        X.finalize()
    _check(got, st.divu, precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_divu_stats_channel(precision):
    from oracle import x3d_oracle as O
    kw = dict(nx=16, ny=33, nz=16, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2, istret=2,
              beta=0.259, re=500.0, dt=0.002, itimescheme=5, isecondder=5)
    p = O.Params(**kw)
    S = O.Schemes(p)
    u = O.init_channel(S, 0.125, 5)
    st = O.State(u[0].copy(), u[1].copy(), u[2].copy(), [u[0].copy()] * 3, [u[1].copy()] * 3,
                 [u[2].copy()] * 3, 0)
    wg = O.wall_grad_zero(p)
    for _ in range(3):
        st, wg = O.step_channel(S, st, wg)
    X.init(X.make_params(**kw))
    try:
        s = X.Solver(X.make_params(**kw), precision)
        s.set_velocity(*[torch.tensor(c, device="cuda") for c in u])
        s.set_forcing(1, 2.0 / 3.0)
        s.step(3)
        got = s.divu_stats()
        s.close()
    finally:
        X.finalize()
    _check(got, st.divu, precision)
