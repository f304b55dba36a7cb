"""GPU parity of the fused time step -- momentum sweeps (one launch per sweep
direction and velocity component, AB3 fused into the x sweep) and projection
sweeps (one launch per direction for the divergence and for gradp + cor_vel)
-- against the oracle's time step (transeq.f90:73-470 +
time_integrators.f90:85-100 + navier.f90), and against the unfused path (one
line-solve launch per derivative, selected with X3D_UNFUSED_MOMENTUM at
solver creation).

Grids: the fused sweeps need nx in {64,128,256,512} and ny, nz <= 512; the
non-cubic cases put a different segment count on each sweep.
FP64 tolerance: 1e-11 relative against the oracle, 1e-12 between the two
device paths.  FP32 bulk: diagnostics within 1e-4 of the FP64 oracle.
"""
import os

import numpy as np
import pytest

from oracle import x3d_oracle as O
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

GRIDS = [(64, 64, 64), (128, 32, 16), (64, 128, 32),
         # the line lengths of the 256^3 step and beyond: x sweeps with 4 and
         # 8 points per lane, y and z sweeps of 32 and 64 segments
         (256, 32, 16), (64, 256, 16), (64, 16, 256), (512, 16, 16), (64, 512, 16), (64, 16, 512)]


def _solver(X, p, prec, unfused=False):
    if unfused:
        os.environ["X3D_UNFUSED_MOMENTUM"] = "1"
    try:
        return X.Solver(p, prec)
    finally:
        os.environ.pop("X3D_UNFUSED_MOMENTUM", None)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_fused_step_matches_oracle(torch_cuda, grid):
    import incompact3d_amd as X
    nx, ny, nz = grid
    kw = dict(nx=nx, ny=ny, nz=nz, re=1600.0, dt=0.005)
    print(f"fused momentum vs oracle {grid}", flush=True)
    p = X.make_params(**kw)
    X.init(p)
    try:
        po = O.Params(**kw)
        S = O.Schemes(po)
        st = O.tgv_state(po)
        for _ in range(6):
            O.step(S, st)
        s = _solver(X, p, 0)
        s.init_tgv()
        s.step(6)
        ux, uy, uz = [t.cpu().numpy() for t in s.velocity()]
        for a, b in ((ux, st.ux), (uy, st.uy), (uz, st.uz)):
            assert rel_err(a, b) <= 1e-11
        ref = O.tgv_diagnostics(S, st.ux, st.uy, st.uz)
        got = s.tgv_stats()
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-11 * abs(r), (got, ref)
        dmax, _ = s.divergence_stats()
        assert dmax < 1e-10
        s.close()
        s32 = _solver(X, p, 1)
        s32.init_tgv()
        s32.step(6)
        got32 = s32.tgv_stats()
        for g, r in zip(got32, ref):
            assert abs(g - r) <= 1e-4 * abs(r), (got32, ref)
        s32.close()
    finally:
        X.finalize()


def test_fused_matches_unfused(torch_cuda):
    """The two device paths compute the same step: 40 steps of TGV 64^3.
    FP64: the paths agree to 1e-12.  FP32 bulk: the paths round differently,
    so each is measured against the FP64 run; the fused path's error stays
    below 1e-4 and within 1.5x of the unfused path's error."""
    import incompact3d_amd as X
    p = X.make_params(64, 64, 64, re=1600.0, dt=0.005)
    X.init(p)
    try:
        vel = {}
        for prec in (0, 1):
            for unfused in (False, True):
                print(f"fused vs unfused momentum, precision {prec}, unfused={unfused}", flush=True)
                s = _solver(X, p, prec, unfused)
                s.init_tgv()
                s.step(40)
                vel[prec, unfused] = [t.cpu().numpy() for t in s.velocity()]
                s.close()
        for a, b in zip(vel[0, False], vel[0, True]):
            assert rel_err(a, b) <= 1e-12
        for c in range(3):
            e_f = rel_err(vel[1, False][c], vel[0, False][c])
            e_u = rel_err(vel[1, True][c], vel[0, False][c])
            assert e_f <= 1e-4 and e_f <= 1.5 * e_u + 1e-7, (c, e_f, e_u)
    finally:
        X.finalize()
