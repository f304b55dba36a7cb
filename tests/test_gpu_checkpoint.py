"""Checkpoint / restart (restart() of tools.f90:205-514) through the C ABI.

* A run of 6 steps equals, bit for bit, 3 steps + x3d_solver_checkpoint +
  x3d_solver_restart into a fresh solver + 3 steps: the periodic TGV on the
  fused AB3 path and the channel (walls in y, stretched, RK3 and AB3, the
  wall planes of the pressure gradient rebuilt from pp), FP64 and FP32 bulk.
* The files are what the reference's decomp_2d_write_one calls write: one
  raw FP64 global array per variable in Fortran order (ux equals the
  velocity, pp has the pressure grid's size, dux-2/-3 only for AB3), and
  restart.info carries the &Time and &NumParam namelists in the reference's
  formats (tools.f90:343-373).
* Errors: a checkpoint before any step, a missing checkpoint directory and a
  file of the wrong size are refused."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")

CASES = {
    "tgv64-ab3": dict(nx=64, ny=64, nz=64, re=1600.0, dt=0.005),
    "channel-rk3-s2": dict(nx=16, ny=33, nz=16, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2,
                           istret=2, beta=0.259, re=500.0, dt=0.002, itimescheme=5, isecondder=5),
    "channel-ab3-s1": dict(nx=32, ny=17, nz=16, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2,
                           istret=1, beta=0.259, re=500.0, dt=0.002, itimescheme=3, isecondder=5),
}


def _solver(kw, precision):
    s = X.Solver(X.make_params(**kw), precision)
    if kw.get("ncly1", 0) == 2:
        s.init_channel(0.125, 5)
        if kw.get("itimescheme", 3) == 5:
            s.set_forcing(1, 2.0 / 3.0)
        else:
            s.set_forcing(2, 2.0 / kw["re"])
        s.set_rotation(0.12, 5)
    else:
        s.init_tgv()
    return s


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", list(CASES))
def test_restart_continues_bit_for_bit(case, precision, tmp_path):
    kw = CASES[case]
    p = X.make_params(**kw)
    X.init(p)
    try:
        a = _solver(kw, precision)
        a.step(6)
        ref = [t.cpu().numpy() for t in a.velocity()]
        a.close()
        b = _solver(kw, precision)
        b.step(3)
        mid = [t.cpu().numpy() for t in b.velocity()]
        b.checkpoint(str(tmp_path), t=3 * kw["dt"])
        b.close()
        c = _solver(kw, precision)
        c.step(1)                      # state to be overwritten by the restart
        t = c.restart(str(tmp_path))
        assert c.itime == 3 and abs(t - round(3 * kw["dt"], 4)) < 1e-12
        c.step(3)
        got = [t.cpu().numpy() for t in c.velocity()]
        c.close()
    finally:
        X.finalize()
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    # the files: raw FP64 global arrays in Fortran order
    ck = tmp_path / "checkpoint"
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    for name, m in zip(("ux", "uy", "uz"), mid):
        assert np.array_equal(np.fromfile(ck / name, dtype=np.float64).reshape(nz, ny, nx), m)
    walls = kw.get("ncly1", 0) == 2
    npp = nx * (ny - 1 if walls else ny) * nz
    assert os.path.getsize(ck / "pp") == 8 * npp
    ab3 = kw.get("itimescheme", 3) == 3
    for v in ("dux", "duy", "duz"):
        for lag in (2, 3):
            assert (ck / f"{v}-{lag}").exists() == ab3
    info = (tmp_path / "restart.info").read_text().splitlines()
    assert info[1] == "&Time" and info[5] == "/End" and info[7] == "&NumParam"
    tline = [ln for ln in info if ln.startswith("tfield=")][0]
    assert len(tline) == 10 + 16 and float(tline[10:]) == round(3 * kw["dt"], 4)
    iline = [ln for ln in info if ln.startswith("itime=")][0]
    assert len(iline) == 10 + 16 and int(iline[10:]) == 3
    assert [ln for ln in info if ln.startswith("itimescheme=")][0] == f"itimescheme={kw.get('itimescheme', 3):14d}"
    assert [ln for ln in info if ln.startswith("beta=")][0] == f"beta=     {kw.get('beta', 0.259):16.12f}"


def test_checkpoint_errors(tmp_path):
    kw = CASES["tgv64-ab3"]
    X.init(X.make_params(**kw))
    try:
        s = _solver(kw, 0)
        with pytest.raises(X.X3DError):
            s.checkpoint(str(tmp_path))          # no step yet
        with pytest.raises(X.X3DError):
            s.restart(str(tmp_path / "absent"))
        s.step(1)
        s.checkpoint(str(tmp_path))
        with open(tmp_path / "checkpoint" / "uy", "r+b") as f:
            f.truncate(8 * 100)
        with pytest.raises(X.X3DError):
            s.restart(str(tmp_path))
        s.close()
    finally:
        X.finalize()
