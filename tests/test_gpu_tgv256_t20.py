"""The Taylor-Green vortex at 256^3 to its full length: Re = 1600, AB3,
dt = 0.002, 10,000 steps to t = 20, in FP64 and in the FP32-bulk form, the
statistics of postprocess_tgv (Case-TGV.f90:246-377) every 10 steps.

* FP32 bulk against FP64: the dissipation eps(t) within 1 % at every
  sampled time (the precision target), and E_k(t) within 1e-4; the
  divergence left by the FP32 projection stays at FP32 rounding.
* FP64 against the reference's 512^3 DNS history
  (examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat, whole history in
  tests/golden/tgv_reference.json): the peak of eps within 2 % and at the
  same time within 0.1, -dE_k/dt from the run's own E_k(t) within 2 % of the
  DNS eps_t peak, and E_k within 2 % over t in [0, 20].  2 % is the
  resolution allowance of 256^3 against 512^3 for this flow (the 6th-order
  compact schemes at 256^3 resolve the dissipation range to k eta ~ 1).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

X = pytest.importorskip("incompact3d_amd")

N, DT, NSTEPS, EVERY = 256, 0.002, 10000, 10
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tgv_reference.json")


def _history(precision):
    p = X.make_params(N, N, N, re=1600.0, dt=DT)
    s = X.Solver(p, precision)
    s.init_tgv()
    rows = [(0.0,) + tuple(s.tgv_stats())]
    for k in range(NSTEPS // EVERY):
        s.step(EVERY)
        rows.append(((k + 1) * EVERY * DT,) + tuple(s.tgv_stats()))
        if (k + 1) % 250 == 0:
            print(f"precision {precision}: t = {rows[-1][0]:.2f}, eps = {rows[-1][2]:.6e}", flush=True)
    div = s.divergence_stats()
    s.close()
    return np.array(rows), div


def test_tgv256_to_t20():
    X.init(X.make_params(N, N, N, re=1600.0, dt=DT))
    try:
        h64, d64 = _history(0)
        h32, d32 = _history(1)
    finally:
        X.finalize()
    assert np.isfinite(h64).all() and np.isfinite(h32).all()
    t, ek64, eps64 = h64[:, 0], h64[:, 1], h64[:, 2]
    ek32, eps32 = h32[:, 1], h32[:, 2]
    err_eps = np.abs(eps32 - eps64) / eps64
    err_ek = np.abs(ek32 - ek64) / ek64
    with open(GOLDEN) as f:
        dns = np.array(json.load(f)["TGV_Re1600_dns512"]["history"])   # t, E_k, eps_t, eps
    i64 = int(np.argmax(eps64))
    j = int(np.argmax(dns[:, 3]))
    peak_err = abs(eps64[i64] - dns[j, 3]) / dns[j, 3]
    # -dE_k/dt by centred differences of the sampled E_k (every 0.02)
    epst = -(ek64[2:] - ek64[:-2]) / (t[2:] - t[:-2])
    k = int(np.argmax(dns[:, 2]))
    epst_err = abs(epst.max() - dns[k, 2]) / dns[k, 2]
    ek_dns = np.interp(t, dns[:, 0], dns[:, 1])
    ek_err = np.max(np.abs(ek64 - ek_dns) / ek_dns)
    print(f"FP32 vs FP64: max eps rel err {err_eps.max():.3e}, max E_k rel err {err_ek.max():.3e}")
    print(f"div max/mean: FP64 {d64[0]:.3e}/{d64[1]:.3e}, FP32 bulk {d32[0]:.3e}/{d32[1]:.3e}")
    print(f"FP64 peak eps {eps64[i64]:.7f} at t={t[i64]:.2f}; DNS {dns[j, 3]:.7f} at t={dns[j, 0]:.2f}; "
          f"rel err {peak_err:.3e}; -dE/dt peak rel err {epst_err:.3e}; max E_k rel err vs DNS {ek_err:.3e}")
    assert err_eps.max() <= 0.01
    assert err_ek.max() <= 1e-4
    # FP64 divergence of the FP64 velocity after the last projection.  In
    # the FP32-bulk form it carries the FP32 rounding of the pressure
    # gradient (~6e-8 of it per point) differentiated at the grid scale
    # (~n/2): ~1e-5 on average and ~2e-4 at the maximum over 16.7 M points.
    assert d64[0] < 1e-9 and d32[0] < 1e-3 and d32[1] < 1e-4
    assert peak_err <= 0.02
    assert abs(t[i64] - dns[j, 0]) <= 0.1 + 1e-9
    assert epst_err <= 0.02
    assert ek_err <= 0.02
