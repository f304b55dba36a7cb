"""Config 1 of the hot-path survey (TGV 64^3, Re=1600, AB3, dt=0.005) run on
the GPU to t=1 in FP64 and FP32 bulk, against the reference's published
512^3 DNS statistics (examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat,
extracted to tests/golden/tgv_reference.json)."""
import pytest

pytestmark = pytest.mark.gpu


def _dns_row(t):
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tgv_reference.json")
    rows = json.load(open(path))["TGV_Re1600_dns512"]["rows"]
    return [r for r in rows if r[0] == t][0]


def test_tgv64_config1_against_reference_dns(torch_cuda):
    """Config 1 (64^3, Re=1600, AB3, dt=0.005) run on the GPU to t=1, FP64 and
    FP32 bulk, against the reference's 512^3 DNS (TGV_Re1600.dat)."""
    import incompact3d_amd as X
    p = X.make_params(64, 64, 64, re=1600.0, dt=0.005)
    X.init(p)
    try:
        row = _dns_row(1.0)
        out = {}
        for prec in (0, 1):
            s = X.Solver(p, prec)
            s.init_tgv()
            s.step(200)
            out[prec] = s.tgv_stats()
            s.close()
        for prec in (0, 1):
            eek, eps = out[prec][0], out[prec][1]
            assert abs(eek - row[1]) / row[1] < 1e-5, (prec, eek, row)
            assert abs(eps - row[3]) / row[3] < 1e-5, (prec, eps, row)
        # FP32 bulk against the FP64 run: dissipation within 1e-5 relative
        assert abs(out[1][1] - out[0][1]) / out[0][1] < 1e-5
    finally:
        X.finalize()
