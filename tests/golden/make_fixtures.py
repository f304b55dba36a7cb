"""Extract the numeric fixtures the tests use from the reference's data files.

Run in a container that has the reference checkout (reads text data only):

    python tests/golden/make_fixtures.py /root/reference

Writes tests/golden/tgv_reference.json with
  * rows of examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat (Incompact3d DNS
    at 512^3): t, E_k, eps_t, eps, enstrophy at selected times, the peak
    of eps_t, and the whole t, E_k, eps_t, eps history (t = 0..20);
  * the full tests/TGV-Taylor-Green-vortex/reference_time_evol.dat (CTest
    fixture: 65^3 free-slip RK3 run, columns t, E_k, eps, eps2, enstrophy)
    and its single-precision twin;
  * the t, E_k, eps_t columns of examples/TGV-Taylor-Green-vortex/
    TGV_Re5000.dat (Incompact3d DNS at 1280^3) for t <= 12 and its eps_t peak.
"""
import json
import os
import sys


def read_table(path):
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            rows.append([float(x) for x in s.split()])
    return rows


def main(ref):
    dns = read_table(os.path.join(ref, "examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat"))
    want = [0.0, 0.5, 1.0, 2.0, 5.0, 10.0, 20.0]
    sel = [r[:5] for r in dns if any(abs(r[0] - t) < 1e-9 for t in want)]
    peak = max(dns, key=lambda r: r[2])
    ctest = read_table(os.path.join(ref, "tests/TGV-Taylor-Green-vortex/reference_time_evol.dat"))
    ctest1 = read_table(os.path.join(ref, "tests/TGV-Taylor-Green-vortex/reference_time_evol_single.dat"))
    re5k = read_table(os.path.join(ref, "examples/TGV-Taylor-Green-vortex/TGV_Re5000.dat"))
    peak5k = max(re5k, key=lambda r: r[2])
    out = {
        "TGV_Re5000_dns1280": {"columns": ["t", "E_k", "eps_t"], "rows": [r[:3] for r in re5k if r[0] <= 12.0 + 1e-9],
                               "peak_eps_t": peak5k[:3],
                               "source": "examples/TGV-Taylor-Green-vortex/TGV_Re5000.dat"},
        "TGV_Re1600_dns512": {"columns": ["t", "E_k", "eps_t", "eps", "enstrophy"], "rows": sel,
                              "peak_eps_t": peak[:5],
                              "history": [r[:4] for r in dns],
                              "history_columns": ["t", "E_k", "eps_t", "eps"],
                              "source": "examples/TGV-Taylor-Green-vortex/TGV_Re1600.dat"},
        "ctest_tgv_freeslip_65": {"columns": ["t", "E_k", "eps", "eps2", "enstrophy"], "rows": ctest,
                                  "rows_single": ctest1,
                                  "source": "tests/TGV-Taylor-Green-vortex/reference_time_evol.dat",
                                  "setup": "65^3, ncl=1 on all faces, L=pi, Re=1600, RK3, dt=0.005"},
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tgv_reference.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, len(sel), "DNS rows,", len(ctest), "CTest rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
