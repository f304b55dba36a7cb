"""The learned closure end to end (TGV, Re=5000): a 256^3 DNS on the device,
box-filtered by 4 onto 64^3; the closure network trained on the filtered
snapshots with the library's BF16-MFMA kernels (forward, backward, Adam);
64^3 runs without a model, with Smagorinsky (Cs=0.17, les_models.f90 jles=1)
and with the learned closure (its eddy-viscosity form, Smagorinsky fallback
where the network's stress would inject more than the Smagorinsky stress
removes, beta = 1).  eps_t = -dK/dt of each run is compared with the
filtered DNS's -dK_bar/dt and with the DNS's own -dK/dt.

The criterion: the learned LES's dissipation peak within 5 % of the filtered
DNS's, where K_bar is the energy of the DNS velocity box-filtered at every
DNS point (the resolved-scale energy).  The energy of the 4^3 block averages
alone, the same filter sampled at the LES points, is reported beside it: it
carries the aliasing of the filtered field's unresolved tail and its
-dK/dt oscillates by up to 10 % between samples near the peak (the
learned run's peak lies 7-8 % below that curve's highest sample).  Beyond
the criterion the learned closure must beat Smagorinsky on the peak (both
filtered forms and the DNS) and on the rms error of eps_t(t), and beat the
model-free run on the rms error."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu


def test_learned_closure_gains_over_smagorinsky(torch_cuda):
    from incompact3d_amd import les
    out = les.experiment(log=lambda m: print(m, flush=True))
    runs = out["runs"]
    print(json.dumps({k: v for k, v in out.items() if k not in ("history", "params")}, indent=1), flush=True)
    # X3D_RESULTS_DIR, when set to a directory, receives the full record
    dest = os.environ.get("X3D_RESULTS_DIR", "")
    if dest and os.path.isdir(dest):
        with open(os.path.join(dest, "les_experiment.json"), "w") as f:
            json.dump({k: v for k, v in out.items() if k != "params"}, f)
    lr, sm, no = runs["learned"], runs["smagorinsky"], runs["none"]
    assert lr["peak_rel_err"] <= 0.05
    assert lr["peak_rel_err"] < sm["peak_rel_err"]
    assert lr["peak_rel_err_vs_block"] < sm["peak_rel_err_vs_block"]
    assert lr["peak_rel_err_vs_dns"] < sm["peak_rel_err_vs_dns"]
    assert lr["eps_rms_rel_err"] < sm["eps_rms_rel_err"]
    assert lr["eps_rms_rel_err"] < no["eps_rms_rel_err"]
