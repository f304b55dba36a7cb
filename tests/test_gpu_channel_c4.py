"""The turbulent channel at Re_tau 180, 256 x 129 x 128, FP32 bulk, 20,000
RK3 steps of dt 0.005 (to t = 100) from the laminar profile with 10 %
seeded noise and the reference's spin-up rotation (wrotation 0.12 for 5000
steps, examples/Channel/input_DNS_Re180_LR_explicittime.i3d).

* Before transition the wall shear is the laminar one: Re_tau =
  sqrt(2 Re) = 91.65 (U = 1 - y^2, dU/dy = 2 at the walls).
* After it, over the last 4000 steps, the friction Reynolds number and the
  ratio of centreline to bulk velocity are those of the Re_tau 180 channel
  DNS of Moser, Kim & Mansour (Phys. Fluids 11, 1999): Re_tau 178.1 and
  U_c / U_b = 18.20 / 15.73 = 1.157, within 6 % and 5 %.
* At every sample the constant flow rate holds the bulk velocity at 2/3
  (1e-4) and the projected velocity is divergence-free to the FP32-bulk
  roundoff through the finest wall cells (5e-4)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_channel_re180_transition_and_friction(torch_cuda):
    from incompact3d_amd import channel
    hist, ms = channel.run(20000, precision=1, log=lambda m: print(m, flush=True))
    print(f"channel 256x129x128 FP32 bulk: {ms:.3f} ms per RK3 step", flush=True)
    steps = np.array([h[0] for h in hist])
    rt = np.array([h[2] for h in hist])
    uc = np.array([h[3] for h in hist])
    bulk = np.array([h[4] for h in hist])
    div = np.array([h[5] for h in hist])
    early = rt[steps <= 3000]
    assert len(early) >= 2 and np.all(np.abs(early - math.sqrt(2 * 4200.0)) <= 0.05 * math.sqrt(2 * 4200.0))
    late = steps > 16000
    assert abs(rt[late].mean() - 178.12) <= 0.06 * 178.12, rt[late]
    assert abs(uc[late].mean() / (2.0 / 3.0) - 1.157) <= 0.05 * 1.157, uc[late]
    assert np.all(np.abs(bulk - 2.0 / 3.0) <= 1e-4)
    assert np.all(np.abs(div) <= 5e-4)
