"""The turbulent channel at Re_tau 180 (examples/Channel/
input_DNS_Re180_LR_explicittime.i3d with the 256 x 129 x 128 grid of the
Re_tau 180 configuration): Lx = 4 pi, Ly = 2 with both walls refined
(istret = 2, beta = 0.259), Lz = 4 pi / 3, Re = 4200 on the laminar
centreline velocity, RK3, dt = 0.005, the hyperviscous second derivative
(nu0/nu = 4, c_nu = 0.44), constant flow rate (bulk 2/3), the laminar
profile plus seeded noise (Case-Channel.f90:96-128) and the spin-up
rotation 0.12 for the first 5000 steps (Case-Channel.f90:413-418).

``friction_reynolds`` takes the wall shear from the plane-averaged
streamwise velocity with the second-order one-sided difference on the
stretched mesh (the three points next to each wall), averaged over both
walls: Re_tau = u_tau h / nu with h = 1 and u_tau = sqrt(nu |dU/dy|)."""
from __future__ import annotations

import math
import time

import numpy as np


def params(**over):
    import incompact3d_amd as X
    kw = dict(nx=256, ny=129, nz=128, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3, ncly1=2, nclyn=2,
              istret=2, beta=0.259, re=4200.0, dt=0.005, itimescheme=5, isecondder=5, nu0nu=4.0, cnu=0.44)
    kw.update(over)
    return X.make_params(**kw)


def _wall_slope(y0, y1, y2, u0, u1, u2):
    h1, h2 = y1 - y0, y2 - y0
    return (u1 * h2 * h2 - u2 * h1 * h1 - u0 * (h2 * h2 - h1 * h1)) / (h1 * h2 * (h2 - h1))


def friction_reynolds(ux, yp, re):
    """(Re_tau, plane-averaged U(y)) of a streamwise velocity (nz, ny, nx)
    device tensor; yp the y coordinates of the velocity points."""
    U = ux.double().mean(dim=(0, 2)).cpu().numpy()
    g0 = _wall_slope(yp[0], yp[1], yp[2], U[0], U[1], U[2])
    g1 = -_wall_slope(yp[-1], yp[-2], yp[-3], U[-1], U[-2], U[-3])
    nu = 1.0 / re
    return math.sqrt(abs(nu * 0.5 * (g0 + g1))) / nu, U


def run(steps=20000, precision=1, init_noise=0.1, seed=7, wrotation=0.12, spinup=5000, every=1000, log=None):
    """Run the channel from the laminar profile; returns the history
    [(step, t, Re_tau, U_centre, bulk, max div)] and ms per step (HIP events
    around the stepping, statistics outside)."""
    import torch
    import incompact3d_amd as X
    p = params()
    X.init(p)
    try:
        stream = torch.cuda.current_stream()
        X.set_stream(stream.cuda_stream)
        yp = X.coef.yp
        s = X.Solver(p, precision)
        s.init_channel(init_noise, seed)
        s.set_forcing(1, 2.0 / 3.0)
        s.set_rotation(wrotation, spinup)
        hist = []
        ms_total = 0.0
        done = 0
        t0 = time.time()
        while done < steps:
            k = min(every, steps - done)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            s.step(k)
            e1.record(stream)
            torch.cuda.synchronize()
            ms_total += e0.elapsed_time(e1)
            done += k
            rt, U = friction_reynolds(s.velocity()[0], yp, p.re)
            dmax, _ = s.divergence_stats()
            hist.append((done, done * p.dt, rt, float(U[p.ny // 2]), s.bulk_velocity(), dmax))
            if log:
                log(f"channel step {done} t={done * p.dt:.1f} Re_tau={rt:.1f} U_c={U[p.ny // 2]:.4f} "
                    f"bulk={hist[-1][4]:.6f} div={dmax:.2e} ({time.time() - t0:.0f} s)")
        s.close()
    finally:
        X.finalize()
    return hist, ms_total / steps
