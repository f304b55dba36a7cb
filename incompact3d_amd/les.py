"""Learned subgrid-stress closure for the Taylor-Green vortex: data, training
and a-posteriori comparison, all on the device through the library's C-ABI.

1. ``dns_filtered``: a DNS (FP64 solver, no model) from the TGV initial
   condition (Case-TGV.f90:64-96); at the snapshot times its velocity is
   box-filtered by ``f`` on the device (x3d_box_filter), giving the filtered
   velocity and tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j); the kinetic energy
   of the filtered field is recorded on a finer cadence, in two forms: K_bar
   of the velocity box-filtered at every DNS point (the filter is translation
   invariant, so K_bar is the energy of the resolved scales, sum over k of
   |G(k)|^2 E(k)), and the energy of the f^3 block averages alone (the same
   filter sampled at the block centres, i.e. at the LES points).  Sampling
   aliases the filtered field's unresolved tail back onto the coarse grid,
   and the block energy then oscillates as structures move relative to the
   blocks: its -dK/dt jumps by up to 10 % between samples 0.05 apart near
   the peak, where the DNS's own -dK/dt changes by under 4 %.
2. ``training_set``: on the coarse grid (its own context) the nine gradients
   of the filtered velocity come from the same compact derivatives the LES
   uses (parities of smag, les_models.f90:210-238); features are the
   gradients over one scale G (their rms) and delta/delta0; targets are the
   deviatoric T_ij = -dev(tau_ij) over delta^2 G^2.
3. ``train``: Adam on BF16-MFMA forward/backward kernels (x3d_mlp_*); the
   target scale of the training set is folded into the last layer
   afterwards, so the network outputs T_ij / (delta^2 G^2) directly.
4. ``run_les``: the coarse solver with no model, Smagorinsky (jles=1) or the
   learned closure (with its Smagorinsky fallback); kinetic energy recorded
   every ``every`` steps.
5. ``dissipation``: eps_t = -dK/dt by centred differences; ``peak``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GRAD_ORDER = ((0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2), (2, 2))   # (comp, dir)


@dataclass
class FilteredDNS:
    n: int
    f: int
    re: float
    dt: float
    snap_times: list
    snaps: list = field(default_factory=list)      # [(ub[3], tau[6])] FP64 device tensors
    ke_t: list = field(default_factory=list)       # times of ke_filtered
    ke_filtered: list = field(default_factory=list)   # K_bar, filter at every DNS point
    ke_block: list = field(default_factory=list)      # energy of the block averages
    ke_dns: list = field(default_factory=list)


def _init(X, p):
    """x3d_init plus the library stream = torch's current stream, so that the
    torch operations of this module and the library's kernels are ordered."""
    import torch
    X.init(p)
    X.set_stream(torch.cuda.current_stream().cuda_stream)


def _ke(ub):
    return float(sum((c * c).mean().item() for c in ub) * 0.5)


def _ke_box(u, f):
    """Kinetic energy of the velocity box-filtered over f points along each
    axis at every grid point (periodic): the translation-invariant filter
    whose values at i = f*I + f-1 are x3d_box_filter's block averages."""
    import torch
    tot = 0.0
    for c in u:
        v = c
        for ax in range(3):
            acc = v.clone()
            for sh in range(1, f):
                acc += torch.roll(v, sh, dims=ax)
            v = acc / f
        tot += float((v * v).mean().item())
    return 0.5 * tot


def dns_filtered(n=256, f=4, re=5000.0, dt=0.002, t_end=12.0, snap_times=(4, 5, 6, 7, 8, 9, 10, 11, 12),
                 ke_every=25, precision=0, log=None):
    """Run the DNS and box-filter it (see the module docstring)."""
    import incompact3d_amd as X
    p = X.make_params(n, n, n, re=re, dt=dt)
    _init(X, p)
    out = FilteredDNS(n, f, re, dt, list(snap_times))
    try:
        s = X.Solver(p, precision)
        s.init_tgv()
        nsteps = int(round(t_end / dt))
        snap_steps = {int(round(t / dt)): t for t in snap_times}
        for it in range(0, nsteps + 1):
            if it % ke_every == 0 or it in snap_steps:
                u = s.velocity()
                ub, tau = X.box_filter(u[0], u[1], u[2], f)
                if it % ke_every == 0:
                    out.ke_t.append(it * dt)
                    out.ke_filtered.append(_ke_box(u, f))
                    out.ke_block.append(_ke(ub))
                    out.ke_dns.append(_ke(u))
                if it in snap_steps:
                    out.snaps.append((ub, tau))
                    if log:
                        log(f"DNS {n}^3 snapshot t={snap_steps[it]:.2f}")
                del u
            if it < nsteps:
                s.step(1)
        s.close()
    finally:
        X.finalize()
    return out


def gradients(X, ub):
    """The nine gradients g[(comp, dir)] of a coarse velocity (context of
    its grid initialised), FP64 device tensors, order GRAD_ORDER."""
    import torch
    g = []
    for comp, d in GRAD_ORDER:
        t = torch.empty_like(ub[comp])
        X.line_op_async(d, "der1", 0 if comp == d else 1, ub[comp], t)
        g.append(t)
    X.synchronize()
    return g


def training_set(dns: FilteredDNS, log=None):
    """(features, targets, G, delta) on the coarse grid; the coarse context
    must be initialised by the caller."""
    import torch
    import incompact3d_amd as X
    nc = dns.n // dns.f
    delta = (2.0 * math.pi / nc)
    feats, targs = [], []
    for ub, tau in dns.snaps:
        g = gradients(X, ub)
        tr = (tau[0] + tau[1] + tau[2]) / 3.0
        t = [-(tau[0] - tr), -(tau[1] - tr), -(tau[2] - tr), -tau[3], -tau[4], -tau[5]]
        feats.append(torch.stack([c.reshape(-1) for c in g], dim=1))
        targs.append(torch.stack([c.reshape(-1) for c in t], dim=1))
    gx = torch.cat(feats)
    ty = torch.cat(targs)
    G = float(torch.sqrt((gx * gx).mean()).item())
    x = torch.cat([gx / G, torch.ones((gx.shape[0], 1), dtype=gx.dtype, device=gx.device)], dim=1)
    y = ty / (delta * delta * G * G)
    if log:
        log(f"training set: {x.shape[0]} points, G={G:.4g}, target rms={float(y.pow(2).mean().sqrt()):.4g}")
    return x.float().contiguous(), y.float().contiguous(), G, delta


LAYERS = ((10, 64), (64, 64), (64, 64), (64, 6))


def init_params(seed=0):
    """Flat FP32 parameters (W1 b1 W2 b2 W3 b3 W4 b4): weights uniform in
    +-sqrt(6/fan_in) (half that for the output layer), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in LAYERS:
        lim = math.sqrt(6.0 / fan_in) * (0.5 if fan_out == 6 else 1.0)
        parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def train(x, y, epochs=30, batch=8192, lr=2e-3, seed=0, log=None):
    """Train a fresh network on (x, y); returns (Mlp, final batch loss).
    The targets are scaled to unit rms for training and the scale is folded
    into the last layer afterwards."""
    import torch
    import incompact3d_amd as X
    ys = float(y.pow(2).mean().sqrt().item())
    yn = (y / ys).contiguous()
    m = X.Mlp(init_params(seed))
    n = x.shape[0]
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    nb = max(1, n // batch)
    total = epochs * nb
    step = 0
    loss = float("nan")
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(x.device)
        xs, ysh = x[perm].contiguous(), yn[perm].contiguous()
        for b in range(nb):
            # cosine learning-rate decay to 5% of lr
            lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * step / total)))
            loss = m.train_step(xs[b * batch:(b + 1) * batch], ysh[b * batch:(b + 1) * batch], lr_t)
            step += 1
        if log and (ep % 5 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: batch loss {loss:.4g} (normalised targets)")
    flat = m.get_params()
    nw4 = 6 * 64
    off4 = flat.size - nw4 - 6
    flat[off4:] *= ys
    m.set_params(flat)
    return m, loss


def run_les(n, re, dt, t_end, model=0, cs=0.17, mlp=None, G=1.0, delta0=1.0, beta=0.0, iconserv=0, every=5,
            precision=0):
    """Coarse TGV run; returns (times, KE, fallback count).  The coarse
    context of n^3 must be initialised by the caller."""
    import incompact3d_amd as X
    p = X.make_params(n, n, n, re=re, dt=dt)
    s = X.Solver(p, precision)
    s.init_tgv()
    if model:
        s.set_les(model, cs=cs, iconserv=iconserv, beta=beta, mlp=mlp, gscale=G, delta0=delta0)
    nsteps = int(round(t_end / dt))
    ts, ke = [], []
    for it in range(0, nsteps + 1):
        if it % every == 0:
            ts.append(it * dt)
            ke.append(s.tgv_stats()[0])
        if it < nsteps:
            s.step(1)
    fb = s.les_fallbacks() if model >= 2 else 0
    s.close()
    return np.array(ts), np.array(ke), fb


def dissipation(t, k):
    """eps_t = -dK/dt, centred differences (one-sided at the ends)."""
    t = np.asarray(t, dtype=np.float64)
    k = np.asarray(k, dtype=np.float64)
    return t, -np.gradient(k, t)


def peak(t, eps, tmin=2.0):
    sel = t >= tmin
    i = int(np.argmax(eps[sel]))
    return float(t[sel][i]), float(eps[sel][i])


def apriori(mlp, x, y_phys_scaled):
    """Correlation of the network's T_ij with the filtered-DNS T_ij over a
    feature set (both scaled by delta^2 G^2), per component."""
    import torch
    pred = mlp.forward(x)
    out = []
    for c in range(6):
        a = pred[:, c].double()
        b = y_phys_scaled[:, c].double()
        a = a - a.mean()
        b = b - b.mean()
        out.append(float((a * b).sum() / torch.sqrt((a * a).sum() * (b * b).sum())))
    return out


def time_les_steps(n, re, dt, model, mlp=None, G=1.0, delta0=1.0, cs=0.17, beta=1.0, precision=1, steps=200):
    """ms per step of the coarse LES (no diagnostics), HIP events on the
    library stream; the coarse context must be initialised by _init."""
    import torch
    import incompact3d_amd as X
    stream = torch.cuda.current_stream()   # the library stream (see _init)
    p = X.make_params(n, n, n, re=re, dt=dt)
    s = X.Solver(p, precision)
    s.init_tgv()
    if model:
        s.set_les(model, cs=cs, beta=beta, mlp=mlp, gscale=G, delta0=delta0)
    s.step(5)
    X.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    s.step(steps)
    X.synchronize()
    e1.record(stream)
    torch.cuda.synchronize()
    s.close()
    return e0.elapsed_time(e1) / steps


def experiment(n_dns=256, f=4, re=5000.0, dt_dns=0.002, dt_les=0.008, t_end=12.0, snap_times=None, epochs=30,
               batch=8192, lr=2e-3, cs=0.17, beta=1.0, seed=0, log=None, extra_betas=(), les_precision=0,
               time_steps=0):
    """The whole learned-closure experiment: filtered DNS, training, and the
    coarse runs without a model, with Smagorinsky (cs), with the learned
    closure in its eddy-viscosity form (model 3, "learned") and as the full
    stress (model 2, "learned_full"), fallback bound beta (and any
    extra_betas for model 3).  Each run's eps_t = -dK/dt is compared with
    the filtered DNS's -dK_bar/dt (peak and rms error over t >= 2) and with
    the DNS's own -dK/dt (peak)."""
    import incompact3d_amd as X
    if snap_times is None:
        snap_times = tuple(float(t) for t in range(3, int(t_end) + 1))
    dns = dns_filtered(n_dns, f, re, dt_dns, t_end, snap_times, log=log)
    nc = n_dns // f
    _init(X, X.make_params(nc, nc, nc, re=re, dt=dt_les))
    try:
        x, y, G, delta = training_set(dns, log=log)
        mlp, loss = train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
        corr = apriori(mlp, x, y)
        runs = {}
        pr = les_precision
        runs["none"] = run_les(nc, re, dt_les, t_end, model=0, precision=pr)
        runs["smagorinsky"] = run_les(nc, re, dt_les, t_end, model=1, cs=cs, precision=pr)
        runs["learned"] = run_les(nc, re, dt_les, t_end, model=3, cs=cs, mlp=mlp, G=G, delta0=delta, beta=beta,
                                  precision=pr)
        runs["learned_full"] = run_les(nc, re, dt_les, t_end, model=2, cs=cs, mlp=mlp, G=G, delta0=delta,
                                       beta=beta, precision=pr)
        les_ms = (time_les_steps(nc, re, dt_les, 3, mlp=mlp, G=G, delta0=delta, cs=cs, beta=beta, precision=pr,
                                 steps=time_steps) if time_steps else None)
        for b in extra_betas:
            runs[f"learned_b{b:g}"] = run_les(nc, re, dt_les, t_end, model=3, cs=cs, mlp=mlp, G=G, delta0=delta,
                                              beta=b)
        params = mlp.get_params()
        mlp.close()
    finally:
        X.finalize()
    tf, ef = dissipation(dns.ke_t, dns.ke_filtered)
    tp, ep = peak(tf, ef)
    tb, eb = dissipation(dns.ke_t, dns.ke_block)
    tpb, epb = peak(tb, eb)
    td, edns = dissipation(dns.ke_t, dns.ke_dns)
    tpd, epd = peak(td, edns)
    summary = {
        "dns": {"grid": n_dns, "re": re, "dt": dt_dns, "t_end": t_end, "filter": f,
                "peak_eps_filtered": [tp, ep], "peak_eps_block": [tpb, epb], "peak_eps_dns": [tpd, epd]},
        "les": {"grid": nc, "dt": dt_les, "cs": cs, "beta": beta, "precision": "fp32" if les_precision else "fp64",
                "ms_per_step_learned": les_ms},
        "training": {"points": int(x.shape[0]), "G": G, "delta": delta, "final_batch_loss": loss,
                     "apriori_corr": corr, "epochs": epochs},
        "runs": {},
    }
    for name, (t, k, fb) in runs.items():
        tt, ee = dissipation(t, k)
        pt, pv = peak(tt, ee)
        ref = np.interp(tt, tf, ef)
        refb = np.interp(tt, tb, eb)
        sel = tt >= 2.0
        summary["runs"][name] = {
            "peak_eps": [pt, pv], "peak_rel_err": abs(pv - ep) / ep, "peak_rel_err_vs_block": abs(pv - epb) / epb,
            "peak_rel_err_vs_dns": abs(pv - epd) / epd,
            "eps_rms_rel_err": float(np.sqrt(np.mean((ee[sel] - ref[sel]) ** 2)) / ep),
            "eps_rms_rel_err_vs_block": float(np.sqrt(np.mean((ee[sel] - refb[sel]) ** 2)) / epb),
            "ke_end": float(k[-1]), "fallback_points": int(fb),
        }
    summary["history"] = {"t_filtered": list(map(float, tf)), "eps_filtered": list(map(float, ef)),
                          "eps_block": list(map(float, eb)),
                          "t_dns": list(map(float, td)), "eps_dns": list(map(float, edns))}
    for name, (t, k, _) in runs.items():
        tt, ee = dissipation(t, k)
        summary["history"][f"t_{name}"] = list(map(float, tt))
        summary["history"][f"eps_{name}"] = list(map(float, ee))
    summary["params"] = params
    return summary
