#!/usr/bin/env python3
"""Throughput of the Xcompact3d hot path on one MI355X.

Workload: Taylor-Green vortex, 256^3 triply periodic box [0,2pi)^3, Re=1600,
6th-order compact schemes, spectral Poisson, AB3 with dt=0.002 -- one "step"
is one full time step (momentum RHS, AB3, divergence, Poisson, gradient,
correction).  Prints ONE JSON line; ``value`` is the FP64 path in
Mpoint-steps/s (grid points x steps / s), the FP32-bulk path is reported
beside it with its dissipation error against the FP64 run.

    python bench.py [--steps K] [--warmup W] [--n 256]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def _events_time(stream, fn, reps):
    import torch
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    e.synchronize()
    return s.elapsed_time(e) / reps   # ms per launch


def run_solver(X, p, precision, warmup, steps, stream):
    import torch
    s = X.Solver(p, precision)
    s.init_tgv()
    s.step(warmup)
    torch.cuda.synchronize()
    if torch.distributed.is_available() and torch.distributed.is_initialized():
        torch.distributed.barrier()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    s.step(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if torch.distributed.is_available() and torch.distributed.is_initialized():
        torch.distributed.barrier()
    stats = s.tgv_stats()
    div = s.divergence_stats()
    s.close()
    return {"wall_ms_per_step": (t1 - t0) * 1e3 / steps,
            "event_ms_per_step": ev0.elapsed_time(ev1) / steps,
            "stats": stats, "div": div}


def run_les_step(X, p, warmup, steps, stream):
    """FP32-bulk step with the learned closure (eddy-viscosity form, model 3)
    at the bench grid: nine gradients, the BF16-MFMA network, its stress
    divergence, then the step.  The network is freshly initialised -- the
    cost does not depend on the weights; its trained accuracy is measured by
    tests/test_gpu_les_learned.py."""
    import torch
    from incompact3d_amd import les
    s = X.Solver(p, 1)
    s.init_tgv()
    m = X.Mlp(les.init_params(0))
    delta = (p.xlx * p.yly * p.zlz / (p.nx * p.ny * p.nz)) ** (1.0 / 3.0)
    s.set_les(3, cs=0.17, beta=1.0, mlp=m, gscale=1.0, delta0=delta)
    s.step(warmup)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    s.step(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    s.close()
    m.close()
    return ev0.elapsed_time(ev1) / steps


def roofline_momentum_yz(X, p, stream, reps=10):
    """Roofline of the step's dominant kernel, the fused y/z momentum sweep
    (mom_yz_kernel, FP64): the six y/z sweeps of one step launched `reps`
    times through x3d_solver_momentum_yz on the library stream, timed with
    HIP events on that stream.  Algorithmic bytes per launch: the fields each
    launch must move once (u_d if it differs from u_c, u_c, the accumulator
    when present, the output), N*8 bytes each, averaged over the six launch
    configurations (19 fields per step / 6)."""
    import torch
    s = X.Solver(p, 0)
    s.init_tgv()
    s.momentum_yz(2)
    torch.cuda.synchronize()
    per_call = _events_time(stream, lambda: s.momentum_yz(1), reps)   # ms per six launches
    nbytes = s.momentum_yz(0)
    s.close()
    ms = per_call / 6.0
    achieved = nbytes / (ms * 1e-3) / 1e9
    n = p.nx
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic("mom_yz_kernel<double, 8>", n),
            "kernel": f"x3d::mom_yz_kernel<double, 8> (fused y/z momentum sweep, {n}^3 FP64, mean of the six "
                      f"launch configurations of a step)",
            "avg_ms": round(ms, 5), "algorithmic_bytes": int(nbytes)}


def roofline_line_kernel(X, n, stream, reps=40):
    """One compact solve on its own: a first derivative along y of an n^3
    FP64 field (dery_00, seg_line_kernel), timed with HIP events on the
    library stream.  Algorithmic bytes per launch: 2*N*8 (one read of u, one
    write of du/dy)."""
    import torch
    u = torch.randn((n, n, n), dtype=torch.float64, device="cuda")
    t = torch.empty_like(u)
    for _ in range(3):
        X.line_op_async(1, "der1", 0, u, t)
    torch.cuda.synchronize()
    ms = _events_time(stream, lambda: X.line_op_async(1, "der1", 0, u, t), reps)
    algo = 2.0 * n ** 3 * 8
    achieved = algo / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic("seg_line_kernel<double, 0, 8, false>", n),
            "kernel": f"x3d::seg_line_kernel<double, 0, 8, false> (dery_00, {n}^3 FP64)",
            "avg_ms": round(ms, 5), "algorithmic_bytes": int(algo)}


def _traffic(kernel, n):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size)."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        k = d.get("kernels", {}).get(kernel)
        if k and int(d.get("n", -1)) == n:
            return k.get("traffic_bytes")
    except (OSError, ValueError):
        pass
    return None


def micro_field(shape, walls, dtype, seed):
    """Kernel-microbench input: seeded uniform[-1,1] noise plus 8 sine modes
    of integer wavenumbers |k|_inf <= 4 (amplitudes U[0.5,1], phases
    U[0,2pi)), on the (nz, ny, nx) grid of the [0,2pi)^3 box (y spans
    [0,2pi] closed when y has walls)."""
    import numpy as np
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2.0 - 1.0
    two_pi = 2.0 * math.pi
    x = torch.arange(nx, device="cuda", dtype=torch.float64) * (two_pi / nx)
    y = torch.arange(ny, device="cuda", dtype=torch.float64) * (two_pi / ((ny - 1) if walls else ny))
    z = torch.arange(nz, device="cuda", dtype=torch.float64) * (two_pi / nz)
    for _ in range(8):
        k = rng.integers(-4, 5, size=3)
        while not k.any():
            k = rng.integers(-4, 5, size=3)
        a, ph = rng.uniform(0.5, 1.0), rng.uniform(0.0, two_pi)
        u += a * torch.sin(k[0] * x[None, None, :] + k[1] * y[None, :, None] + k[2] * z[:, None, None] + ph)
    return u.to(dtype)


def microbench(X, n, walls, stream, reps=10):
    """Kernel microbench on an n^3 field (n x (n+1) x n with Dirichlet walls,
    ncl=2, in y): first and second derivatives along x, y, z, one pressure
    Poisson solve on the zero-mean pressure-grid field, one pencil transpose;
    FP64 and FP32.  Times are HIP-event averages over `reps` launches on the
    library stream.  Algorithmic bytes: 2*N*w per derivative and transpose,
    10*N*w per Poisson solve (five FFT passes, divide fused)."""
    import torch
    ny = n + 1 if walls else n
    p = X.make_params(n, ny, n, ncly1=2 if walls else 0, nclyn=2 if walls else 0, re=1600.0, dt=0.002)
    X.init(p)
    X.set_stream(stream.cuda_stream)
    out = []
    try:
        for prec, dt in ((0, torch.float64), (1, torch.float32)):
            w = 8 if prec == 0 else 4
            with torch.cuda.stream(stream):
                u = micro_field((n, ny, n), walls, dt, 20261202)
                t = torch.empty_like(u)
                N = u.numel()
                for d, name in ((0, "x"), (1, "y"), (2, "z")):
                    for op, lab in (("der1", "der"), ("der2", "der" + name)):
                        nm = f"{lab}{name}" if op == "der1" else f"der{name}{name}"
                        X.line_op_async(d, op, 1, u, t)
                        ms = _events_time(stream, lambda: X.line_op_async(d, op, 1, u, t), reps)
                        gbs = 2.0 * N * w / (ms * 1e-3) / 1e9
                        out.append({"op": nm, "dtype": "f64" if prec == 0 else "f32", "ms": round(ms, 4),
                                    "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
                rhs = u[:, : (ny - 1 if walls else ny), :].contiguous()
                rhs -= rhs.mean()
                Np = rhs.numel()
                X.poisson_async(rhs)
                ms = _events_time(stream, lambda: X.poisson_async(rhs), reps)
                gbs = 10.0 * Np * w / (ms * 1e-3) / 1e9
                out.append({"op": "poisson_010" if walls else "poisson_000", "dtype": "f64" if prec == 0 else "f32",
                            "ms": round(ms, 4), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
                X.transpose_async(u, t)
                ms = _events_time(stream, lambda: X.transpose_async(u, t), reps)
                gbs = 2.0 * N * w / (ms * 1e-3) / 1e9
                out.append({"op": "transpose", "dtype": "f64" if prec == 0 else "f32", "ms": round(ms, 4),
                            "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
                del u, t, rhs
                torch.cuda.synchronize()
    finally:
        X.finalize()
    return {"grid": [n, ny, n], "y": "Dirichlet walls (ncl=2)" if walls else "periodic", "ops": out}


def c1_bench(X, stream, steps=200):
    """The 64^3 TGV of the CPU-runnable configuration (Re 1600, AB3, dt 0.005,
    200 steps): HIP-event ms per step over the whole run, FP64 and FP32 bulk
    (each kernel is a few microseconds at this size, so the step is bound by
    the ~20 launches, not by HBM)."""
    p = X.make_params(64, 64, 64, re=1600.0, dt=0.005)
    out = {"workload": "TGV 64^3 Re=1600, AB3 dt=0.005, 200 steps"}
    X.init(p)
    X.set_stream(stream.cuda_stream)
    try:
        for prec, lab in ((0, "fp64"), (1, "fp32_bulk")):
            s = X.Solver(p, prec)
            s.init_tgv()
            ms = _events_time(stream, lambda: s.step(steps), 1) / steps
            s.close()
            out[lab] = {"ms_per_step": round(ms, 4), "mpoint_steps_per_s": round(64 ** 3 / (ms * 1e-3) / 1e6, 1)}
    finally:
        X.finalize()
    return out


def channel_step_bench(X, stream, warmup=20, steps=200):
    """The Re_tau 180 channel step (256 x 129 x 128, stretched y, RK3,
    hyperviscous second derivative, constant flow rate, spin-up rotation on)
    from the laminar profile with 10 % noise: HIP-event ms per step, FP64
    and FP32 bulk.  The full 20,000-step run and its friction Reynolds
    number are a GPU test of their own (tests/test_gpu_channel_c4.py)."""
    from incompact3d_amd import channel
    p = channel.params()
    npts = p.nx * p.ny * p.nz
    out = {"workload": "channel Re_tau 180, 256x129x128, istret 2, RK3 dt 0.005, per-operator step"}
    X.init(p)
    X.set_stream(stream.cuda_stream)
    try:
        for prec, lab in ((0, "fp64"), (1, "fp32_bulk")):
            s = X.Solver(p, prec)
            s.init_channel(0.1, 7)
            s.set_forcing(1, 2.0 / 3.0)
            s.set_rotation(0.12, 5000)
            s.step(warmup)
            ms = _events_time(stream, lambda: s.step(steps), 1) / steps
            s.close()
            out[lab] = {"ms_per_step": round(ms, 4), "mpoint_steps_per_s": round(npts / (ms * 1e-3) / 1e6, 1)}
    finally:
        X.finalize()
    return out


def stretched_poisson_bench(X, stream, reps=10):
    """Pressure Poisson solve on the channel grid 256 x 129 x 128 (walls in y,
    ncl=2), uniform (poisson_010, istret=0) and stretched (istret=2,
    beta=0.259: matrice_refinement's pentadiagonal systems), FP64 and FP32,
    HIP-event averages over `reps` solves."""
    import torch
    out = {"grid": [256, 129, 128], "y": "Dirichlet walls (ncl=2)", "beta": 0.259}
    for istret in (0, 2):
        p = X.make_params(256, 129, 128, xlx=4.0 * math.pi, yly=2.0, zlz=4.0 * math.pi / 3.0, ncly1=2, nclyn=2,
                          istret=istret, beta=0.259)
        X.init(p)
        X.set_stream(stream.cuda_stream)
        try:
            for dt, lab in ((torch.float64, "f64"), (torch.float32, "f32")):
                with torch.cuda.stream(stream):
                    rhs = micro_field((128, 128, 256), True, dt, 20261203)
                    rhs -= rhs.mean()
                    X.poisson_async(rhs)
                    ms = _events_time(stream, lambda: X.poisson_async(rhs), reps)
                    out[f"istret{istret}_{lab}_ms"] = round(ms, 4)
                    del rhs
                    torch.cuda.synchronize()
        finally:
            X.finalize()
    return out


def learned_les_comparison(X, stream, fp64_ms_256, n_dns):
    """Time to solution at equal dissipation error, TGV Re=5000 to t=12:
    the FP64 plain solver on the 256^3 DNS grid against the FP32-bulk solver
    with the learned closure on 64^3.  The closure is trained in this run on
    the box-filtered DNS (incompact3d_amd.les.experiment); the errors are the
    dissipation-peak errors against the filtered DNS (K_bar of the velocity
    box-filtered at every DNS point), against the energy of the block
    averages alone, and against the DNS.  The FP64 step time is this run's
    256^3 measurement (Re changes nothing in the cost)."""
    import time as _t
    from incompact3d_amd import les
    t0 = _t.perf_counter()
    with __import__("torch").cuda.stream(stream):
        out = les.experiment(les_precision=1, time_steps=200)
    wall = _t.perf_counter() - t0
    dt_dns, dt_les, t_end = out["dns"]["dt"], out["les"]["dt"], out["dns"]["t_end"]
    les_ms = out["les"]["ms_per_step_learned"]
    dns_s = fp64_ms_256 * 1e-3 * t_end / dt_dns
    les_s = les_ms * 1e-3 * t_end / dt_les
    r = out["runs"]
    return {
        "workload": f"TGV Re=5000 to t={t_end:g}: FP64 plain {n_dns}^3 (dt {dt_dns}) vs FP32-bulk + learned closure "
                    f"{out['les']['grid']}^3 (dt {dt_les})",
        "fp64_plain_seconds_to_t": round(dns_s, 3),
        "fp32_learned_les_seconds_to_t": round(les_s, 4),
        "fp32_learned_les_ms_per_step": round(les_ms, 4),
        "time_to_solution_speedup": round(dns_s / les_s, 1),
        "peak_eps_dns": out["dns"]["peak_eps_dns"],
        "peak_eps_filtered_dns": out["dns"]["peak_eps_filtered"],
        "peak_eps_block_filtered_dns": out["dns"]["peak_eps_block"],
        "peak_eps_err_vs_dns": {k: round(v["peak_rel_err_vs_dns"], 4) for k, v in r.items()},
        "peak_eps_err_vs_filtered_dns": {k: round(v["peak_rel_err"], 4) for k, v in r.items()},
        "peak_eps_err_vs_block_filtered_dns": {k: round(v["peak_rel_err_vs_block"], 4) for k, v in r.items()},
        "eps_rms_err_vs_filtered_dns": {k: round(v["eps_rms_rel_err"], 4) for k, v in r.items()},
        "training_points": out["training"]["points"],
        "experiment_wall_seconds": round(wall, 1),
    }


def cpu_baseline(n, steps=20):
    """The CPU restatement of the step (oracle/cport: C++ with OpenMP on every
    host thread the process is given, FP64, checked against the numpy oracle
    to 1e-12 by tests/test_cport.py) on a bounded sample of the same
    workload: `steps` TGV time steps at n^3 after one untimed step.  It
    stands in for the reference's CPU build, which needs 2DECOMP&FFT (not in
    the reference tree) and is not built here."""
    import numpy as np
    from oracle import cport_step
    from oracle import x3d_oracle as O
    p = O.Params(n, n, n, re=1600.0, dt=0.002)
    S = O.Schemes(p)
    c = cport_step.CportStep(S)
    u = [np.ascontiguousarray(a) for a in O.init_tgv(p)]
    c.step(*u, nsteps=1)
    t0 = time.perf_counter()
    c.step(*u, nsteps=steps)
    dt = time.perf_counter() - t0
    threads = c.threads()
    c.close()
    return {"value": round(n ** 3 * steps / dt / 1e6, 4), "unit": "Mpoint-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle/cport (C++/OpenMP FP64 restatement of the step, {threads} threads) TGV {n}^3, "
                      f"{steps} AB3 steps after 1 warm-up step, {dt:.1f} s"}


def c3_history(X, stream, n=256, nsteps=10000, every=10):
    """Config 3 to its full length: TGV n^3, Re=1600, AB3 dt=0.002, 10,000
    steps to t=20, FP64 and FP32 bulk, statistics every 10 steps; the FP32
    dissipation error against FP64 over the whole history, and the FP64 run
    against the reference's 512^3 DNS (tests/golden/tgv_reference.json)."""
    import numpy as np
    import torch
    p = X.make_params(n, n, n, re=1600.0, dt=0.002)
    X.init(p)
    X.set_stream(stream.cuda_stream)
    hist, wall = {}, {}
    try:
        with torch.cuda.stream(stream):
            for prec in (0, 1):
                t0 = time.perf_counter()
                s = X.Solver(p, prec)
                s.init_tgv()
                rows = [(0.0,) + tuple(s.tgv_stats())]
                for k in range(nsteps // every):
                    s.step(every)
                    rows.append(((k + 1) * every * 0.002,) + tuple(s.tgv_stats()))
                s.close()
                hist[prec] = np.array(rows)
                wall[prec] = time.perf_counter() - t0
    finally:
        X.finalize()
    h64, h32 = hist[0], hist[1]
    with open(os.path.join(ROOT, "tests", "golden", "tgv_reference.json")) as f:
        dns = np.array(json.load(f)["TGV_Re1600_dns512"]["history"])   # t, E_k, eps_t, eps
    t, ek, eps = h64[:, 0], h64[:, 1], h64[:, 2]
    i, j = int(np.argmax(eps)), int(np.argmax(dns[:, 3]))
    epst = -(ek[2:] - ek[:-2]) / (t[2:] - t[:-2])
    k = int(np.argmax(dns[:, 2]))
    return {
        "workload": f"TGV {n}^3 Re=1600, AB3 dt=0.002, {nsteps} steps to t={nsteps * 0.002:g}, statistics every "
                    f"{every} steps",
        "fp32_vs_fp64_eps_max_rel_err": float(np.max(np.abs(h32[:, 2] - eps) / eps)),
        "fp32_vs_fp64_ke_max_rel_err": float(np.max(np.abs(h32[:, 1] - ek) / ek)),
        "fp64_peak_eps": [round(float(t[i]), 3), float(eps[i])],
        "dns512_peak_eps": [float(dns[j, 0]), float(dns[j, 3])],
        "fp64_peak_eps_rel_err_vs_dns512": float(abs(eps[i] - dns[j, 3]) / dns[j, 3]),
        "fp64_peak_minus_dEk_dt_rel_err_vs_dns512": float(abs(epst.max() - dns[k, 2]) / dns[k, 2]),
        "wall_seconds": {"fp64": round(wall[0], 1), "fp32_bulk": round(wall[1], 1)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the full-length (t=20) FP64 / FP32-bulk accuracy run (about 40 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the channel and 64^3 TGV step timings (a few seconds)")
    ap.add_argument("--no-les", action="store_true",
                    help="skip the learned-LES time-to-solution comparison (TGV Re=5000, about 40 s)")
    ap.add_argument("--micro-n", type=int, default=512,
                    help="kernel microbench grid (0 = skip): n^3 periodic and n x (n+1) x n with y walls")
    ap.add_argument("--only", choices=["fp64", "fp32", "both"], default="both")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline kernel (and a copy of known size) -- for counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU slab decomposition is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import torch
    import incompact3d_amd as X

    n = args.n
    p = X.make_params(n, n, n, re=1600.0, dt=0.002)
    X.init(p)
    stream = torch.cuda.Stream()
    X.set_stream(stream.cuda_stream)
    if args.roofline_only:
        with torch.cuda.stream(stream):
            roof = roofline_momentum_yz(X, p, stream, reps=2)
            line = roofline_line_kernel(X, n, stream, reps=10)
            # calibration: one pencil copy (transpose_x_to_y) of n^3 FP64 words,
            # i.e. n^3*8 bytes read and n^3*8 bytes written
            a = torch.randn((n, n, n), dtype=torch.float64, device="cuda")
            b = torch.empty_like(a)
            X.transpose_x_to_y(a, b)
            torch.cuda.synchronize()
        X.finalize()
        print(json.dumps({"roofline": roof, "line_solve": line, "copy_bytes_each_way": n ** 3 * 8}))
        return
    with torch.cuda.stream(stream):
        r64 = run_solver(X, p, 0, args.warmup, args.steps, stream) if args.only != "fp32" else None
        r32 = run_solver(X, p, 1, args.warmup, args.steps, stream) if args.only != "fp64" else None
        roof = roofline_momentum_yz(X, p, stream)
        line = roofline_line_kernel(X, n, stream)
        les_ms = run_les_step(X, p, args.warmup, args.steps, stream) if args.only == "both" else None
    npts = n ** 3
    main_r = r64 if r64 is not None else r32
    ms = main_r["event_ms_per_step"]
    value = npts / (ms * 1e-3) / 1e6
    out = {
        "metric": "Mpoint-steps/s on TGV, FP64 plain vs FP32+learned LES at equal dissipation err",
        "value": round(value, 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if r64 is not None else "f32",
        "data": "synthetic (analytic Taylor-Green initial condition)",
        "config": {"workload": f"TGV {n}^3 triply periodic, Re=1600, AB3 dt=0.002, 6th-order compact, "
                               f"spectral Poisson", "grid": [n, n, n], "precision": "fp64"},
        "roofline": roof,
        "line_solve_roofline": line,
    }
    if r64 is not None:
        out["fp64"] = {"ms_per_step": round(r64["event_ms_per_step"], 4),
                       "wall_ms_per_step": round(r64["wall_ms_per_step"], 4),
                       "tgv_stats": r64["stats"], "div_max_mean": r64["div"]}
    if r32 is not None:
        v32 = npts / (r32["event_ms_per_step"] * 1e-3) / 1e6
        f32 = {"value": round(v32, 2), "ms_per_step": round(r32["event_ms_per_step"], 4),
               "tgv_stats": r32["stats"], "div_max_mean": r32["div"]}
        if r64 is not None:
            f32["speedup_vs_fp64"] = round(v32 / value, 3)
            e64, e32 = r64["stats"][1], r32["stats"][1]
            f32["eps_rel_err_vs_fp64"] = abs(e32 - e64) / abs(e64)
            f32["ke_rel_err_vs_fp64"] = abs(r32["stats"][0] - r64["stats"][0]) / abs(r64["stats"][0])
        out["fp32_bulk"] = f32
    if les_ms is not None:
        out["fp32_learned_les"] = {"value": round(npts / (les_ms * 1e-3) / 1e6, 2), "ms_per_step": round(les_ms, 4),
                                   "model": "closure network 10-64-64-64-6 on BF16 MFMA, eddy-viscosity form, "
                                            "Smagorinsky fallback (Cs 0.17, beta 1)",
                                   "speedup_vs_fp64": round((npts / (les_ms * 1e-3) / 1e6) / value, 3)}
    out["device"] = X.device_name()
    X.finalize()
    if not args.no_les and r64 is not None:
        out["learned_les_equal_error"] = learned_les_comparison(X, stream, r64["event_ms_per_step"], n)
    if not args.no_c3 and r64 is not None and r32 is not None and n == 256:
        c3 = c3_history(X, stream)
        out["fp32_bulk"]["t20_accuracy"] = c3
    if args.micro_n > 0:
        out["microbench"] = [microbench(X, args.micro_n, False, stream), microbench(X, args.micro_n, True, stream)]
        out["stretched_poisson"] = stretched_poisson_bench(X, stream)
    if not args.no_c4:
        out["channel_c4"] = channel_step_bench(X, stream)
        out["tgv64_c1"] = c1_bench(X, stream)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
        out["cpu_baseline"]["host"] = f"{os.cpu_count()} logical CPUs visible"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
