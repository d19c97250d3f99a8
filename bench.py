#!/usr/bin/env python3
"""One-GPU throughput of the periodic Brinkman Ewald evaluator.

Workload (single system): N = 1e5 regularized point forces uniform in the unit
box with N(0,1) components, M = 1e5 separate uniform targets, alpha = 10,
eps = 1e-3, 128^3 grid, P = 12, target accuracy 1e-6 relative L2.  Inputs come
from the library's deterministic generator (seed 0x5EED0001) and are resident
in HBM before the timed region.  A step is one full velocity evaluation
(binning, k space, real space) of all targets; the headline is the mixed
FP32/FP64 form, the plain FP64 form is timed beside it.

Accuracy: both forms are compared with an FP64 reference evaluated on the GPU
with tighter parameters (a different split d, cutoff 5.6 d, P = 20), itself
checked by split invariance and, unless --no-cpu, against the FP64 oracle on a
subsample of targets.

Batched leg (reported beside the headline, in "batched_form"): 512
independent periodic systems of uneven size (N log-uniform in [100, 20000],
box side U[0.5, 2), alpha log-uniform in [1, 50], eps = 1e-3 L, targets =
sources, generator seed 0x5EED0003), every system autotuned, in one call that
replays a captured graph; the mixed form against the FP64 (table) and plain
FP64 (closed forms) forms of the same systems, errors against an FP64
evaluation of every system at tuner tolerance 1e-11.

Clustered leg ("clustered_form"): 1e6 sources in 64 Gaussian clusters (sigma
0.02), targets = sources, alpha 25, eps 5e-4, 256^3, P 16 (seed 0x5EED0002).
Multi-RHS leg ("multirhs_form"): 2e5 uniform particles with 64 FP32 force
sets, targets = sources, 128^3, window P from the autotuner (10) (seed
0x5EED0004).

Prints ONE JSON line (rank 0, N = 1 only).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NS = NT = 100_000
ALPHA, EPS, GRID, P = 10.0, 1e-3, (128, 128, 128), 12
# The accuracy requirement is rel. L2 <= 1e-6 against the FP64 oracle.  The
# tuner's error model (pb_tune: rc = d sqrt(ln(1/tol) + 1), d from the grid's
# Nyquist and window) is conservative here: measured (mixed form) tol 2e-6
# gives 4.0e-7, 3e-6 5.2e-7, 4e-6 6.4e-7, 6e-6 8.9e-7, 1e-5 1.4e-6.  The
# benchmark takes tol 4e-6; the measured errors are reported and checked.
ACCURACY, TOL = 1e-6, 4e-6
# The batched and clustered legs: their windows P come from the tuner, and
# tol < e^{-pi 8 / 2} = 3.5e-6 keeps P >= 10 (P = 8 alone gives ~2e-6)
BATCH_TOL = 3e-6
MRHS_TOL = 2e-6  # multi-RHS: the autotuner's window and split at the configuration's 128^3 grid


def multirhs_params(pb, n):
    """Parameters of the multi-RHS workload: the autotuner's choice for n
    particles at MRHS_TOL (it picks the configuration's 128^3 grid, with P 10);
    should it pick another grid, 128^3 with its window rule instead."""
    q = pb.Params(alpha=10.0, eps=1e-3, tol=MRHS_TOL, precision="mixed")
    q.autotune(n, n)
    if tuple(q.grid) != (128, 128, 128):
        q = pb.Params(alpha=10.0, eps=1e-3, grid=(128, 128, 128), P=10, tol=MRHS_TOL, precision="mixed").tune(n, n)
    return q
CLUSTER_TOL = 3e-6
# Algorithmic work of the pair kernel: 200 FLOP per (target, source image)
# inside rc -- about 80 arithmetic operations (geometry, radial functions,
# 3x3 apply) plus six exp/erfc at 20 FLOP each.
FLOP_PER_PAIR = 200.0
PEAK_FP64, PEAK_FP32 = 78.6, 157.3  # TFLOP/s, MI355X FP64 (vector = matrix) and FP32 peaks
COPY_BYTES = 1 << 29  # copy-calibration size (512 MiB read + 512 MiB written)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def timed_steps(ctx, p, xs, fs, ys, u, steps, warmup):
    """ms per step over `steps` timed steps with no events in the stream; the
    pair kernel's ms per launch from HIP events around it in as many steps
    again (two events per step, on the pair kernel's stream: each event costs
    the stream about 6 us, so these steps are not the ones timed for the
    throughput; the pair kernel runs concurrently with the k-space pipeline in
    both); and the phase breakdown from a few further steps with every phase
    boundary recorded (phases run one after another then: "pair_alone" is the
    pair kernel without the concurrent k space)."""
    import torch
    for _ in range(warmup):
        ctx.velocity(p, xs, fs, ys, u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.velocity(p, xs, fs, ys, u)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing("pair")
    t2 = time.perf_counter()
    for _ in range(steps):
        ctx.velocity(p, xs, fs, ys, u)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    pair_ms = ctx.last_timings()["pair"]
    ctx.set_timing(True)
    for _ in range(min(steps, 5)):
        ctx.velocity(p, xs, fs, ys, u)
    phases = ctx.last_timings()
    ctx.set_timing(False)
    phases["pair_alone"] = phases["pair"]
    phases["pair"] = pair_ms
    phases["step_with_pair_events"] = (t3 - t2) * 1e3 / steps
    return (t1 - t0) * 1e3 / steps, phases


def batched_leg(ctx, steps, warmup):
    """Time the batched systems in the mixed, FP64 and plain FP64 forms."""
    import torch

    import periodicbrinkman_amd as pb

    systems, x, f, off = pb.batched_systems(pb.config_seed(3))
    X = torch.as_tensor(x, device="cuda:0")
    F = torch.as_tensor(f, device="cuda:0")
    U = torch.empty_like(X)

    def params(precision, tol, kernel="table"):
        return [pb.Params(L=(s["L"],) * 3, alpha=s["alpha"], eps=s["eps"], precision=precision, tol=tol,
                          kernel=kernel) for s in systems]

    def run(ps):
        for _ in range(max(2, warmup)):  # the first call runs eagerly, the second captures the graph
            ctx.velocity_batched(ps, X, F, off, X, off, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity_batched(ps, X, F, off, X, off, U)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, U.cpu().numpy().copy()

    pm = params("mixed", BATCH_TOL)
    ms_m, u_m = run(pm)
    log(f"batched mixed: {ms_m:.3f} ms/call")
    ms_64, u_64 = run(params("fp64", BATCH_TOL))
    log(f"batched fp64: {ms_64:.3f} ms/call")
    ms_pl, u_pl = run(params("fp64", BATCH_TOL, "direct"))
    log(f"batched plain fp64: {ms_pl:.3f} ms/call")
    pr = params("fp64", 1e-11)
    u_ref = ctx.velocity_batched(pr, X, F, off, X, off).cpu().numpy()
    per = [relL2(u_m[off[i]:off[i + 1]], u_ref[off[i]:off[i + 1]]) for i in range(len(systems))]
    nt = int(off[-1])
    grids = np.array([q.grid[0] for q in pm])
    return {
        "workload": "512 independent periodic systems, N log-uniform in [100, 20000] (sum %d), box side "
                    "U[0.5, 2), alpha log-uniform in [1, 50], eps 1e-3 L, targets = sources; grids and "
                    "split autotuned per system (tuner tol %g)" % (nt, BATCH_TOL),
        "value": nt / (ms_m / 1e3), "unit": "velocity evals/s", "ms_per_call": ms_m, "calls": steps,
        "rel_l2_mixed": relL2(u_m, u_ref), "max_system_rel_l2_mixed": float(max(per)),
        "meets_accuracy": bool(relL2(u_m, u_ref) <= ACCURACY),
        "rel_l2_fp64": relL2(u_64, u_ref), "rel_l2_plain_fp64": relL2(u_pl, u_ref),
        "fp64_form": {"value": nt / (ms_64 / 1e3), "ms_per_call": ms_64},
        "plain_fp64_form": {"value": nt / (ms_pl / 1e3), "ms_per_call": ms_pl},
        "speedup_mixed_vs_fp64": ms_64 / ms_m, "speedup_mixed_vs_plain_fp64": ms_pl / ms_m,
        "grids_mixed": {str(int(g)): int((grids == g).sum()) for g in np.unique(grids)},
        "note": "each call replays the graph captured from the second identical call: one launch per "
                "phase and kernel class for all systems; wall time per call over the timed calls",
    }


def clustered_leg(ctx, steps):
    """The clustered large system: 1e6 sources in 64 Gaussian clusters, targets = sources."""
    import torch

    import periodicbrinkman_amd as pb

    n = 1_000_000
    x, f = pb.clustered_system(pb.config_seed(2), n, 64, 0.02)
    X = torch.as_tensor(x, device="cuda:0")
    F = torch.as_tensor(f, device="cuda:0")
    U = torch.empty_like(X)
    base = pb.Params(alpha=25.0, eps=5e-4, grid=(256, 256, 256), P=16, tol=CLUSTER_TOL)

    def run(p):
        ctx.velocity(p, X, F, X, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity(p, X, F, X, U)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, U.cpu().numpy().copy()

    pm = base.copy(precision="mixed").tune(n, n)
    ms_m, u_m = run(pm)
    stats = ctx.last_pair_stats()
    log(f"clustered mixed: {ms_m:.2f} ms/step, pairs {stats[0] + stats[1]}")
    ms_64, u_64 = run(base.copy(precision="fp64").tune(n, n))
    log(f"clustered fp64: {ms_64:.2f} ms/step")
    ms_pl, u_pl = run(base.copy(precision="fp64", kernel="direct").tune(n, n))
    log(f"clustered plain fp64: {ms_pl:.2f} ms/step")
    pr = pb.Params(alpha=25.0, eps=5e-4, grid=(256, 256, 256), P=16, tol=1e-10, precision="fp64").tune(n, n)
    u_ref = ctx.velocity(pr, X, F, X).cpu().numpy()
    return {
        "workload": "1e6 sources in 64 Gaussian clusters (sigma 0.02, centres uniform, wrapped), targets = "
                    "sources, alpha 25, eps 5e-4, 256^3 grid, P 16, tuner tol %g" % CLUSTER_TOL,
        "value": n / (ms_m / 1e3), "unit": "velocity evals/s", "ms_per_step": ms_m, "steps": steps,
        "d": pm.d, "rc": pm.rc, "pairs": stats[0] + stats[1],
        "rel_l2_mixed": relL2(u_m, u_ref), "meets_accuracy": bool(relL2(u_m, u_ref) <= ACCURACY),
        "rel_l2_fp64": relL2(u_64, u_ref),
        "fp64_form": {"value": n / (ms_64 / 1e3), "ms_per_step": ms_64},
        "plain_fp64_form": {"value": n / (ms_pl / 1e3), "ms_per_step": ms_pl},
        "speedup_mixed_vs_plain_fp64": ms_pl / ms_m,
        "note": "errors against an FP64 evaluation at tuner tol 1e-10 (a smaller split d)",
    }


def multirhs_leg(ctx, steps):
    """Many right-hand sides: 2e5 particles, 64 FP32 force sets, targets = sources."""
    import torch

    import periodicbrinkman_amd as pb

    n, R = 200_000, 64
    x, _, _ = pb.uniform_system(pb.config_seed(4), n, 0)
    r = pb.Rng(pb.config_seed(4) + (1 << 32))
    Fh = r.normal(3 * n * R).reshape(R, n, 3).astype(np.float32)
    X = torch.as_tensor(x, device="cuda:0")
    F = torch.as_tensor(Fh, device="cuda:0")
    p = multirhs_params(pb, n)
    U = ctx.velocity_multirhs(p, X, F, X)
    U = ctx.velocity_multirhs(p, X, F, X, U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.velocity_multirhs(p, X, F, X, U)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"multi-RHS: {ms:.2f} ms/call")
    u = U.cpu().numpy()
    # the same sets one call each (mixed single-set form), and FP64 references for three sets
    F64 = [torch.as_tensor(Fh[k].astype(np.float64), device="cuda:0") for k in range(R)]
    V = torch.empty_like(X)
    ctx.velocity(p, X, F64[0], X, V)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(R):
        ctx.velocity(p, X, F64[k], X, V)
    torch.cuda.synchronize()
    ms_single = (time.perf_counter() - t0) * 1e3
    pairs = sum(ctx.last_pair_stats()[:2])  # (pair-images inside rc: the same for every force set)
    pr = pb.Params(alpha=10.0, eps=1e-3, grid=(128, 128, 128), P=20, tol=1e-11, precision="fp64").tune(n, n)
    errs = [relL2(u[k], ctx.velocity(pr, X, F64[k], X).cpu().numpy()) for k in (0, R // 2, R - 1)]
    return {
        "workload": "2e5 uniform particles, 64 force sets (FP32, N(0,1)), targets = sources, alpha 10, "
                    "eps 1e-3, %d^3 grid, P %d (autotuned), tuner tol %g" % (p.grid[0], p.P, MRHS_TOL),
        "value": n * R / (ms / 1e3), "unit": "velocity evals/s (targets x force sets)", "ms_per_call": ms,
        "calls": steps, "pairs": pairs, "P": p.P, "d": p.d, "rc": p.rc, "rel_l2_sets_0_32_63": errs, "meets_accuracy": bool(max(errs) <= ACCURACY),
        "single_set_calls_ms": ms_single, "speedup_vs_single_set_calls": ms_single / ms,
        "note": "real space as the MFMA contraction of FP32 kernel tiles with the 64 sets; errors of three "
                "sets against FP64 (P 20, tuner tol 1e-11)",
    }


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the share of the host given to one GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_reference_algorithm(x, f, y, p, n_real, threads):
    """The reference algorithm (FP64 oracle) on the host with `threads`
    threads: real space over all source pairs x 27 images (realShells,
    periodicBrinkman.h:564-693) on the first n_real targets, extrapolated
    linearly to all targets; k space by the spectral fftBrinkman algorithm
    (fftBrinkman.h:34-287) in full.  Returns (evals/s, real s, k s, real s
    per target)."""
    from oracle import pyoracle as po
    po.set_threads(threads)
    t0 = time.perf_counter()
    po.real_velocity_cube(y[:n_real], x, f, p.L, p.alpha, p.d, p.eps, (1, 1, 1))
    per_target = (time.perf_counter() - t0) / n_real
    t_real = per_target * y.shape[0]
    t0 = time.perf_counter()
    po.kspace_spectral(y, x, f, p.L, p.alpha, p.d, p.grid, (p.P,) * 3, p.m)
    t_k = time.perf_counter() - t0
    return y.shape[0] / (t_real + t_k), t_real, t_k, per_target


def cpu_baseline(x, f, y, p, n1, nN):
    """The reference algorithm on the host at 1 thread (the reference is
    single-threaded) and at every thread this process may use: real space
    sampled on n1 and nN targets, k space in full."""
    threads = host_threads()
    v1, r1, k1, pt1 = cpu_reference_algorithm(x, f, y, p, n1, 1)
    log(f"cpu 1 thread: {v1:.3f} evals/s (real {r1:.1f} s extrapolated, k {k1:.1f} s)")
    vN, rN, kN, ptN = cpu_reference_algorithm(x, f, y, p, nN, threads)
    log(f"cpu {threads} threads: {vN:.3f} evals/s (real {rN:.1f} s extrapolated, k {kN:.1f} s)")
    return {
        "value": vN, "unit": "velocity evals/s", "cores": threads, "kind": "port",
        "sample": f"FP64 oracle of the reference algorithm, {threads} threads (OpenMP over targets, spreading over "
                  f"grid slabs): real space {nN} targets x {NS} sources x 27 images (realShells), extrapolated "
                  f"to {NT} targets ({rN:.1f} s); k space in full, 128^3 spectral fftBrinkman ({kN:.1f} s)",
        "single_thread": {
            "value": v1, "cores": 1,
            "sample": f"real space {n1} targets x {NS} sources x 27 images, extrapolated to {NT} targets "
                      f"({r1:.1f} s); k space in full ({k1:.1f} s)"},
        "real_space_s_per_target": {"1_thread_%d_targets" % n1: pt1,
                                    "%d_threads_%d_targets_x_threads" % (threads, nN): ptN * threads},
    }


def config1_leg(ctx, steps, cpu):
    """The 1000-source workload (targets = sources, alpha 10, eps 0.01, 32^3,
    P 8, split for 1e-8): GPU mixed and FP64 forms; with cpu, the reference
    algorithm in full on the host (27 images, all pairs; spectral k space) at
    1 thread and at every host thread, and the errors against the oracle."""
    import torch

    import periodicbrinkman_amd as pb

    n = 1000
    x, f, _ = pb.uniform_system(pb.config_seed(0), n, 0)
    X = torch.as_tensor(x, device="cuda:0")
    F = torch.as_tensor(f, device="cuda:0")
    U = torch.empty_like(X)
    p = pb.Params(alpha=10.0, eps=0.01, grid=(32, 32, 32), P=8, tol=1e-8).tune(n, n)
    out = {"workload": "1000 sources uniform in the unit box, targets = sources, alpha 10, eps 0.01, 32^3 grid, "
                       "P 8, split for 1e-8 (d %.4f, rc %.4f)" % (p.d, p.rc)}
    res = {}
    for form, q in (("mixed", p.copy(precision="mixed")), ("fp64", p.copy(precision="fp64"))):
        for _ in range(5):
            ctx.velocity(q, X, F, X, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity(q, X, F, X, U)
        torch.cuda.synchronize()
        out["gpu_%s_ms" % form] = (time.perf_counter() - t0) * 1e3 / steps
        res[form] = U.cpu().numpy().copy()
    if cpu:
        from oracle import pyoracle as po
        want = po.velocity_ewald(x, x, f, p.L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
        for form in res:
            out["rel_l2_%s_vs_oracle" % form] = relL2(res[form], want)
        threads = host_threads()
        for th in (1, threads):
            po.set_threads(th)
            t0 = time.perf_counter()
            po.real_velocity_cube(x, x, f, p.L, p.alpha, p.d, p.eps, (1, 1, 1))
            po.kspace_spectral(x, x, f, p.L, p.alpha, p.d, p.grid, (p.P,) * 3, p.m)
            out["cpu_%d_threads_ms" % th] = (time.perf_counter() - t0) * 1e3
        out["gpu_mixed_vs_cpu_%d_threads" % threads] = out["cpu_%d_threads_ms" % threads] / out["gpu_mixed_ms"]
    return out


def oracle_subsample_check(x, f, y, pref, uref, nsub):
    """FP64 oracle (real cutoff + spectral) on nsub targets vs the GPU reference."""
    from oracle import pyoracle as po
    po.set_threads(0)
    idx = np.linspace(0, y.shape[0] - 1, nsub).astype(np.int64)
    want = po.velocity_ewald(y[idx], x, f, pref.L, pref.alpha, pref.d, pref.eps, pref.rc, pref.grid,
                             (pref.P,) * 3, pref.m)
    return relL2(uref[idx], want)


def profile_workload(name, steps, warmup):
    """Run one workload's forms for `steps` calls after `warmup`, with no
    reference evaluation and no CPU work, so a profiler run covers only the
    measured kernels: "headline" (mixed, FP64-table and plain FP64 forms of
    the 1e5 -> 1e5 system), "batched" (the 512 systems, mixed, graph replays),
    "clustered" (mixed), "multirhs" (64 force sets)."""
    import torch

    import periodicbrinkman_amd as pb

    ctx = pb.Context(0)
    out = {"profile": name, "steps": steps, "warmup": warmup}
    if name == "headline":
        x, f, y = pb.uniform_system(pb.config_seed(1), NS, NT)
        dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")
        xs, fs, ys = dev(x), dev(f), dev(y)
        u = torch.empty((NT, 3), dtype=torch.float64, device="cuda:0")
        p = pb.Params(alpha=ALPHA, eps=EPS, grid=GRID, P=P, tol=TOL).tune(NS, NT)
        for form, q in (("mixed", p.copy(precision="mixed")), ("fp64", p.copy(precision="fp64")),
                        ("plain_fp64", p.copy(precision="fp64", kernel="direct"))):
            for _ in range(warmup):
                ctx.velocity(q, xs, fs, ys, u)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.velocity(q, xs, fs, ys, u)
            torch.cuda.synchronize()
            out[form + "_ms"] = (time.perf_counter() - t0) * 1e3 / steps
            out[form + "_pairs"] = sum(ctx.last_pair_stats()[:2])
    elif name == "batched":
        systems, x, f, off = pb.batched_systems(pb.config_seed(3))
        X = torch.as_tensor(x, device="cuda:0")
        F = torch.as_tensor(f, device="cuda:0")
        U = torch.empty_like(X)
        ps = [pb.Params(L=(s["L"],) * 3, alpha=s["alpha"], eps=s["eps"], precision="mixed", tol=BATCH_TOL)
              for s in systems]
        for _ in range(max(2, warmup)):
            ctx.velocity_batched(ps, X, F, off, X, off, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity_batched(ps, X, F, off, X, off, U)
        torch.cuda.synchronize()
        out["mixed_ms"] = (time.perf_counter() - t0) * 1e3 / steps
    elif name == "clustered":
        n = 1_000_000
        x, f = pb.clustered_system(pb.config_seed(2), n, 64, 0.02)
        X = torch.as_tensor(x, device="cuda:0")
        F = torch.as_tensor(f, device="cuda:0")
        U = torch.empty_like(X)
        q = pb.Params(alpha=25.0, eps=5e-4, grid=(256, 256, 256), P=16, tol=CLUSTER_TOL,
                      precision="mixed").tune(n, n)
        for _ in range(warmup):
            ctx.velocity(q, X, F, X, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity(q, X, F, X, U)
        torch.cuda.synchronize()
        out["mixed_ms"] = (time.perf_counter() - t0) * 1e3 / steps
        out["mixed_pairs"] = sum(ctx.last_pair_stats()[:2])
    else:
        n, R = 200_000, 64
        x, _, _ = pb.uniform_system(pb.config_seed(4), n, 0)
        Fh = pb.Rng(pb.config_seed(4) + (1 << 32)).normal(3 * n * R).reshape(R, n, 3).astype(np.float32)
        X = torch.as_tensor(x, device="cuda:0")
        F = torch.as_tensor(Fh, device="cuda:0")
        q = multirhs_params(pb, n)
        U = ctx.velocity_multirhs(q, X, F, X)
        for _ in range(warmup):
            ctx.velocity_multirhs(q, X, F, X, U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.velocity_multirhs(q, X, F, X, U)
        torch.cuda.synchronize()
        out["mixed_ms"] = (time.perf_counter() - t0) * 1e3 / steps
        out["pairs"] = sum(ctx.last_pair_stats()[:2])
    ctx.close()
    print(json.dumps(out))


def copy_calibration():
    """A plain device-to-device copy of COPY_BYTES, three times: the HBM
    counters (FETCH_SIZE, WRITE_SIZE) of its kernel scale the pair kernel's."""
    import torch
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "copies": 3}))


COUNTERS_FILE = os.path.join(ROOT, "profiles", "kernel_counters.json")
PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E


def kernel_counters(workload):
    """Per-dispatch counter averages of one workload's kernels from the
    committed rocprofv3 summary (bench.py --profile <workload> under
    rocprofv3), or {}."""
    try:
        with open(COUNTERS_FILE) as fh:
            return json.load(fh)["workloads"].get(workload, {})
    except (OSError, ValueError, KeyError):
        return {}


def _entry(name, kernels, ms, counters, bound, algorithmic=None, note=None):
    """One per-kernel roofline entry: the counted work per launch (sum over
    `kernels`) over the launch duration `ms`, against the MI355X peaks."""
    c = [counters[k] for k in kernels if k in counters]
    if not c or not ms:
        return None
    t = ms / 1e3
    f64 = sum(e["valu_fp64_flop"] for e in c)
    f32 = sum(e["valu_fp32_flop"] for e in c)
    mf = sum(e["mfma_fp32_flop"] for e in c)
    hbm = sum(e["hbm_bytes"] for e in c)
    e = {"kernel": name, "kernels": kernels, "ms": ms, "bound": bound,
         "counted": {"valu_fp64_tflops": f64 / t / 1e12, "valu_fp32_tflops": f32 / t / 1e12,
                     "mfma_fp32_tflops": mf / t / 1e12, "hbm_gbps": hbm / t / 1e9, "hbm_bytes": hbm}}
    if bound == "valu-fp64":
        e["achieved"], e["peak"], e["unit"] = f64 / t / 1e12, PEAK_FP64, "TFLOP/s"
    elif bound == "valu-fp32":
        e["achieved"], e["peak"], e["unit"] = f32 / t / 1e12, PEAK_FP32, "TFLOP/s"
    elif bound == "mfma":
        e["achieved"], e["peak"], e["unit"] = mf / t / 1e12, PEAK_FP32, "TFLOP/s"
    else:
        e["achieved"], e["peak"], e["unit"] = hbm / t / 1e9, PEAK_HBM, "GB/s"
    e["frac"] = e["achieved"] / e["peak"]
    busy = [x["mfma_busy_frac"] for x in c if x.get("mfma_busy_frac")]
    if bound == "mfma" and busy:
        e["mfma_busy_frac"] = busy[0]
    # busy fractions of the units a launch can saturate (counter runs, one
    # kernel at a time; the largest over `kernels`): the vector ALU issue,
    # and the texture address (TA) and data (TD) units of the vector memory
    # path, through which every global load passes
    if all("td_busy_frac" in x for x in c):
        e["unit_busy"] = {u: max(x[u + "_busy_frac"] for x in c) for u in ("valu", "ta", "td")}
    if algorithmic:
        work, unit, peak = algorithmic
        e["algorithmic"] = {"per_launch": work, "rate": work / t / (1e12 if unit == "TFLOP/s" else 1e9),
                            "unit": unit, "frac": work / t / (1e12 if unit == "TFLOP/s" else 1e9) / peak}
    if note:
        e["note"] = note
    return e


def headline_kernel_rooflines(ph, pairs, prec="float", pair_kernel="k_pair_table<false, 6>"):
    """Counter-based roofline of every hot kernel of the headline step: work
    counted by rocprofv3 (committed summary) over the kernel's live duration
    in this run (HIP events: the pair kernel beside the k space and alone, the
    other phases one after another)."""
    cnt = kernel_counters("headline")
    if not cnt:
        return []
    n3 = GRID[0] * GRID[1] * GRID[2]
    ngc = GRID[0] * GRID[1] * (GRID[2] // 2 + 1)
    spread_flop = 7.0 * P ** 3 * NS
    gather_flop = 7.0 * P ** 3 * NT
    fft_bytes = 2 * 3 * ngc * 8  # one transform of the three fields, read + write (FP32 complex, half spectrum)
    out = [
        _entry("pair kernel (beside the k space)", [pair_kernel], ph["pair"], cnt, "valu-fp64",
               (FLOP_PER_PAIR * pairs, "TFLOP/s", PEAK_FP64)),
        _entry("pair kernel (alone)", [pair_kernel], ph["pair_alone"], cnt, "valu-fp64",
               (FLOP_PER_PAIR * pairs, "TFLOP/s", PEAK_FP64)),
        _entry("spread", ["k_spread_regions<%s>" % prec], ph["spread"], cnt, "mfma" if prec == "float" else "valu-fp64",
               (spread_flop, "TFLOP/s", PEAK_FP32 if prec == "float" else PEAK_FP64),
               "the one-set FP32 spread's multiply-adds run on the matrix cores (v_mfma_f32_16x16x4_f32; "
               "counted.mfma_fp32_tflops), its window weights and staging on the vector ALUs; "
               "algorithmic: 7 P^3 FLOP per source; its bytes (24 B per source + 12 B per grid point) are in "
               "counted.hbm_gbps's terms %.1f MB" % ((24 * NS + 12 * n3) / 1e6)),
        _entry("FFT forward (z r2c + y)", ["k_fft_z_r2c<%s, 128>" % prec, "k_fft_y<%s, 128>" % prec],
               ph["fft_forward"], cnt, "hbm"),
        _entry("FFT x pass with the Brinkman multiply (forward + inverse)", ["k_fft_x_mult<%s, 128>" % prec],
               ph["multiply"], cnt, "hbm"),
        _entry("FFT inverse (y + z c2r)", ["k_fft_y<%s, 128>" % prec, "k_fft_z_c2r<%s, 128>" % prec],
               ph["fft_inverse"], cnt, "hbm"),
        _entry("FFT and multiply, all five passes (forward and inverse transforms)",
               ["k_fft_z_r2c<%s, 128>" % prec, "k_fft_y<%s, 128>" % prec, "k_fft_x_mult<%s, 128>" % prec,
                "k_fft_y<%s, 128>" % prec, "k_fft_z_c2r<%s, 128>" % prec],
               ph["fft_forward"] + ph["multiply"] + ph["fft_inverse"], cnt, "hbm",
               (2 * fft_bytes, "GB/s", PEAK_HBM),
               "algorithmic: 2 transforms x 2 (read + write) x 3 fields x n^2 (n/2 + 1) x 8 B"),
        _entry("gather", ["k_interp<%s, 12>" % prec], ph["gather"], cnt, "valu-fp32",
               (gather_flop, "TFLOP/s", PEAK_FP32),
               "bound in practice by the vector memory path: the texture address and data units are busy "
               "most of the kernel's cycles (unit_busy) for the 12-byte loads of the P^3 window points per target"),
        _entry("binning", ["k_fill_int2", "k_bin_count2", "k_bin_sums2", "k_bin_apply2", "k_bin_scatter2"],
               ph["binning"], cnt, "hbm"),
    ]
    return [e for e in out if e]


def leg_kernel_rooflines(workload, spec):
    """Per-kernel roofline of a leg from the committed counters and kernel
    trace (durations: the trace average of the profiled run).  spec: (label,
    kernels, bound, algorithmic or None) per entry."""
    cnt = kernel_counters(workload)
    out = []
    for label, kernels, bound, alg in spec:
        ms = sum(cnt[k]["trace_avg_us"] for k in kernels if k in cnt) / 1e3
        e = _entry(label, kernels, ms, cnt, bound, alg,
                   "duration: rocprofv3 trace average (launches beside other kernels included); at_min_duration: "
                   "the same work over the trace minimum, the kernel with the GPU to itself")
        if e:
            ms_min = sum(cnt[k]["trace_min_us"] for k in kernels if k in cnt) / 1e3
            m = _entry(label, kernels, ms_min, cnt, bound, alg)
            e["at_min_duration"] = {"ms": ms_min, "achieved": m["achieved"], "frac": m["frac"]}
            if "algorithmic" in m:
                e["at_min_duration"]["algorithmic_frac"] = m["algorithmic"]["frac"]
            out.append(e)
    return out


def pair_traffic(kernel="k_pair_table<false, 6>"):
    """HBM bytes per launch of the headline's pair kernel from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE of bench.py --profile headline,
    the same parameters as the timed mixed form, scaled by what the counters
    read for a copy of known size), or None."""
    e = kernel_counters("headline").get(kernel)
    return e.get("hbm_bytes") if e else None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-systems leg")
    ap.add_argument("--no-more", action="store_true", help="skip the clustered and multi-RHS legs")
    ap.add_argument("--cpu-targets", type=int, default=32,
                    help="targets in the single-thread CPU real-space sample (all 1e5 sources x 27 images each)")
    ap.add_argument("--cpu-targets-threads", type=int, default=1024,
                    help="targets in the CPU real-space sample at every host thread")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of known size (for HBM counter calibration)")
    ap.add_argument("--only", choices=("batched", "clustered", "multirhs"),
                    help="run one extra leg alone and print its JSON")
    ap.add_argument("--profile", choices=("headline", "batched", "clustered", "multirhs"),
                    help="only run one workload's forms for --steps calls (no references, no CPU work: "
                         "for rocprofv3)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile:
        profile_workload(args.profile, args.steps, args.warmup)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import periodicbrinkman_amd as pb

    if args.only:
        ctx = pb.Context(0)
        leg = {"batched": lambda: batched_leg(ctx, args.steps, args.warmup),
               "clustered": lambda: clustered_leg(ctx, 3), "multirhs": lambda: multirhs_leg(ctx, 3)}[args.only]
        print(json.dumps({args.only + "_form": leg()}))
        ctx.close()
        return

    x, f, y = pb.uniform_system(pb.config_seed(1), NS, NT)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")
    xs, fs, ys = dev(x), dev(f), dev(y)
    u = torch.empty((NT, 3), dtype=torch.float64, device="cuda:0")
    ctx = pb.Context(0)

    p = pb.Params(alpha=ALPHA, eps=EPS, grid=GRID, P=P, tol=TOL).tune(NS, NT)
    pm, p64 = p.copy(precision="mixed"), p.copy(precision="fp64")
    pplain = p.copy(precision="fp64", kernel="direct")
    pfar32 = p.copy(precision="mixed", kernel="table_far32")

    log(f"params {p}")
    ms_mixed, ph_mixed = timed_steps(ctx, pm, xs, fs, ys, u, args.steps, args.warmup)
    log(f"mixed form: {ms_mixed:.3f} ms/step {ph_mixed}")
    u_mixed = u.cpu().numpy().copy()
    n64, n32 = ctx.last_pair_counts()
    stats = ctx.last_pair_stats()
    log(f"pair work: fp64 pairs {stats[0]} fp32 pairs {stats[1]} candidates tested {stats[2]} "
        f"read {stats[3]}")
    ms_fp64, ph_fp64 = timed_steps(ctx, p64, xs, fs, ys, u, args.steps, args.warmup)
    u_fp64 = u.cpu().numpy().copy()
    n64b, _ = ctx.last_pair_counts()
    log(f"fp64 form: {ms_fp64:.3f} ms/step {ph_fp64}")
    ms_plain, ph_plain = timed_steps(ctx, pplain, xs, fs, ys, u, args.steps, args.warmup)
    u_plain = u.cpu().numpy().copy()
    log(f"plain fp64 (closed forms) form: {ms_plain:.3f} ms/step {ph_plain}")
    ms_far32, ph_far32 = timed_steps(ctx, pfar32, xs, fs, ys, u, args.steps, args.warmup)
    u_far32 = u.cpu().numpy().copy()
    log(f"mixed form with FP32 far-field table: {ms_far32:.3f} ms/step {ph_far32}")

    # FP64 reference with tighter parameters; a second split checks it
    lt = np.log(1e13)
    kmax = np.pi * GRID[0]
    ref = []
    for dref in (2 * np.sqrt(lt + 2) / kmax, 2 * np.sqrt(lt + 4) / kmax):
        pr = pb.Params(alpha=ALPHA, eps=EPS, d=dref, rc=dref * np.sqrt(lt + 2), grid=GRID, P=20,
                       precision="fp64", tol=1e-13, kernel="direct").tune(NS, NT)
        ref.append((pr, ctx.velocity(pr, xs, fs, ys).cpu().numpy()))
    torch.cuda.synchronize()
    pref, uref = ref[0]
    err_mixed = relL2(u_mixed, uref)
    err_fp64 = relL2(u_fp64, uref)
    err_plain = relL2(u_plain, uref)
    err_far32 = relL2(u_far32, uref)
    ref_split = relL2(ref[1][1], uref)
    log(f"rel L2 mixed {err_mixed:.3e} fp64 {err_fp64:.3e} reference split check {ref_split:.3e}")

    value = NT / (ms_mixed / 1e3)
    pairs = n64 + n32
    t_pair = ph_mixed["pair"] / 1e3
    achieved = FLOP_PER_PAIR * pairs / t_pair / 1e12
    f64 = n64 / pairs
    peak = 1.0 / (f64 / PEAK_FP64 + (1 - f64) / PEAK_FP32)
    workload = (f"single periodic system, {NS} sources -> {NT} separate targets, alpha {ALPHA:g}, "
                f"eps {EPS:g}, {GRID[0]}^3 grid, P {P}, mixed FP32/FP64")
    result = {
        "metric": "velocity evals/sec at <=1e-6 rel. L2 vs FP64 Ewald oracle; % of roofline",
        "value": value,
        "unit": "velocity evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (xoshiro256** seed 0x5EED0001: uniform positions, N(0,1) forces)",
        "config": {"workload": workload,
                   "ns": NS, "nt": NT, "d": p.d, "rc": p.rc, "grid": list(GRID), "P": P,
                   "m": p.m},
        "rel_l2_mixed": err_mixed,
        "meets_accuracy": bool(err_mixed <= ACCURACY),
        "tuner_tol": TOL,
        "rel_l2_fp64": err_fp64,
        "reference_split_check": ref_split,
        "rel_l2_plain_fp64": err_plain,
        "fp64_form": {"value": NT / (ms_fp64 / 1e3), "ms_per_step": ms_fp64,
                      "phases_ms": ph_fp64, "pairs_fp64": n64b,
                      "note": "every step FP64; real-space radial functions from the FP64 table"},
        "mixed_far32_form": {"value": NT / (ms_far32 / 1e3), "ms_per_step": ms_far32,
                             "phases_ms": ph_far32, "rel_l2": err_far32,
                             "note": "mixed form with pairs beyond max(3d, 8e) from an FP32 table"},
        "plain_fp64_form": {"value": NT / (ms_plain / 1e3), "ms_per_step": ms_plain,
                            "phases_ms": ph_plain,
                            "note": "every step FP64; closed forms (exp, erfcx) evaluated per pair"},
        "speedup_mixed_vs_fp64": ms_fp64 / ms_mixed,
        "speedup_mixed_vs_plain_fp64": ms_plain / ms_mixed,
        "phases_ms": ph_mixed,
        "pairs": {"fp64": n64, "fp32": n32},
        "roofline": {"bound": "valu-fp64", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": pair_traffic(),
                     "kernel": "k_pair_table (real-space pair kernel)",
                     "per_kernel": headline_kernel_rooflines(ph_mixed, pairs),
                     "note": "the FP64 vector-ALU roof (78.6 TFLOP/s on MI355X) of the pair kernel, which "
                             "computes on the vector ALUs (no matrix cores); achieved = the frozen "
                             "%d FLOP per pair-image inside rc x %d pair-images per launch / "
                             "%.4f ms per launch (HIP events on the launch stream around each "
                             "launch, in as many steps as the timed ones, where it runs "
                             "concurrently with the k-space pipeline on another stream; alone "
                             "it takes %.4f ms, frac %.3f); FP64 pair fraction %.3f"
                             % (FLOP_PER_PAIR, pairs, ph_mixed["pair"], ph_mixed["pair_alone"],
                                FLOP_PER_PAIR * pairs / (ph_mixed["pair_alone"] / 1e3) / 1e12 / peak, f64)},
    }
    if not args.no_batched:
        result["batched_form"] = batched_leg(ctx, args.steps, args.warmup)
        result["batched_form"]["per_kernel_roofline"] = leg_kernel_rooflines("batched", [
            ("pair kernel (degree-6 table; every system's class)", ["k_gpair_table<false, 6>"], "valu-fp64", None),
            ("spread", ["k_gspread<float>"], "mfma", None),
            ("gather", ["k_ginterp<float, 10>"], "valu-fp32", None),
            ("FFT passes, 64^3 systems", ["k_gfft<float, 64, %d>" % q for q in (5, 2, 6)], "hbm", None),
            ("FFT passes, 32^3 systems", ["k_gfft<float, 32, %d>" % q for q in (5, 2, 6)], "hbm", None)])
    if not args.no_more:
        result["clustered_form"] = clustered_leg(ctx, 3)
        result["clustered_form"]["per_kernel_roofline"] = leg_kernel_rooflines("clustered", [
            ("pair kernel", ["k_pair_table<false, 6>"], "valu-fp64", None),
            ("spread (count, chunks and regions, chunk reduction)",
             ["k_spread_count", "k_spread_items<float>", "k_spread_reduce<float>"], "mfma", None),
            ("gather", ["k_interp<float, 16>"], "valu-fp32", None)])
        result["multirhs_form"] = multirhs_leg(ctx, 3)
        mr = result["multirhs_form"]
        result["multirhs_form"]["per_kernel_roofline"] = leg_kernel_rooflines("multirhs", [
            ("contraction (MFMA)", ["k_pair_mrhs<4, 2, 6>"], "mfma",
             (18.0 * 64 * mr.get("pairs", 0), "TFLOP/s", PEAK_FP32)),
            ("spread, 8 sets per workgroup", ["k_gspread_sets<8>"], "valu-fp32", None),
            ("gather", ["k_ginterp<float, %d>" % mr.get("P", 10)], "valu-fp32", None),
            ("FFT x pass with the multiply", ["k_gfft<float, 128, 2>"], "hbm", None)])
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(x, f, y, p64, args.cpu_targets, args.cpu_targets_threads)
        log("oracle subsample check")
        result["reference_vs_oracle_subsample"] = oracle_subsample_check(x, f, y, pref, uref, 256)
    result["config1_form"] = config1_leg(ctx, 100, not args.no_cpu)
    if "cpu_baseline" in result:
        result["cpu_baseline"]["config1_full"] = {
            k: v for k, v in result["config1_form"].items() if k.startswith("cpu_") or k == "workload"}
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
