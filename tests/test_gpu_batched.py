"""Independent periodic systems of uneven size in one call
(pb_velocity_batched): each system against the FP64 oracle with its own
parameters, and against a single-system call."""
import numpy as np
import pytest

import periodicbrinkman_amd as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def relL2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_batched_systems_match_oracle(precision, tol):
    import torch
    rng = np.random.default_rng(21)
    sizes = [(120, 100), (2500, 900), (700, 1300), (40, 40), (1600, 1600), (300, 500)]
    params, xs, fs, ys = [], [], [], []
    for i, (ns, nt) in enumerate(sizes):
        Ls = float(rng.uniform(0.5, 2.0))
        L = (Ls, Ls * float(rng.uniform(0.8, 1.25)), Ls)
        alpha = float(np.exp(rng.uniform(0, np.log(50))))
        x, f, y = pb.uniform_system(pb.config_seed(100 + i), ns, nt, L)
        q = pb.Params(L=L, alpha=alpha, eps=1e-3 * Ls, precision=precision, tol=1e-8 if precision == "fp64" else 1e-7)
        if i % 2 == 0:  # half the systems with a fixed grid, half autotuned
            q.grid, q.P = (32, 32, 32), 12
            q.tune(ns, nt)
        params.append(q)
        xs.append(x), fs.append(f), ys.append(y)
    so = np.concatenate([[0], np.cumsum([s[0] for s in sizes])])
    to = np.concatenate([[0], np.cumsum([s[1] for s in sizes])])
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    ctx = pb.Context(0)
    u = ctx.velocity_batched(params, dev(np.concatenate(xs)), dev(np.concatenate(fs)), so,
                             dev(np.concatenate(ys)), to).cpu().numpy()
    torch.cuda.synchronize()
    for i, q in enumerate(params):
        assert q.d > 0 and min(q.grid) >= 4
        ui = u[to[i]:to[i + 1]]
        want = po.velocity_ewald(ys[i], xs[i], fs[i], q.L, q.alpha, q.d, q.eps, q.rc, q.grid, (q.P,) * 3, q.m)
        assert relL2(ui, want) < tol, (i, relL2(ui, want))
        single = ctx.velocity(q, dev(xs[i]), dev(fs[i]), dev(ys[i])).cpu().numpy()
        # the same kernels; in mixed precision the FP32 spread sums in the
        # (atomic, run-dependent) source order of the counting sort
        assert relL2(ui, single) < (1e-12 if precision == "fp64" else 1e-6), i
    ctx.close()


def test_batched_empty_and_invalid():
    import torch
    ctx = pb.Context(0)
    z = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    u = ctx.velocity_batched([], z, z, [0], z, [0])
    assert u.shape == (0, 3)
    with pytest.raises(pb.PBError) as ei:
        ctx.velocity_batched([pb.Params(alpha=1.0)], z, z, [1, 0], z, [0, 0])
    assert ei.value.status == pb.PB_ERR_INVALID
    ctx.close()


def test_batched_graph_replays_follow_the_inputs():
    # the second identical call captures a graph, later ones replay it; the
    # replays read the current forces (scaled in place between calls), with
    # eager device work and synchronizations in between
    import torch
    sizes = [(400, 300), (900, 800), (150, 150)]
    params, xs, fs, ys = [], [], [], []
    for i, (ns, nt) in enumerate(sizes):
        L = (1.0 + 0.2 * i,) * 3
        x, f, y = pb.uniform_system(pb.config_seed(300 + i), ns, nt, L)
        params.append(pb.Params(L=L, alpha=5.0 + 3 * i, eps=1e-3, precision="fp64", tol=1e-8))
        xs.append(x), fs.append(f), ys.append(y)
    so = np.concatenate([[0], np.cumsum([s[0] for s in sizes])])
    to = np.concatenate([[0], np.cumsum([s[1] for s in sizes])])
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    X, F, Y = dev(np.concatenate(xs)), dev(np.concatenate(fs)), dev(np.concatenate(ys))
    U = torch.empty((Y.shape[0], 3), dtype=torch.float64, device="cuda")
    ctx = pb.Context(0)
    ref = ctx.velocity_batched(params, X, F, so, Y, to, U).cpu().numpy()  # eager
    for k, scale in enumerate([1.0, 2.0, -0.5, 1.0]):
        F.copy_(dev(np.concatenate(fs) * scale))
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        got = ctx.velocity_batched(params, X, F, so, Y, to, U).cpu().numpy()
        assert relL2(got, scale * ref) < 1e-12, (k, scale)
    del junk
    ctx.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_batched_grouped_and_single_systems_mixed(precision, tol):
    # systems whose cell lists fit one scan workgroup run grouped (one launch
    # per phase for all of them); the middle system's fine cell grid (rc small
    # against L: ncell + target keys > 2^18) sends it through the per-system
    # path.  Every system against the oracle and the single-system call; the
    # pair statistics of the call cover the grouped systems.
    import torch
    rng = np.random.default_rng(5)
    sizes = [(600, 500), (400, 300), (700, 900), (64, 1)]
    params, xs, fs, ys = [], [], [], []
    for i, (ns, nt) in enumerate(sizes):
        L = (1.0, 1.0, 1.0)
        x, f, y = pb.uniform_system(pb.config_seed(500 + i), ns, nt, L)
        if i == 1:
            q = pb.Params(L=L, alpha=4.0, eps=1e-3, d=0.03, rc=0.06, grid=(64, 64, 64), P=8,
                          m=float(np.sqrt(np.pi * 8)), precision=precision)
        else:
            q = pb.Params(L=L, alpha=float(rng.uniform(1, 30)), eps=1e-3, precision=precision,
                          tol=1e-8 if precision == "fp64" else 1e-7)
        params.append(q)
        xs.append(x), fs.append(f), ys.append(y)
    so = np.concatenate([[0], np.cumsum([s[0] for s in sizes])])
    to = np.concatenate([[0], np.cumsum([s[1] for s in sizes])])
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    ctx = pb.Context(0)
    X, F, Y = dev(np.concatenate(xs)), dev(np.concatenate(fs)), dev(np.concatenate(ys))
    u = ctx.velocity_batched(params, X, F, so, Y, to).cpu().numpy()
    stats = ctx.last_pair_stats()
    u2 = ctx.velocity_batched(params, X, F, so, Y, to).cpu().numpy()  # captured
    u3 = ctx.velocity_batched(params, X, F, so, Y, to).cpu().numpy()  # replayed
    for i, q in enumerate(params):
        ui = u[to[i]:to[i + 1]]
        want = po.velocity_ewald(ys[i], xs[i], fs[i], q.L, q.alpha, q.d, q.eps, q.rc, q.grid, (q.P,) * 3, q.m)
        assert relL2(ui, want) < tol, (i, relL2(ui, want))
        single = ctx.velocity(q, dev(xs[i]), dev(fs[i]), dev(ys[i])).cpu().numpy()
        assert relL2(ui, single) < (1e-12 if precision == "fp64" else 1e-6), i
    assert relL2(u2, u) < (1e-12 if precision == "fp64" else 1e-6)
    assert relL2(u3, u) < (1e-12 if precision == "fp64" else 1e-6)
    assert stats[0] + stats[1] > 0
    ctx.close()


def test_batched_group_edge_systems():
    # the group with systems that have no sources or no targets, FP64 and
    # mixed systems side by side, and one target: every system against the
    # oracle (a system without sources has zero velocity)
    import torch
    sizes = [(0, 50), (300, 0), (1, 1), (500, 400), (0, 0), (200, 1), (350, 280)]
    precs = ["fp64", "mixed", "fp64", "mixed", "fp64", "fp64", "mixed"]
    params, xs, fs, ys = [], [], [], []
    for i, ((ns, nt), pr) in enumerate(zip(sizes, precs)):
        L = (0.8 + 0.1 * i, 1.0, 1.2)
        x, f, y = pb.uniform_system(pb.config_seed(700 + i), ns, nt, L)
        params.append(pb.Params(L=L, alpha=3.0 + 2 * i, eps=1e-3, precision=pr,
                                tol=1e-8 if pr == "fp64" else 1e-7))
        xs.append(x.reshape(-1, 3)), fs.append(f.reshape(-1, 3)), ys.append(y.reshape(-1, 3))
    so = np.concatenate([[0], np.cumsum([s[0] for s in sizes])])
    to = np.concatenate([[0], np.cumsum([s[1] for s in sizes])])
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    ctx = pb.Context(0)
    X, F, Y = dev(np.concatenate(xs)), dev(np.concatenate(fs)), dev(np.concatenate(ys))
    for call in range(3):  # eager, captured, replayed
        u = ctx.velocity_batched(params, X, F, so, Y, to).cpu().numpy()
        for i, q in enumerate(params):
            ui = u[to[i]:to[i + 1]]
            if sizes[i][1] == 0:
                continue
            if sizes[i][0] == 0:
                assert np.all(ui == 0), i
                continue
            want = po.velocity_ewald(ys[i], xs[i], fs[i], q.L, q.alpha, q.d, q.eps, q.rc, q.grid, (q.P,) * 3, q.m)
            assert relL2(ui, want) < (1e-10 if precs[i] == "fp64" else 1e-6), (call, i, relL2(ui, want))
    ctx.close()
