"""Many right-hand sides on one particle set (pb_velocity_multirhs): every
force set against the FP64 oracle and against the single-set mixed path."""
import numpy as np
import pytest

import periodicbrinkman_amd as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def relL2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("R,same", [(5, False), (13, False), (40, True)])
def test_multirhs_matches_oracle(R, same):
    import torch
    ns, nt = 3000, 2000
    L = (1.0, 1.0, 1.0)
    x, _, y = pb.uniform_system(pb.config_seed(4), ns, nt, L)
    if same:
        y = x[:nt].copy()  # targets on sources: self pairs
    rng = np.random.default_rng(7)
    F = rng.standard_normal((R, ns, 3)).astype(np.float32)
    p = pb.Params(L=L, alpha=10.0, eps=1e-3, grid=(32, 32, 32), P=12, tol=1e-7,
                  precision="mixed").tune(ns, nt)
    ctx = pb.Context(0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    u = ctx.velocity_multirhs(p, dev(x), dev(F), dev(y)).cpu().numpy()
    torch.cuda.synchronize()
    assert u.shape == (R, nt, 3)
    for r in sorted({0, R // 2, R - 1}):
        f64 = F[r].astype(np.float64)
        want = po.velocity_ewald(y, x, f64, L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
        assert relL2(u[r], want) < 1e-6, r
        single = ctx.velocity(p, dev(x), dev(f64), dev(y)).cpu().numpy()
        assert relL2(u[r], single) < 1e-6, r
    # linearity across sets: the sum of two sets is the velocity of their sum
    F2 = np.stack([F[0] + F[1]])
    u2 = ctx.velocity_multirhs(p, dev(x), dev(F2.astype(np.float32)), dev(y)).cpu().numpy()
    assert relL2(u2[0], u[0] + u[1]) < 1e-6
    ctx.close()


def test_multirhs_rejects_closed_form_kernel():
    import torch
    ctx = pb.Context(0)
    p = pb.Params(alpha=10.0, eps=1e-3, grid=(16, 16, 16), P=8, tol=1e-5, kernel="direct").tune(10, 10)
    x = torch.rand((10, 3), dtype=torch.float64, device="cuda")
    F = torch.rand((2, 10, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(pb.PBError) as ei:
        ctx.velocity_multirhs(p, x, F, x)
    assert ei.value.status == pb.PB_ERR_UNSUPPORTED
    ctx.close()


def test_multirhs_dense_staging_matches_single_sets():
    # cells much smaller than a spreading region plus its window (grid 32,
    # P 12, rc 0.1): nearly every candidate source of a region is staged, so
    # chunks of 256 candidates stage whole after a flush -- the case the
    # grouped spread's staging capacity must hold.  Every set against the
    # single-set call.
    import torch
    n, R = 20000, 16
    L = (1.0, 1.0, 1.0)
    x, _, _ = pb.uniform_system(pb.config_seed(8), n, 0, L)
    rng = np.random.default_rng(9)
    F = rng.standard_normal((R, n, 3)).astype(np.float32)
    p = pb.Params(L=L, alpha=5.0, eps=1e-3, d=0.07, rc=0.1, grid=(32, 32, 32), P=12, m=float(np.sqrt(np.pi * 12)),
                  precision="mixed")
    ctx = pb.Context(0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    X = dev(x)
    u = ctx.velocity_multirhs(p, X, dev(F), X).cpu().numpy()
    assert np.all(np.isfinite(u))
    for r in range(R):
        single = ctx.velocity(p, X, dev(F[r].astype(np.float64)), X).cpu().numpy()
        assert relL2(u[r], single) < 1e-6, (r, relL2(u[r], single))
    ctx.close()


@pytest.mark.parametrize("P,grid", [(6, 16), (8, 16), (10, 32), (16, 32)])
def test_multirhs_windows_match_single_sets(P, grid):
    # window sizes other than the 12 of the cases above (the gather's one and
    # four targets per wavefront, grids of 16 and 32): every set against the
    # single-set call
    import torch
    ns, nt, R = 4000, 3000, 6
    L = (1.0, 1.0, 1.0)
    x, _, y = pb.uniform_system(pb.config_seed(10), ns, nt, L)
    rng = np.random.default_rng(11)
    F = rng.standard_normal((R, ns, 3)).astype(np.float32)
    p = pb.Params(L=L, alpha=8.0, eps=1e-3, d=0.05, rc=0.12, grid=(grid,) * 3, P=P, m=float(np.sqrt(np.pi * P)),
                  precision="mixed")
    ctx = pb.Context(0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    X, Y = dev(x), dev(y)
    u = ctx.velocity_multirhs(p, X, dev(F), Y).cpu().numpy()
    assert np.all(np.isfinite(u))
    for r in range(R):
        single = ctx.velocity(p, X, dev(F[r].astype(np.float64)), Y).cpu().numpy()
        assert relL2(u[r], single) < 1e-6, (r, relL2(u[r], single))
    ctx.close()


def test_multirhs_repeated_calls_follow_the_forces():
    # a repeated call with the same layout reuses the uploaded descriptors:
    # new forces in the same buffer, and a second force buffer, must both
    # give their own velocities
    import torch
    ns, nt, R = 3000, 2000, 8
    L = (1.0, 1.0, 1.0)
    x, _, y = pb.uniform_system(pb.config_seed(12), ns, nt, L)
    rng = np.random.default_rng(13)
    p = pb.Params(L=L, alpha=8.0, eps=1e-3, d=0.05, rc=0.12, grid=(32,) * 3, P=12, m=float(np.sqrt(np.pi * 12)),
                  precision="mixed")
    ctx = pb.Context(0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    X, Y = dev(x), dev(y)
    F = torch.empty((R, ns, 3), dtype=torch.float32, device="cuda")
    U = torch.empty((R, nt, 3), dtype=torch.float64, device="cuda")
    for k in range(3):
        Fh = rng.standard_normal((R, ns, 3)).astype(np.float32)
        F.copy_(dev(Fh))
        src = F if k < 2 else dev(Fh)  # the third call from another buffer
        ctx.velocity_multirhs(p, X, src, Y, U)
        u = U.cpu().numpy()
        for r in (0, R - 1):
            single = ctx.velocity(p, X, dev(Fh[r].astype(np.float64)), Y).cpu().numpy()
            assert relL2(u[r], single) < 1e-6, (k, r, relL2(u[r], single))
    ctx.close()


def test_multirhs_more_than_64_sets():
    # R = 80: two k-space chunks are not needed (the grids fit), but the
    # real space runs a 64-set and a 32-set contraction (RP = 96) and the
    # spreading takes 10 groups of 8 sets; the last set (79) against the
    # single-set call and the oracle
    import torch
    ns, nt, R = 3000, 2000, 80
    L = (1.0, 1.0, 1.0)
    x, _, y = pb.uniform_system(pb.config_seed(14), ns, nt, L)
    rng = np.random.default_rng(15)
    F = rng.standard_normal((R, ns, 3)).astype(np.float32)
    p = pb.Params(L=L, alpha=10.0, eps=1e-3, grid=(32, 32, 32), P=12, tol=1e-7, precision="mixed").tune(ns, nt)
    ctx = pb.Context(0)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    X, Y = dev(x), dev(y)
    u = ctx.velocity_multirhs(p, X, dev(F), Y).cpu().numpy()
    assert u.shape == (R, nt, 3) and np.all(np.isfinite(u))
    for r in (0, 31, 63, 64, 79):
        single = ctx.velocity(p, X, dev(F[r].astype(np.float64)), Y).cpu().numpy()
        assert relL2(u[r], single) < 1e-6, (r, relL2(u[r], single))
    f79 = F[79].astype(np.float64)
    want = po.velocity_ewald(y, x, f79, L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
    assert relL2(u[79], want) < 1e-6
    ctx.close()


@pytest.mark.parametrize("grid,P", [((4, 4, 4), 12), ((8, 8, 8), 20), ((16, 12, 16), 8), ((16, 16, 16), 0)])
def test_invalid_grid_or_window_rejected_before_any_work(grid, P):
    # checked before binning or any launch, on every velocity entry point: a
    # window wider than the grid would make the gather read outside the grid
    import torch
    ctx = pb.Context(0)
    n = 50
    x = torch.rand((n, 3), dtype=torch.float64, device="cuda")
    F = torch.rand((2, n, 3), dtype=torch.float32, device="cuda")
    p = pb.Params(alpha=10.0, eps=1e-3, d=0.05, rc=0.2, grid=grid, P=P, m=float(np.sqrt(np.pi * P)),
                  precision="mixed")
    u = torch.full((n, 3), 3.0, dtype=torch.float64, device="cuda")
    U = torch.full((2, n, 3), 3.0, dtype=torch.float64, device="cuda")
    calls = [lambda: ctx.velocity(p, x, x, x, u), lambda: ctx.velocity_kspace(p, x, x, x, u),
             lambda: ctx.velocity_multirhs(p, x, F, x, U),
             lambda: ctx.velocity_batched([p], x, x, [0, n], x, [0, n], u)]
    for call in calls:
        with pytest.raises(pb.PBError) as ei:
            call()
        assert ei.value.status in (pb.PB_ERR_INVALID, pb.PB_ERR_UNSUPPORTED)
    torch.cuda.synchronize()
    assert bool(torch.all(u == 3.0)) and bool(torch.all(U == 3.0))
    ctx.close()


def test_python_wrappers_reject_mismatched_shapes():
    import torch
    ctx = pb.Context(0)
    p = pb.Params(alpha=10.0, eps=1e-3, d=0.05, rc=0.2, grid=(16, 16, 16), P=8, m=float(np.sqrt(np.pi * 8)),
                  precision="mixed")
    x = torch.rand((40, 3), dtype=torch.float64, device="cuda")
    for args in [(x, x[:39], x), (x, x, x, torch.empty((39, 3), dtype=torch.float64, device="cuda"))]:
        with pytest.raises(pb.PBError) as ei:
            ctx.velocity(p, *args)
        assert ei.value.status == pb.PB_ERR_INVALID
    F = torch.rand((2, 39, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(pb.PBError):
        ctx.velocity_multirhs(p, x, F, x)
    with pytest.raises(pb.PBError):
        ctx.velocity_batched([p], x, x, [0, 41], x, [0, 40])
    ctx.close()
