"""pb_velocity_noalloc (the entry point that allocates nothing) against
pb_velocity through the C ABI: two calls in a row, the second on a smaller
input with other parameters, in both precisions; a workload beyond the
scratch limit is rejected before anything is enqueued."""
import numpy as np
import pytest

import periodicbrinkman_amd as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def relL2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


# (ns, nt, grid, P, tol): the first call at the 1e5 x 1e5, 128^3 workload,
# the second smaller, on another grid and window
CALLS = [(100_000, 100_000, 128, 12, 4e-6), (20_000, 12_000, 64, 10, 1e-6)]


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_noalloc_matches_velocity_on_two_calls(precision, bar):
    import torch
    ctx = pb.Context(0)
    for k, (ns, nt, n, P, tol) in enumerate(CALLS):
        print(f"noalloc {precision} call {k}: ns {ns} nt {nt} grid {n} P {P}", flush=True)
        x, f, y = pb.uniform_system(pb.config_seed(1) + k, ns, nt)
        p = pb.Params(alpha=10.0, eps=1e-3, grid=(n, n, n), P=P, tol=tol, precision=precision).tune(ns, nt)
        X, F, Y = dev(x), dev(f), dev(y)
        u_na = torch.full((nt, 3), np.nan, dtype=torch.float64, device="cuda")
        ctx.velocity_noalloc(p, X, F, Y, u_na)
        u_ref = ctx.velocity(p, X, F, Y)
        torch.cuda.synchronize()
        got, want = u_na.cpu().numpy(), u_ref.cpu().numpy()
        assert np.all(np.isfinite(got))
        # the same kernels; summation order differs with the atomic binning
        # order and the pair kernel's atomics (and, mixed, the FP32 spread order)
        assert relL2(got, want) <= bar, (precision, k, relL2(got, want))
        if precision == "mixed" and k == 0:
            idx = np.linspace(0, nt - 1, 128).astype(np.int64)
            ora = po.velocity_ewald(y[idx], x, f, p.L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
            assert relL2(got[idx], ora) <= 1e-6
    ctx.close()


def test_noalloc_rejects_oversize_before_any_work():
    import torch
    ctx = pb.Context(0)
    ns = nt = 1000
    x, f, y = pb.uniform_system(pb.config_seed(1), ns, nt)
    u = torch.full((nt, 3), 7.0, dtype=torch.float64, device="cuda")
    # a 256^3 FP64 grid alone needs 512 MiB (a large cutoff keeps the cell
    # grid small); a small cutoff needs more than 2^18 cell and target keys
    big = pb.Params(alpha=10.0, eps=1e-3, d=0.06, rc=0.3, grid=(256, 256, 256), P=12, precision="fp64")
    fine = pb.Params(alpha=10.0, eps=1e-3, d=0.004, rc=0.012, grid=(32, 32, 32), P=8, precision="mixed")
    for q, what in ((big, "MiB"), (fine, "2^18")):
        with pytest.raises(pb.PBError) as ei:
            ctx.velocity_noalloc(q, dev(x), dev(f), dev(y), u)
        assert ei.value.status == pb.PB_ERR_UNSUPPORTED
        assert what in str(ei.value), str(ei.value)
    torch.cuda.synchronize()
    assert bool(torch.all(u == 7.0))
    # a valid call afterwards on the same context
    p2 = pb.Params(alpha=10.0, eps=1e-3, grid=(32, 32, 32), P=8, tol=1e-5, precision="mixed").tune(ns, nt)
    ctx.velocity_noalloc(p2, dev(x), dev(f), dev(y), u)
    want = ctx.velocity(p2, dev(x), dev(f), dev(y))
    torch.cuda.synchronize()
    assert relL2(u.cpu().numpy(), want.cpu().numpy()) <= 1e-6
    ctx.close()
