"""Edge regimes of the Ewald evaluator against the FP64 oracle: strongly
anisotropic boxes, small alpha L (the 1/alpha^2 cancellation of the closed
forms grows), the singular Brinkmanlet (eps = 0) with targets on the
sources, and clustered sources (cell-list load imbalance)."""
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


@pytest.fixture(scope="module")
def ctx():
    c = pb.Context(0)
    yield c
    c.close()


def _check(ctx, x, f, y, p, tol_fp64, tol_mixed=1e-6):
    want = po.velocity_ewald(y, x, f, p.L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
    got = ctx.velocity(p, dev(x), dev(f), dev(y)).cpu().numpy()
    assert relL2(got, want) < tol_fp64, relL2(got, want)
    gm = ctx.velocity(p.copy(precision="mixed"), dev(x), dev(f), dev(y)).cpu().numpy()
    assert relL2(gm, want) < tol_mixed, relL2(gm, want)


def test_strongly_anisotropic_box(ctx):
    L = (1.0, 2.5, 0.4)
    x, f, y = pb.uniform_system(pb.config_seed(40), 2500, 1500, L)
    p = pb.Params(L=L, alpha=8.0, eps=2e-3, grid=(32, 64, 16), P=12, tol=1e-8).tune(2500, 1500)
    _check(ctx, x, f, y, p, 1e-11)


@pytest.mark.parametrize("alpha,tol", [(0.05, 2e-9), (0.5, 1e-10)])
def test_small_alpha_L(ctx, alpha, tol):
    # at alpha d ~ 5e-3 the oracle's FP64 closed forms lose ~(alpha d)^-2
    # rounding units to cancellation (its golden-vector tolerance there is
    # 1e-9); the device evaluates that range from a long-double table
    L = (1.0, 1.0, 1.0)
    x, f, y = pb.uniform_system(pb.config_seed(41), 2000, 1500, L)
    p = pb.Params(L=L, alpha=alpha, eps=1e-3, grid=(32, 32, 32), P=12, tol=1e-8).tune(2000, 1500)
    _check(ctx, x, f, y, p, tol)


def test_singular_brinkmanlet_on_sources(ctx):
    L = (1.0, 1.0, 1.0)
    x, f, _ = pb.uniform_system(pb.config_seed(42), 2000, 0, L)
    p = pb.Params(L=L, alpha=10.0, eps=0.0, grid=(32, 32, 32), P=12, tol=1e-8).tune(2000, 2000)
    _check(ctx, x, f, x.copy(), p, 1e-11)


def test_clustered_sources(ctx):
    # sources in 8 Gaussian clusters (sigma 0.02, wrapped), targets on the sources
    L = (1.0, 1.0, 1.0)
    x, f = pb.clustered_system(pb.config_seed(43), 4000, 8, 0.02, L)
    p = pb.Params(L=L, alpha=25.0, eps=5e-4, grid=(64, 64, 64), P=12, tol=1e-8).tune(4000, 4000)
    _check(ctx, x, f, x.copy(), p, 1e-11)


def test_clustered_full_size_split_invariance(ctx):
    # 1e6 sources in 64 Gaussian clusters (sigma 0.02), targets on the sources,
    # alpha 25, eps 5e-4, 256^3 grid, P 16: too large for the oracle, so two
    # different splits d must give the same field (the reference's own check,
    # periodicBrinkman.cu:184-203); mixed precision meets 1e-6 against FP64
    import torch
    n = 1_000_000
    x, f = pb.clustered_system(pb.config_seed(2), n, 64, 0.02)
    X, F = dev(x), dev(f)
    us, ds = [], []
    for tol in (1e-8, 1e-10):
        p = pb.Params(alpha=25.0, eps=5e-4, grid=(256, 256, 256), P=16, tol=tol).tune(n, n)
        ds.append(p.d)
        us.append(ctx.velocity(p, X, F, X).cpu().numpy())
    torch.cuda.synchronize()
    assert ds[1] > 1.05 * ds[0]
    assert relL2(us[0], us[1]) < 1e-7
    pm = pb.Params(alpha=25.0, eps=5e-4, grid=(256, 256, 256), P=16, tol=1e-10, precision="mixed").tune(n, n)
    um = ctx.velocity(pm, X, F, X).cpu().numpy()
    assert relL2(um, us[1]) < 1e-6


def test_one_context_through_changing_cell_grids(ctx):
    # one context, cell grids that grow, shrink and grow again (the binning's
    # count arrays are cleared by each call's scan and refilled only when
    # they are reallocated): every call against the oracle
    L = (1.0, 1.0, 1.0)
    x, f, y = pb.uniform_system(pb.config_seed(44), 1500, 1000, L)
    for rc in (0.45, 0.2, 0.45, 0.12):
        p = pb.Params(L=L, alpha=10.0, eps=1e-3, d=rc / 4.5, rc=rc, grid=(32, 32, 32), P=12,
                      m=float(np.sqrt(np.pi * 12)))
        want = po.velocity_ewald(y, x, f, p.L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)
        got = ctx.velocity(p, dev(x), dev(f), dev(y)).cpu().numpy()
        assert relL2(got, want) < 1e-10, (rc, relL2(got, want))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_dense_clusters_split_spreading_regions(ctx, precision, tol):
    # 3e5 sources in 8 clusters of sigma 0.004 (a quarter of a grid spacing at
    # 64^3) with P 16: 3e5 x 16^3 window points take the balanced spreading,
    # and each cluster's region holds ~4e4 candidates, so it is split into
    # chunks whose partial sums are added in order.  The k-space part against
    # the oracle's spectral sum on 256 targets (all sources), and the full
    # velocity on the same targets.
    import torch
    n = 300_000
    L = (1.0, 1.0, 1.0)
    x, f = pb.clustered_system(pb.config_seed(45), n, 8, 0.004, L)
    y = np.vstack([x[:128], pb.uniform_system(pb.config_seed(46), 1, 128, L)[2]])
    p = pb.Params(L=L, alpha=10.0, eps=1e-3, d=0.05, rc=0.02, grid=(64, 64, 64), P=16,
                  m=float(np.sqrt(np.pi * 16)), precision=precision)
    want = po.kspace_spectral(y, x, f, p.L, p.alpha, p.d, p.grid, (p.P,) * 3, p.m)
    got = ctx.velocity_kspace(p, dev(x), dev(f), dev(y)).cpu().numpy()
    torch.cuda.synchronize()
    assert relL2(got, want) < tol, relL2(got, want)
    want_all = want + po.real_velocity_cutoff(y, x, f, p.L, p.alpha, p.d, p.eps, p.rc)
    got_all = ctx.velocity(p, dev(x), dev(f), dev(y)).cpu().numpy()
    assert relL2(got_all, want_all) < tol, relL2(got_all, want_all)
