"""The tuner's error model against the FP64 oracle (host only; no GPU).

pb_tune and pb_autotune choose the split d, the cutoff rc and the window
shape m from a calibrated model (window m^2 = pi P - 5; k-space criterion
X = (d^2/4 - eta/2) k_max^2 = max(ln(1/tol) - 7.5, 0.4 ln(1/tol)); rc = d
sqrt(ln(1/tol) + 1)).  Here the velocity the oracle computes with those
parameters is compared with a converged Ewald sum (window P 24, tolerance
1e-13, hence a wider split, a longer cutoff and a negligible window error):
the relative L2 error must not exceed the tolerance asked for.  A mistake in
the model (a dropped term in d, a wrong sign on eta, the old m = sqrt(pi P)
paired with the relaxed criterion) shows up as an error above tol here.
"""
import numpy as np
import pytest

import periodicbrinkman_amd as pb
from oracle import pyoracle as po


def relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle(y, x, f, p):
    return po.velocity_ewald(y, x, f, p.L, p.alpha, p.d, p.eps, p.rc, p.grid, (p.P,) * 3, p.m)


@pytest.fixture(scope="module", autouse=True)
def threads():
    po.set_threads(0)
    yield


def converged(y, x, f, L, alpha, eps, grid):
    pc = pb.Params(L=L, alpha=alpha, eps=eps, grid=grid, P=24, tol=1e-13).tune(len(x), len(y))
    return oracle(y, x, f, pc)


@pytest.mark.parametrize("tol", [1e-5, 1e-6, 1e-8, 1e-10])
def test_tuned_uniform_system_meets_tol(tol):
    n = 20_000
    x, f, y = pb.uniform_system(pb.config_seed(7), n, 256)
    ref = converged(y, x, f, (1.0,) * 3, 10.0, 1e-3, (64,) * 3)
    # the smallest window whose truncation e^{-m^2/2} meets tol (as pb_autotune requires)
    P = next(q for q in (4, 6, 8, 10, 12, 16, 20, 24) if np.exp(-(np.pi * q - 5) / 2) <= tol)
    p = pb.Params(alpha=10.0, eps=1e-3, grid=(64, 64, 64), P=P, tol=tol).tune(n, 256)
    assert relL2(oracle(y, x, f, p), ref) <= tol


@pytest.mark.parametrize("L,alpha,tol", [(1.5, 30.0, 3e-6), (0.6, 2.0, 3e-6), (2.0, 45.0, 1e-7)])
def test_autotuned_system_meets_tol(L, alpha, tol):
    n = 4000
    x, f, _ = pb.uniform_system(pb.config_seed(8), n, 0, (L, L, L))
    p = pb.Params(L=(L,) * 3, alpha=alpha, eps=1e-3 * L, tol=tol)
    p.autotune(n, n)
    k = np.linspace(0, n - 1, 128).astype(np.int64)
    ref = converged(x[k], x, f, (L,) * 3, alpha, 1e-3 * L, p.grid)
    assert relL2(oracle(x[k], x, f, p), ref) <= tol


def test_tuned_clustered_system_meets_tol():
    n = 20_000
    x, f = pb.clustered_system(pb.config_seed(9), n, 8, 0.02)
    k = np.linspace(0, n - 1, 128).astype(np.int64)
    ref = converged(x[k], x, f, (1.0,) * 3, 25.0, 5e-4, (64,) * 3)
    p = pb.Params(alpha=25.0, eps=5e-4, grid=(64, 64, 64), P=12, tol=1e-6).tune(n, n)
    assert relL2(oracle(x[k], x, f, p), ref) <= 1e-6
