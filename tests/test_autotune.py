"""pb_autotune: the chosen parameters meet the error model's bounds, move the
right way with the tolerance and the box, and reject invalid input (host
only; no GPU)."""
import math

import pytest

import periodicbrinkman_amd as pb


def _tuned(tol, L=(1.0, 1.0, 1.0), ns=100_000, nt=100_000, alpha=10.0, precision="mixed"):
    p = pb.Params(L=L, alpha=alpha, eps=1e-3, precision=precision, tol=tol)
    rep = p.autotune(ns, nt)
    return p, rep


@pytest.mark.parametrize("tol", [1e-4, 1e-6, 1e-8, 1e-10])
def test_bounds_met(tol):
    p, rep = _tuned(tol)
    assert rep["candidates"] > 0
    assert max(rep["err_real"], rep["err_kspace"], rep["err_window"]) <= tol * (1 + 1e-12)
    # the error model's own relations (window m = sqrt(pi P - 5), rc = d sqrt(ln(1/tol) + 1))
    assert math.isclose(p.m, math.sqrt(math.pi * p.P - 5), rel_tol=1e-12)
    assert math.isclose(p.rc, p.d * math.sqrt(math.log(1 / tol) + 1), rel_tol=1e-12)
    assert all(g & (g - 1) == 0 and 4 <= g <= 512 for g in p.grid)
    assert p.P in (4, 6, 8, 10, 12, 16, 20, 24) and p.P <= min(p.grid)
    assert rep["t_total_ms"] == pytest.approx(rep["t_real_ms"] + rep["t_kspace_ms"] + rep["t_bin_ms"])


def test_tighter_tolerance_costs_more():
    prev = None
    for tol in (1e-4, 1e-6, 1e-8, 1e-10):
        p, rep = _tuned(tol)
        if prev is not None:
            assert rep["t_total_ms"] > prev[1]["t_total_ms"]
            assert p.P >= prev[0].P
        prev = (p, rep)


def test_predicted_pairs_follow_the_cutoff():
    p, rep = _tuned(1e-6)
    want = 1e5 * 1e5 * 4 / 3 * math.pi * p.rc ** 3
    assert rep["pairs"] == pytest.approx(want, rel=1e-12)


def test_anisotropic_box_grid_follows_lengths():
    p, _ = _tuned(1e-6, L=(1.0, 2.0, 0.5))
    h = [L / g for L, g in zip(p.L, p.grid)]
    assert max(h) <= min(h) * 2 + 1e-12  # one spacing, each axis rounded to a power of two
    assert p.grid[1] >= p.grid[0] >= p.grid[2]


def test_more_particles_move_work_to_k_space():
    # denser systems favour a finer grid (smaller d, fewer pairs per particle)
    p1, _ = _tuned(1e-6, ns=10_000, nt=10_000)
    p2, _ = _tuned(1e-6, ns=1_000_000, nt=1_000_000)
    assert p2.grid[0] >= p1.grid[0] and p2.d <= p1.d


def test_invalid_input_is_a_status():
    with pytest.raises(pb.PBError) as ei:
        pb.Params(alpha=-1.0).autotune(10, 10)
    assert ei.value.status == pb.PB_ERR_INVALID
