"""The reference driver's split sweep (pb_split_sweep and the pb_driver
program) against the FP64 oracle's shell sums."""
import os
import subprocess

import numpy as np
import pytest

import periodicbrinkman_amd as pb
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle_sweep(p, x):
    L = (p["Lx"], p["Ly"], p["Lz"])
    x0 = np.array([[0.5, 0.5, 0.5]])
    nt = x.shape[0]

    def matrix(d):
        A = np.zeros((3 * nt, 3), order="F")
        B = np.zeros((3 * nt, 3), order="F")
        M = [p["M"]] * 3
        po.realShells(A, B, x, x0, [0, 0, 0], [0, 0, 0], L, p["alpha"], d, p["e"])
        po.realShells(A, B, x, x0, M, [0, 0, 0], L, p["alpha"], d, p["e"])
        po.fourierShells(A, B, x, x0, M, [0, 0, 0], L, p["alpha"], d, p["e"])
        return A

    A0 = matrix(p["d"])
    dval = np.array([(0.1 + 0.95 * i) * p["d"] for i in range(8)])
    err = np.array([np.max(np.abs(matrix(dv) - A0)) for dv in dval])
    return dval, err


def test_split_sweep_matches_oracle():
    p = dict(pb.driver_defaults(), Lx=1.0, Ly=1.3, Lz=0.8, d=0.3, e=0.02, alpha=2.0, M=2, boxSize=3)
    rng = np.random.default_rng(2)
    x = rng.random((27, 3)) * (p["Lx"], p["Ly"], p["Lz"])
    dval, err = pb.split_sweep(p, x)
    wd, we = _oracle_sweep(p, x)
    assert np.allclose(dval, wd, rtol=0, atol=1e-15)
    # the differences are truncation errors of the shells (small where the
    # split is well resolved); they must agree with the oracle's to rounding
    scale = np.max(np.abs(we))
    assert np.max(np.abs(err - we)) <= 1e-12 * scale + 1e-14


def test_driver_program_writes_reference_log(tmp_path):
    exe = os.path.join(ROOT, "periodicbrinkman_amd", "bin", "pb_driver")
    (tmp_path / "param.txt").write_text("Lx: 1\nLy: 1\nLz: 1\nd: 0.3\ne: 0.01\nM: 2\nalpha: 3\nboxSize: 2\n")
    r = subprocess.run([exe, "param.txt"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "log.txt").read_text().splitlines()
    assert lines[0] == "For Lx = 1, Ly = 1, Lz = 1, and a = 3, we have: "
    assert len(lines) == 9
    vals = [tuple(float(v) for v in ln.split(" : ")) for ln in lines[1:]]
    assert np.allclose([v[0] for v in vals], [(0.1 + 0.95 * i) * 0.3 for i in range(8)], rtol=1e-10)
    assert all(v[1] == 0 for v in vals) and all(v[2] >= 0 for v in vals)
    assert "Done!" in r.stdout
