"""The C ABI without a GPU: symbols, status codes, host-only helpers."""
import ctypes
import math
import re
import subprocess

import numpy as np
import pytest

import periodicbrinkman_amd as pb


def declared_functions():
    text = open(pb.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for ref in ("periodicBrinkman", "fftBrinkman", "realShells", "fourierShells", "vR", "vF"):
        assert ref in names
    assert "pb_velocity" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = pb.lib()
    exported = subprocess.run(["nm", "-D", "--defined-only", pb.LIB_PATH], capture_output=True,
                              text=True, check=True).stdout
    syms = set(line.split()[-1] for line in exported.splitlines() if " T " in line)
    for name in declared_functions():
        assert name in syms, name
        assert hasattr(lib, name)


def test_status_strings_and_null_arguments():
    lib = pb.lib()
    assert lib.pb_status_string(0) == b"ok"
    assert lib.pb_status_string(4) == b"unsupported configuration"
    # argument checks happen before any device call
    assert lib.pb_destroy(None) == 0
    assert lib.vR(None, None, 1.0, 1.0, 0.0) == pb.PB_ERR_INVALID
    h = ctypes.c_void_p()
    assert lib.pb_create(ctypes.byref(h), -1) != 0
    assert lib.pb_velocity(None, None, 0, None, None, 0, None, None, None) == pb.PB_ERR_INVALID


def test_tuner_error_model():
    p = pb.Params(alpha=10.0, eps=1e-3, grid=(128, 128, 128), P=12, tol=1e-6).tune(100000, 100000)
    kmax = math.pi * 128
    eta = (12 / 128 / p.m) ** 2 / 2
    assert abs(p.m - math.sqrt(math.pi * 12 - 5)) < 1e-12
    # the dropped modes' weight at k_max, e^{-X}, is tol e^{7.5} (X = ln(1/tol) - 7.5)
    X = (p.d ** 2 / 4 - eta / 2) * kmax ** 2
    assert X == pytest.approx(math.log(1e6) - 7.5, rel=1e-12)
    assert math.exp(-(p.rc / p.d) ** 2) <= 1e-6 / math.e * 1.0001
    # loose tolerances keep X >= 0.4 ln(1/tol)
    r = pb.Params(alpha=10.0, eps=1e-3, grid=(128, 128, 128), P=12, tol=1e-3).tune(100000, 100000)
    X = (r.d ** 2 / 4 - eta / 2) * kmax ** 2
    assert X == pytest.approx(0.4 * math.log(1e3), rel=1e-12)
    q = pb.Params(alpha=1.0, tol=1e-8).tune(5000, 5000)
    assert all(g >= 16 and (g & (g - 1)) == 0 for g in q.grid)
    with pytest.raises(pb.PBError):
        pb.Params(alpha=-1.0).tune(10, 10)


def _xoshiro_ref(seed, n):
    # independent restatement of SplitMix64 seeding + xoshiro256** (Vigna & Blackman)
    M = (1 << 64) - 1
    z = seed
    s = []
    for _ in range(4):
        z = (z + 0x9E3779B97F4A7C15) & M
        t = z
        t = ((t ^ (t >> 30)) * 0xBF58476D1CE4E5B9) & M
        t = ((t ^ (t >> 27)) * 0x94D049BB133111EB) & M
        s.append(t ^ (t >> 31))
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & M
    out = []
    for _ in range(n):
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        out.append((r >> 11) * 2.0 ** -53)
    return out


def test_generator_matches_independent_restatement():
    seed = pb.config_seed(1)
    got = pb.Rng(seed).uniform(16)
    assert np.array_equal(got, np.array(_xoshiro_ref(seed, 16)))
    u = _xoshiro_ref(seed, 4)
    nrm = pb.Rng(seed).normal(4)
    r = math.sqrt(-2 * math.log(1 - u[0]))
    assert nrm[0] == pytest.approx(r * math.cos(2 * math.pi * u[1]), rel=1e-15)
    assert nrm[1] == pytest.approx(r * math.sin(2 * math.pi * u[1]), rel=1e-15)
    x, f, y = pb.uniform_system(seed, 10, 5)
    assert x.shape == (10, 3) and f.shape == (10, 3) and y.shape == (5, 3)
    assert np.array_equal(x.reshape(-1), np.array(_xoshiro_ref(seed, 30)))


def test_batched_workload_generator():
    # the batched workload: sizes, boxes and alphas from one stream in the
    # stated ranges, each system's points from its own stream, reproducible
    seed = pb.config_seed(3)
    systems, x, f, off = pb.batched_systems(seed, nsys=40)
    assert len(systems) == 40 and off[0] == 0 and off[-1] == x.shape[0] == f.shape[0]
    u = np.array(_xoshiro_ref(seed, 3 * 40)).reshape(40, 3)
    for i, s in enumerate(systems):
        assert off[i + 1] - off[i] == s["n"]
        assert s["n"] == int(np.floor(np.exp(np.log(100) + u[i, 0] * (np.log(20000) - np.log(100)))))
        assert 100 <= s["n"] < 20000 and 0.5 <= s["L"] < 2.0 and 1.0 <= s["alpha"] < 50.0
        assert s["eps"] == pytest.approx(1e-3 * s["L"], rel=1e-15)
        xi = x[off[i]:off[i + 1]]
        assert np.all((xi >= 0) & (xi < s["L"]))
    x0, _, _ = pb.uniform_system(seed + (1 << 32), systems[0]["n"], 0, (systems[0]["L"],) * 3)
    assert np.array_equal(x[:off[1]], x0)
    _, x2, _, _ = pb.batched_systems(seed, nsys=40)
    assert np.array_equal(x, x2)
