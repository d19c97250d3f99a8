"""Ewald sums of the oracle against invariants and an independent formula.

* Split invariance -- the reference's own self-check (periodicBrinkman.cu:184-203):
  the total does not depend on d.
* Pure Fourier series: at d = e the real-space kernel S_e - S_d vanishes, so
  the total equals (1/tau) sum_{k != 0} phi_e(k)(I - k k^T/k^2)/(k^2+a^2) e^{ik.(y-x)} f,
  computed here directly in numpy (no oracle code).  Split invariance then ties
  every d to it.  This pins 1/tau, the k = 0 exclusion, the image sum and the
  sign/normalisation of both halves.
* Spectral k-space (fftBrinkman algorithm) equals the direct k-sum.
* The cutoff real-space sum equals the image-cube sum; matrix symmetry.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def fourier_series_total(y, x, f, L, a, e, lmax):
    L = np.asarray(L, float)
    tau = L.prod()
    ax = [np.arange(-lmax[i], lmax[i] + 1) * 2 * np.pi / L[i] for i in range(3)]
    K = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    K = K[np.any(K != 0, axis=1)]
    k2 = (K ** 2).sum(1)
    om = -np.expm1(-a * a * e * e / 4)
    phi = (1 + om * k2 / (a * a)) * np.exp(-e * e * k2 / 4)
    b = phi / (k2 + a * a) / tau
    S = np.exp(-1j * K @ x.T) @ f  # structure factors (nk, 3)
    PS = b[:, None] * (S - K * ((K * S).sum(1) / k2)[:, None])
    u = np.real(np.exp(1j * y @ K.T) @ PS)
    return u


def rand_system(seed, ns, nt, L):
    r = np.random.default_rng(seed)
    x = r.random((ns, 3)) * L
    f = r.standard_normal((ns, 3))
    y = r.random((nt, 3)) * L
    return x, f, y


@pytest.mark.parametrize("L,a,e", [((1.0, 1.3, 0.8), 2.0, 0.05), ((1.0, 1.0, 1.0), 10.0, 0.04)])
def test_total_equals_pure_fourier_series(L, a, e):
    x, f, y = rand_system(1, 4, 3, L)
    y = np.vstack([y, x[:1]])  # include a self pair
    want = fourier_series_total(y, x, f, L, a, e, lmax=(int(14 * L[0] / e / 6.3) + 1,
                                                      int(14 * L[1] / e / 6.3) + 1,
                                                      int(14 * L[2] / e / 6.3) + 1))
    for d in (0.12, 0.2):
        M = [int(np.ceil(6.5 * d / Li)) for Li in L]
        K = [int(np.ceil(14 / d * Li / (2 * np.pi))) for Li in L]
        got = po.velocity_direct(y, x, f, L, a, d, e, M, K)
        assert np.max(np.abs(got - want)) < 2e-9 * np.max(np.abs(want)), d


@pytest.mark.parametrize("e", [0.0, 0.02])
def test_split_invariance_anisotropic(e):
    L = (1.0, 1.3, 0.8)
    x, f, y = rand_system(2, 5, 4, L)
    y = np.vstack([y, x[2:3]])
    us = []
    for d in (0.1, 0.15, 0.25):
        M = [int(np.ceil(7 * d / Li)) for Li in L]
        K = [int(np.ceil(15 / d * Li / (2 * np.pi))) for Li in L]
        us.append(po.velocity_direct(y, x, f, L, 2.0, d, e, M, K))
    scale = np.max(np.abs(us[0]))
    assert np.max(np.abs(us[1] - us[0])) < 1e-11 * scale
    assert np.max(np.abs(us[2] - us[0])) < 1e-11 * scale


def test_dropping_tau_breaks_split_invariance():
    # Pins the 1/tau of fourierShells (periodicBrinkman.h:718,745): scaling the
    # k-part by tau, as the reference's kernelFourier effectively does, makes
    # the total d-dependent on an anisotropic box.
    L = (1.0, 1.3, 0.8)
    tau = np.prod(L)
    x, f, y = rand_system(3, 3, 3, L)
    tot = []
    for d in (0.12, 0.25):
        M = [int(np.ceil(7 * d / Li)) for Li in L]
        K = [int(np.ceil(15 / d * Li / (2 * np.pi))) for Li in L]
        ur = po.real_velocity_cube(y, x, f, L, 2.0, d, 0.0, M)
        uk = po.kspace_direct(y, x, f, L, 2.0, d, K)
        tot.append((ur + uk, ur + tau * uk))
    good = np.max(np.abs(tot[0][0] - tot[1][0]))
    bad = np.max(np.abs(tot[0][1] - tot[1][1]))
    assert good < 1e-11 * np.max(np.abs(tot[0][0])) and bad > 1e3 * good


def test_spectral_kspace_matches_direct():
    L = (1.0, 1.0, 1.0)
    x, f, y = rand_system(4, 40, 30, L)
    a, d = 10.0, 0.1
    K = [int(np.ceil(16 / d / (2 * np.pi)))] * 3
    want = po.kspace_direct(y, x, f, L, a, d, K)
    got = po.kspace_spectral(y, x, f, L, a, d, (48, 48, 48), (24, 24, 24), 8.7)
    assert np.linalg.norm(got - want) < 1e-12 * np.linalg.norm(want)


def test_spectral_gridding_error_decreases_with_P():
    L = (1.0, 1.0, 1.0)
    x, f, y = rand_system(5, 60, 40, L)
    a, d = 10.0, 0.1
    K = [int(np.ceil(16 / d / (2 * np.pi)))] * 3
    want = po.kspace_direct(y, x, f, L, a, d, K)
    errs = []
    for P in (8, 12, 16):
        got = po.kspace_spectral(y, x, f, L, a, d, (48, 48, 48), (P, P, P), np.sqrt(np.pi * P))
        errs.append(np.linalg.norm(got - want) / np.linalg.norm(want))
    assert errs[0] > 10 * errs[1] > 100 * errs[2] and errs[1] < 1e-6 and errs[2] < 2e-9, errs


def test_cutoff_real_space_equals_image_cube():
    L = (1.0, 0.9, 1.1)
    x, f, y = rand_system(6, 50, 30, L)
    y = np.vstack([y, x[:3]])
    a, d, e = 10.0, 0.08, 0.01
    rc = 0.55  # larger than L/2: several images of one source inside the sphere
    M = [3, 3, 3]
    cube = po.real_velocity_cube(y, x, f, L, a, d, e, M)
    cut = po.real_velocity_cutoff(y, x, f, L, a, d, e, rc)
    # the cube includes pairs beyond rc, whose kernel is below e^{-(rc/d)^2} ~ 1e-20
    assert np.max(np.abs(cube - cut)) < 1e-13 * np.max(np.abs(cube))


def test_matrix_symmetry_and_velocity_consistency():
    L = (1.0, 1.0, 1.0)
    x, f, _ = rand_system(7, 6, 0, L)
    a, d, e = 5.0, 0.2, 0.03
    M, K = [2, 2, 2], [6, 6, 6]
    A = po.ewald_matrix(x, x, L, a, d, e, M, K)
    assert np.max(np.abs(A - A.T)) < 1e-14 * np.max(np.abs(A))
    u = po.velocity_direct(x, x, f, L, a, d, e, M, K)
    assert np.max(np.abs(A @ f.reshape(-1) - u.reshape(-1))) < 1e-13 * np.max(np.abs(u))


def test_shells_semantics():
    # realShells(newM=oldM) overwrites with the home box; layers are then added;
    # fourierShells adds vF/tau over the k layers (periodicBrinkman.h:578-595,
    # :599-690, :719-821).  Home + layer(0->1) + layer(1->2) = the full cube.
    L = (1.0, 1.2, 0.9)
    x, _, y = rand_system(8, 3, 2, L)
    a, d, e = 3.0, 0.3, 0.05
    A = np.full((6, 9), 7.0, order="F")
    absA = np.zeros_like(A)
    po.realShells(A, absA, y, x, [0, 0, 0], [0, 0, 0], L, a, d, e)
    po.realShells(A, absA, y, x, [1, 1, 1], [0, 0, 0], L, a, d, e)
    po.realShells(A, absA, y, x, [2, 2, 2], [1, 1, 1], L, a, d, e)
    po.fourierShells(A, absA, y, x, [3, 3, 3], [0, 0, 0], L, a, d, e)
    po.fourierShells(A, absA, y, x, [5, 5, 5], [3, 3, 3], L, a, d, e)
    want = po.ewald_matrix(y, x, L, a, d, e, [2, 2, 2], [5, 5, 5])
    assert np.max(np.abs(A - want)) < 1e-13 * np.max(np.abs(want))
    assert np.all(absA >= np.abs(A) - 1e-12)
    B = A.copy(order="F")
    po.fourierShells(B, absA, y, x, [3, 3, 2], [3, 3, 1], L, a, d, e)  # newM == oldM in 0,1: no-op
    assert np.array_equal(A, B)


def test_spectral_kspace_matches_direct_anisotropic():
    # box and grid both uneven: per-axis spacing h_i = L_i / n_i in the window
    # offsets and wave numbers (fftBrinkman.h:101-103), the Gaussian width from
    # axis 0 (:113-114); the direct k-sum over (2K_i + 1) modes per axis is the
    # reference's fourierShells with 1/tau
    L = (1.0, 1.3, 0.8)
    x, f, y = rand_system(9, 40, 30, L)
    y = np.vstack([y, x[:2]])
    a, d = 3.0, 0.15
    K = [int(np.ceil(17 / d * Li / (2 * np.pi))) for Li in L]
    want = po.kspace_direct(y, x, f, L, a, d, K)
    got = po.kspace_spectral(y, x, f, L, a, d, (48, 64, 40), (24, 24, 24), np.sqrt(np.pi * 24))
    assert np.linalg.norm(got - want) < 1e-12 * np.linalg.norm(want)
    # a transposed box and grid give the transposed field (axis bookkeeping)
    Lt, perm = (0.8, 1.0, 1.3), [2, 0, 1]
    got_t = po.kspace_spectral(y[:, perm], x[:, perm], f[:, perm], Lt, a, d, (40, 48, 64), (24, 24, 24),
                               np.sqrt(np.pi * 24) * (0.8 / 40) / (1.0 / 48))
    assert np.linalg.norm(got_t - want[:, perm]) < 1e-11 * np.linalg.norm(want)


@pytest.mark.parametrize("shape", [(6, 10, 15), (7, 12, 9), (8, 8, 8), (5, 1, 3)])
def test_oracle_fft3d_matches_numpy(shape):
    # the oracle's own mixed-radix FFT (it replaces FFTW's fftw_plan_dft_3d,
    # fftBrinkman.h:72-76) against numpy, forward and unnormalised inverse,
    # on sizes with factors 2, 3, 5 and 7
    r = np.random.default_rng(sum(shape))
    a = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    fwd = po.fft3d(a.copy(), -1)
    assert np.max(np.abs(fwd - np.fft.fftn(a))) < 1e-13 * np.max(np.abs(fwd))
    inv = po.fft3d(a.copy(), +1)
    assert np.max(np.abs(inv - np.fft.ifftn(a) * a.size)) < 1e-13 * np.max(np.abs(inv))
