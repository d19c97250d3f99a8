"""ctypes binding of the FP64 CPU oracle (oracle/pb_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker.  Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py import this
module; the product package never does.  Arrays follow the reference's
layout: points are 3 x n column-major, i.e. numpy shape (n, 3) C-order.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc only; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    dd = ctypes.c_double
    L.pbo_experf.restype = dd
    L.pbo_experf.argtypes = [dd, dd, dd]
    for name in ("pbo_singH_literal",):
        getattr(L, name).argtypes = [dd, dd, dd, _D, _D]
    for name in ("pbo_regH_literal", "pbo_diffH_literal"):
        getattr(L, name).argtypes = [dd, dd, dd, dd, _D, _D]
    L.pbo_singH.argtypes = [dd, dd, _D, _D]
    L.pbo_regH.argtypes = [dd, dd, dd, _D, _D]
    L.pbo_diffH.argtypes = [dd, dd, dd, _D, _D]
    L.pbo_regH_series_coeffs.argtypes = [dd, dd, ctypes.c_int, _D, _D]
    L.pbo_vR.argtypes = [_D, _D, dd, dd, dd]
    L.pbo_vF.argtypes = [_D, _D, dd, dd, _D]
    L.pbo_blob_hat.restype = dd
    L.pbo_blob_hat.argtypes = [dd, dd, dd]
    shell = [_D, _D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _I, _I, _D, dd, dd, dd]
    L.pbo_realShells.argtypes = shell
    L.pbo_fourierShells.argtypes = shell
    L.pbo_ewald_matrix.argtypes = [_D, ctypes.c_int, _D, ctypes.c_int, _D, ctypes.c_int, _D,
                                   dd, dd, dd, _I, _I]
    L.pbo_velocity_direct.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, dd, dd, dd,
                                      _I, _I]
    L.pbo_real_velocity_cube.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, dd, dd,
                                         dd, _I]
    L.pbo_kspace_direct.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, dd, dd, _I]
    L.pbo_real_velocity_cutoff.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, dd,
                                           dd, dd, dd]
    L.pbo_kspace_spectral.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, dd, dd,
                                      _I, _I, dd]
    L.pbo_fftBrinkman.restype = ctypes.c_int
    L.pbo_fftBrinkman.argtypes = [_D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, _I, _I,
                                  ctypes.c_int, ctypes.c_int, dd, dd, dd, dd]
    L.pbo_fft3d.argtypes = [_D, _I, ctypes.c_int]
    L.pbo_set_threads.restype = ctypes.c_int
    L.pbo_set_threads.argtypes = [ctypes.c_int]


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    return a, a.ctypes.data_as(_I)


def _pts(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
    return a


def _vec3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def set_threads(n):
    return lib().pbo_set_threads(int(n))


def experf(x, z, c=2.0):
    return lib().pbo_experf(float(x), float(z), float(c))


def _two(fn, *args):
    h1 = ctypes.c_double()
    h2 = ctypes.c_double()
    fn(*[float(a) for a in args], ctypes.byref(h1), ctypes.byref(h2))
    return h1.value, h2.value


def singH(r2, a):
    return _two(lib().pbo_singH, r2, a)


def regH(r2, a, d):
    return _two(lib().pbo_regH, r2, a, d)


def diffH(r2, a, d):
    return _two(lib().pbo_diffH, r2, a, d)


def singH_literal(r2, a, eps=0.0):
    return _two(lib().pbo_singH_literal, r2, a, eps)


def regH_literal(r2, a, d, eps=0.0):
    return _two(lib().pbo_regH_literal, r2, a, d, eps)


def diffH_literal(r2, a, d, eps=0.0):
    return _two(lib().pbo_diffH_literal, r2, a, d, eps)


def regH_series_coeffs(a, d, n=12):
    c1 = np.zeros(n)
    c2 = np.zeros(n)
    lib().pbo_regH_series_coeffs(float(a), float(d), n, _dp(c1), _dp(c2))
    return c1, c2


def vR(xh, a, d, e):
    A = np.zeros(9)
    x = _vec3(xh)
    lib().pbo_vR(_dp(A), _dp(x), float(a), float(d), float(e))
    return A.reshape(3, 3).T.copy()  # column-major 3x3 -> numpy [row, col]


def vF(xh, a, d, k):
    A = np.zeros(9)
    x = _vec3(xh)
    kk = _vec3(k)
    lib().pbo_vF(_dp(A), _dp(x), float(a), float(d), _dp(kk))
    return A.reshape(3, 3).T.copy()


def blob_hat(k2, a, d):
    return lib().pbo_blob_hat(float(k2), float(a), float(d))


def _shell(fn, A, absA, x, x0, newM, oldM, L, a, d, e):
    x = _pts(x)
    x0 = _pts(x0)
    L3 = _vec3(L)
    nm, nmp = _ip(newM)
    om, omp = _ip(oldM)
    assert A.flags.f_contiguous and absA.flags.f_contiguous
    fn(_dp(A), _dp(absA), A.shape[0], _dp(x), x.shape[0], _dp(x0), x0.shape[0], nmp, omp,
       _dp(L3), float(a), float(d), float(e))


def realShells(A, absA, x, x0, newM, oldM, L, a, d, e):
    """In-place on Fortran-ordered A, absA of shape (3 nt, 3 ns)."""
    _shell(lib().pbo_realShells, A, absA, x, x0, newM, oldM, L, a, d, e)


def fourierShells(A, absA, x, x0, newM, oldM, L, a, d, e):
    _shell(lib().pbo_fourierShells, A, absA, x, x0, newM, oldM, L, a, d, e)


def ewald_matrix(y, x, L, a, d, e, M, K):
    y = _pts(y)
    x = _pts(x)
    A = np.zeros((3 * y.shape[0], 3 * x.shape[0]), order="F")
    Mv, Mp = _ip(M)
    Kv, Kp = _ip(K)
    lib().pbo_ewald_matrix(_dp(A), A.shape[0], _dp(y), y.shape[0], _dp(x), x.shape[0],
                           _dp(_vec3(L)), float(a), float(d), float(e), Mp, Kp)
    return A


def velocity_direct(y, x, f, L, a, d, e, M, K):
    y, x, f = _pts(y), _pts(x), _pts(f)
    u = np.zeros_like(y)
    Mv, Mp = _ip(M)
    Kv, Kp = _ip(K)
    lib().pbo_velocity_direct(_dp(u), _dp(y), y.shape[0], _dp(x), _dp(f), x.shape[0],
                              _dp(_vec3(L)), float(a), float(d), float(e), Mp, Kp)
    return u


def real_velocity_cube(y, x, f, L, a, d, e, M):
    y, x, f = _pts(y), _pts(x), _pts(f)
    u = np.zeros_like(y)
    Mv, Mp = _ip(M)
    lib().pbo_real_velocity_cube(_dp(u), _dp(y), y.shape[0], _dp(x), _dp(f), x.shape[0],
                                 _dp(_vec3(L)), float(a), float(d), float(e), Mp)
    return u


def kspace_direct(y, x, f, L, a, d, K):
    y, x, f = _pts(y), _pts(x), _pts(f)
    u = np.zeros_like(y)
    Kv, Kp = _ip(K)
    lib().pbo_kspace_direct(_dp(u), _dp(y), y.shape[0], _dp(x), _dp(f), x.shape[0],
                            _dp(_vec3(L)), float(a), float(d), Kp)
    return u


def real_velocity_cutoff(y, x, f, L, a, d, e, rc):
    y, x, f = _pts(y), _pts(x), _pts(f)
    u = np.zeros_like(y)
    lib().pbo_real_velocity_cutoff(_dp(u), _dp(y), y.shape[0], _dp(x), _dp(f), x.shape[0],
                                   _dp(_vec3(L)), float(a), float(d), float(e), float(rc))
    return u


def kspace_spectral(y, x, f, L, a, d, grid, P, m):
    y, x, f = _pts(y), _pts(x), _pts(f)
    u = np.zeros_like(y)
    gv, gp = _ip(grid)
    Pv, Pp = _ip(P)
    lib().pbo_kspace_spectral(_dp(u), _dp(y), y.shape[0], _dp(x), _dp(f), x.shape[0],
                              _dp(_vec3(L)), float(a), float(d), gp, Pp, float(m))
    return u


def fftBrinkman(y, f, xn, L, M, P, msd, C, e, tol, d, alpha=1.0):
    y, f, xn = _pts(y), _pts(f), _pts(xn)
    U = np.zeros_like(y)
    Mv, Mp = _ip(M)
    Pv, Pp = _ip(P)
    rc = lib().pbo_fftBrinkman(_dp(U), _dp(y), y.shape[0], _dp(f), _dp(xn), xn.shape[0],
                               _dp(_vec3(L)), Mp, Pp, int(msd), int(C), float(e), float(tol),
                               float(d), float(alpha))
    assert rc == 0
    return U


def fft3d(data, sign):
    """In-place complex 3D FFT of a C-contiguous complex128 array."""
    assert data.dtype == np.complex128 and data.flags.c_contiguous and data.ndim == 3
    n, npp = _ip(data.shape)
    lib().pbo_fft3d(data.ctypes.data_as(_D), npp, int(sign))
    return data


def velocity_ewald(y, x, f, L, a, d, e, rc, grid, P, m):
    """Real space with cutoff rc plus spectral k-space: the full evaluator."""
    return real_velocity_cutoff(y, x, f, L, a, d, e, rc) + kspace_spectral(y, x, f, L, a, d,
                                                                            grid, P, m)
