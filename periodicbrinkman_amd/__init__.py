"""periodicbrinkman_amd -- MI355X (gfx950) Ewald evaluator for periodic Brinkman flow.

Python host side of the C ABI in include/periodicbrinkman.h (library
periodicbrinkman_amd/lib/libpbamd.so).  The functions mirror the reference's
interface (hoangngan82/periodicBrinkman, cuda/periodicBrinkman.h and
cuda/fftBrinkman.h): same names, argument meaning and layout (points are
3 x n column-major, i.e. arrays of shape (n, 3)), with status codes raised as
PBError.  All arithmetic runs in the HIP library; there is no CPU fallback:
if the library is missing or no GPU is present, calls raise.

PyTorch is used only to hold device memory and streams.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpbamd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "periodicbrinkman.h")

PB_OK, PB_ERR_INVALID, PB_ERR_HIP, PB_ERR_NOMEM, PB_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
KERNEL_TABLE, KERNEL_DIRECT, KERNEL_TABLE_FAR32 = 0, 1, 2
_KERNELS = {"table": KERNEL_TABLE, "direct": KERNEL_DIRECT, "table_far32": KERNEL_TABLE_FAR32}

_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U64 = ctypes.POINTER(ctypes.c_ulonglong)
_VP = ctypes.c_void_p


class PBError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"periodicbrinkman status {status}: {msg}")


class _Params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double * 3), ("alpha", ctypes.c_double), ("eps", ctypes.c_double),
                ("d", ctypes.c_double), ("rc", ctypes.c_double), ("grid", ctypes.c_int * 3),
                ("P", ctypes.c_int), ("m", ctypes.c_double), ("precision", ctypes.c_int),
                ("tol", ctypes.c_double), ("kernel", ctypes.c_int)]


class _DriverParams(ctypes.Structure):
    _fields_ = [("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
                ("d", ctypes.c_double), ("e", ctypes.c_double), ("alpha", ctypes.c_double),
                ("M", ctypes.c_int), ("boxSize", ctypes.c_int), ("toPrint", ctypes.c_int)]


class _TuneReport(ctypes.Structure):
    _fields_ = [("err_real", ctypes.c_double), ("err_kspace", ctypes.c_double),
                ("err_window", ctypes.c_double), ("pairs", ctypes.c_double),
                ("t_real_ms", ctypes.c_double), ("t_kspace_ms", ctypes.c_double),
                ("t_bin_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double),
                ("candidates", ctypes.c_int)]


def build(jobs=8):
    """Compile libpbamd.so for gfx950 in-tree (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", os.path.join(_HERE, "csrc")], check=True)


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run periodicbrinkman_amd.build() "
                               "(hipcc --offload-arch=gfx950) first")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    dd, ii = ctypes.c_double, ctypes.c_int
    L.periodicBrinkman.argtypes = [_VP, ii, _VP, ii, _VP, ii, dd, dd, dd, ii, ii, _D, ii]
    L.fftBrinkman.argtypes = [_D, _D, ii, _D, _D, ii, _D, _I, _I, ii, ii, dd, dd, dd, dd]
    shell = [_D, _D, ii, _D, ii, _D, ii, _I, _I, _D, dd, dd, dd]
    L.realShells.argtypes = shell
    L.fourierShells.argtypes = shell
    L.vR.argtypes = [_D, _D, dd, dd, dd]
    L.vF.argtypes = [_D, _D, dd, dd, _D]
    L.pb_vR_batch.argtypes = [_VP, _VP, ii, dd, dd, dd, _VP]
    L.pb_vF_batch.argtypes = [_VP, _VP, _VP, ii, dd, dd, _VP]
    L.pb_params_default.argtypes = [ctypes.POINTER(_Params)]
    L.pb_params_default.restype = None
    L.pb_tune.argtypes = [ctypes.POINTER(_Params), ii, ii]
    L.pb_driver_defaults.argtypes = [ctypes.POINTER(_DriverParams)]
    L.pb_driver_defaults.restype = None
    L.pb_read_param_file.argtypes = [ctypes.c_char_p, ctypes.POINTER(_DriverParams)]
    L.pb_split_sweep.argtypes = [ctypes.POINTER(_DriverParams), _D, ii, _D, _D]
    L.pb_autotune.argtypes = [ctypes.POINTER(_Params), ii, ii, ctypes.POINTER(_TuneReport)]
    L.pb_create.argtypes = [ctypes.POINTER(_VP), ii]
    L.pb_destroy.argtypes = [_VP]
    L.pb_last_error.argtypes = [_VP]
    L.pb_last_error.restype = ctypes.c_char_p
    L.pb_status_string.argtypes = [ii]
    L.pb_status_string.restype = ctypes.c_char_p
    vel = [_VP, ctypes.POINTER(_Params), ii, _VP, _VP, ii, _VP, _VP, _VP]
    L.pb_velocity.argtypes = vel
    L.pb_velocity_real.argtypes = vel
    L.pb_velocity_kspace.argtypes = vel
    L.pb_velocity_noalloc.argtypes = vel
    L.pb_velocity_batched.argtypes = [_VP, ii, ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_int), _VP, _VP,
                                      ctypes.POINTER(ctypes.c_int), _VP, _VP, _VP]
    L.pb_velocity_multirhs.argtypes = [_VP, ctypes.POINTER(_Params), ii, _VP, ii, _VP, ii, _VP, _VP, _VP]
    L.pb_last_pair_stats.argtypes = [_VP, ctypes.POINTER(ctypes.c_longlong)]
    L.pb_last_pair_counts.argtypes = [_VP, ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_longlong)]
    L.pb_set_timing.argtypes = [_VP, ii]
    L.pb_last_timings.argtypes = [_VP, ctypes.POINTER(ctypes.c_float), ii]
    L.pb_rng_seed.argtypes = [_U64, ctypes.c_ulonglong]
    L.pb_rng_seed.restype = None
    L.pb_rng_uniform.argtypes = [_U64, ctypes.c_long, _D, dd]
    L.pb_rng_uniform.restype = None
    L.pb_rng_normal.argtypes = [_U64, ctypes.c_long, _D]
    L.pb_rng_normal.restype = None


def _check(status, ctx=None):
    if status != PB_OK:
        msg = lib().pb_status_string(status).decode()
        if ctx is not None and ctx.handle:
            detail = lib().pb_last_error(ctx.handle).decode()
            if detail:
                msg += ": " + detail
        raise PBError(status, msg)


def _dp(a):
    return a.ctypes.data_as(_D)


def _iv(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))
    return a


def _hpts(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))


def _dptr(t):
    """Device pointer of a contiguous float64 torch CUDA tensor."""
    import torch
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise PBError(PB_ERR_INVALID, "expected a contiguous float64 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


def _shape(t, shape, name):
    """Raise PBError(PB_ERR_INVALID) unless tensor t has the given shape (a
    mismatch would make the kernels read or write outside the arrays)."""
    if tuple(t.shape) != tuple(int(v) for v in shape):
        raise PBError(PB_ERR_INVALID, f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# ---- parameters ---------------------------------------------------------------
class Params:
    """Evaluator parameters (pb_params).  precision: 'fp64' or 'mixed' (FP32
    k-space); kernel: 'table' (real-space radial functions from the FP64
    piecewise-polynomial table), 'direct' (closed forms per pair; FP32 beyond
    max(3d, 8e) in mixed precision) or 'table_far32' (table; FP32 copy beyond
    max(3d, 8e) in mixed precision)."""

    def __init__(self, L=(1.0, 1.0, 1.0), alpha=1.0, eps=0.0, d=0.0, rc=0.0, grid=(0, 0, 0), P=12,
                 m=0.0, precision="fp64", tol=1e-6, kernel="table"):
        self.L = tuple(float(v) for v in L)
        self.alpha, self.eps, self.d, self.rc = float(alpha), float(eps), float(d), float(rc)
        self.grid = tuple(int(g) for g in grid)
        self.P, self.m, self.tol = int(P), float(m), float(tol)
        self.precision = precision
        self.kernel = kernel

    def _struct(self):
        p = _Params()
        for i in range(3):
            p.L[i] = self.L[i]
            p.grid[i] = self.grid[i]
        p.alpha, p.eps, p.d, p.rc = self.alpha, self.eps, self.d, self.rc
        p.P, p.m, p.tol = self.P, self.m, self.tol
        p.precision = PREC_MIXED if self.precision == "mixed" else PREC_FP64
        if self.kernel not in _KERNELS:
            raise PBError(PB_ERR_INVALID, f"kernel must be one of {sorted(_KERNELS)}")
        p.kernel = _KERNELS[self.kernel]
        return p

    def tune(self, ns, nt):
        """Fill d, rc, grid and m (those unset) from the error model (pb_tune)."""
        p = self._struct()
        _check(lib().pb_tune(ctypes.byref(p), int(ns), int(nt)))
        self.d, self.rc, self.m = p.d, p.rc, p.m
        self.grid = tuple(p.grid[i] for i in range(3))
        return self

    def autotune(self, ns, nt):
        """Choose grid, P, m, d and rc for tol by the error model and the MI355X
        cost model (pb_autotune); returns the report as a dict."""
        p = self._struct()
        r = _TuneReport()
        _check(lib().pb_autotune(ctypes.byref(p), int(ns), int(nt), ctypes.byref(r)))
        self.d, self.rc, self.m, self.P = p.d, p.rc, p.m, p.P
        self.grid = tuple(p.grid[i] for i in range(3))
        return {k: getattr(r, k) for k, _ in _TuneReport._fields_}

    def copy(self, **kw):
        q = Params(self.L, self.alpha, self.eps, self.d, self.rc, self.grid, self.P, self.m,
                   self.precision, self.tol, self.kernel)
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def __repr__(self):
        return (f"Params(L={self.L}, alpha={self.alpha}, eps={self.eps}, d={self.d:.6g}, "
                f"rc={self.rc:.6g}, grid={self.grid}, P={self.P}, m={self.m:.6g}, "
                f"precision={self.precision!r}, tol={self.tol}, kernel={self.kernel!r})")


# ---- evaluator context ----------------------------------------------------------
class Context:
    """pb_context: owns device scratch; one per host thread."""

    def __init__(self, device=0):
        self.handle = None
        h = _VP()
        _check(lib().pb_create(ctypes.byref(h), int(device)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().pb_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vel(self, fn, params, xn, f, y, u=None, stream=None):
        import torch
        ns, nt = xn.shape[0], y.shape[0]
        _shape(xn, (ns, 3), "xn")
        _shape(f, (ns, 3), "f")
        _shape(y, (nt, 3), "y")
        if u is None:
            u = torch.empty((nt, 3), dtype=torch.float64, device=y.device)
        _shape(u, (nt, 3), "u")
        p = params._struct()
        st = fn(self.handle, ctypes.byref(p), int(xn.shape[0]), _dptr(xn), _dptr(f), int(nt),
                _dptr(y), _dptr(u), _stream_ptr(stream))
        _check(st, self)
        return u

    def velocity(self, params, xn, f, y, u=None, stream=None):
        """u = real + k-space velocity at y from forces f at xn (device tensors (n, 3))."""
        return self._vel(lib().pb_velocity, params, xn, f, y, u, stream)

    def velocity_noalloc(self, params, xn, f, y, u, stream=None):
        """velocity() without allocating device memory (pb_velocity_noalloc):
        u must be given; the library's fixed device scratch (shared by every
        call on the device) holds everything else."""
        return self._vel(lib().pb_velocity_noalloc, params, xn, f, y, u, stream)

    def velocity_real(self, params, xn, f, y, u=None, stream=None):
        return self._vel(lib().pb_velocity_real, params, xn, f, y, u, stream)

    def velocity_kspace(self, params, xn, f, y, u=None, stream=None):
        return self._vel(lib().pb_velocity_kspace, params, xn, f, y, u, stream)

    def velocity_batched(self, params, xn, f, src_off, y, tgt_off, u=None, stream=None):
        """Independent periodic systems in one call (pb_velocity_batched).
        params: list of Params (systems with unset d or grid are autotuned and
        updated in place); xn, f (sum ns, 3) and y (sum nt, 3) float64 device
        tensors; src_off, tgt_off: host sequences of nsys + 1 offsets."""
        import torch
        n = len(params)
        if len(src_off) != n + 1 or len(tgt_off) != n + 1:
            raise PBError(PB_ERR_INVALID, f"src_off and tgt_off need {n + 1} entries")
        structs = (_Params * n)(*[q._struct() for q in params])
        so = (ctypes.c_int * (n + 1))(*[int(v) for v in src_off])
        to = (ctypes.c_int * (n + 1))(*[int(v) for v in tgt_off])
        _shape(xn, (int(src_off[-1]), 3), "xn")
        _shape(f, (int(src_off[-1]), 3), "f")
        _shape(y, (int(tgt_off[-1]), 3), "y")
        if u is None:
            u = torch.empty((y.shape[0], 3), dtype=torch.float64, device=y.device)
        _shape(u, (int(tgt_off[-1]), 3), "u")
        st = lib().pb_velocity_batched(self.handle, n, structs, so, _dptr(xn), _dptr(f), to, _dptr(y), _dptr(u),
                                       _stream_ptr(stream))
        _check(st, self)
        for q, sp in zip(params, structs):
            q.d, q.rc, q.m, q.P = sp.d, sp.rc, sp.m, sp.P
            q.grid = tuple(sp.grid[i] for i in range(3))
        return u

    def velocity_multirhs(self, params, xn, F, y, u=None, stream=None):
        """Velocities of R force sets on one particle set (pb_velocity_multirhs).
        xn (ns, 3) and y (nt, 3) float64, F (R, ns, 3) float32 device tensors;
        returns u (R, nt, 3) float64."""
        import torch
        if not (F.is_cuda and F.dtype == torch.float32 and F.is_contiguous() and F.dim() == 3):
            raise PBError(PB_ERR_INVALID, "expected a contiguous float32 CUDA tensor (R, ns, 3)")
        R, ns, nt = F.shape[0], xn.shape[0], y.shape[0]
        _shape(xn, (ns, 3), "xn")
        _shape(F, (R, ns, 3), "F")
        _shape(y, (nt, 3), "y")
        if u is None:
            u = torch.empty((R, nt, 3), dtype=torch.float64, device=y.device)
        _shape(u, (R, nt, 3), "u")
        p = params._struct()
        st = lib().pb_velocity_multirhs(self.handle, ctypes.byref(p), int(xn.shape[0]), _dptr(xn), int(R),
                                        ctypes.c_void_p(F.data_ptr()), int(nt), _dptr(y), _dptr(u),
                                        _stream_ptr(stream))
        _check(st, self)
        return u

    PHASES = ("binning", "spread", "fft_forward", "multiply", "fft_inverse", "gather", "pair")

    def set_timing(self, enable=True):
        """Record HIP events around each phase of velocity() on its stream
        (enable="pair": only around the pair kernel; each event costs the
        stream about 6 us)."""
        mode = 2 if enable == "pair" else (1 if enable else 0)
        _check(lib().pb_set_timing(self.handle, mode), self)

    def last_timings(self):
        """Average milliseconds per phase of the velocity() calls since set_timing()."""
        ms = (ctypes.c_float * 7)()
        _check(lib().pb_last_timings(self.handle, ms, 7), self)
        return dict(zip(self.PHASES, [float(v) for v in ms]))

    def last_pair_counts(self):
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().pb_last_pair_counts(self.handle, ctypes.byref(a), ctypes.byref(b)), self)
        return a.value, b.value

    def last_pair_stats(self):
        """(FP64 pairs, FP32 pairs, candidates tested, candidates read) of the
        last real-space pass."""
        out = (ctypes.c_longlong * 4)()
        _check(lib().pb_last_pair_stats(self.handle, out), self)
        return tuple(int(v) for v in out)


# ---- reference-signature entry points -----------------------------------------
def periodicBrinkman(dA, dx, dx0, a, d, e, M, N, L, maxBlockSize=256):
    """dA += Ewald matrix (device tensors; dA is (3nt, 3ns) column-major, i.e. a
    Fortran-ordered view: pass dA.T of a contiguous (3ns, 3nt) tensor)."""
    nt, ns = dx.shape[0], dx0.shape[0]
    _shape(dx, (nt, 3), "dx")
    _shape(dx0, (ns, 3), "dx0")
    if dA.numel() != 9 * nt * ns or not dA.is_contiguous():
        raise PBError(PB_ERR_INVALID, f"dA needs 9 nt ns = {9 * nt * ns} contiguous elements")
    Lh = np.ascontiguousarray(np.asarray(L, dtype=np.float64).reshape(3))
    lda = 3 * nt
    st = lib().periodicBrinkman(ctypes.c_void_p(dA.data_ptr()), lda, _dptr(dx), nt, _dptr(dx0), ns,
                                float(a), float(d), float(e), int(M), int(N), _dp(Lh),
                                int(maxBlockSize))
    _check(st)
    return dA


def fftBrinkman(y, f, xn, L, M, P, m, C, e, tol, d, alpha=1.0):
    """k-space velocity at y (host arrays (n, 3)); returns U of shape (m, 3)."""
    y, f, xn = _hpts(y), _hpts(f), _hpts(xn)
    U = np.zeros_like(y)
    Lh = np.ascontiguousarray(np.asarray(L, dtype=np.float64).reshape(3))
    Mv, Pv = _iv(M), _iv(P)
    st = lib().fftBrinkman(_dp(U), _dp(y), y.shape[0], _dp(f), _dp(xn), xn.shape[0], _dp(Lh),
                           Mv.ctypes.data_as(_I), Pv.ctypes.data_as(_I), int(m), int(C), float(e),
                           float(tol), float(d), float(alpha))
    _check(st)
    return U


def _shells(fn, A, absA, x, x0, newM, oldM, L, a, d, e):
    x, x0 = _hpts(x), _hpts(x0)
    assert A.flags.f_contiguous and absA.flags.f_contiguous and A.dtype == np.float64
    nm, om = _iv(newM), _iv(oldM)
    Lh = np.ascontiguousarray(np.asarray(L, dtype=np.float64).reshape(3))
    st = fn(_dp(A), _dp(absA), A.shape[0], _dp(x), x.shape[0], _dp(x0), x0.shape[0],
            nm.ctypes.data_as(_I), om.ctypes.data_as(_I), _dp(Lh), float(a), float(d), float(e))
    _check(st)


def realShells(A, absA, x, x0, newM, oldM, L, a, d, e):
    """In place on Fortran-ordered host arrays A, absA of shape (3nt, 3ns)."""
    _shells(lib().realShells, A, absA, x, x0, newM, oldM, L, a, d, e)


def fourierShells(A, absA, x, x0, newM, oldM, L, a, d, e):
    _shells(lib().fourierShells, A, absA, x, x0, newM, oldM, L, a, d, e)


def vR(xh, a, d, e):
    A = np.zeros(9)
    x = np.ascontiguousarray(np.asarray(xh, dtype=np.float64).reshape(3))
    _check(lib().vR(_dp(A), _dp(x), float(a), float(d), float(e)))
    return A.reshape(3, 3).T.copy()


def vF(xh, a, d, k):
    A = np.zeros(9)
    x = np.ascontiguousarray(np.asarray(xh, dtype=np.float64).reshape(3))
    kk = np.ascontiguousarray(np.asarray(k, dtype=np.float64).reshape(3))
    _check(lib().vF(_dp(A), _dp(x), float(a), float(d), _dp(kk)))
    return A.reshape(3, 3).T.copy()


def vR_batch(xh, a, d, e, stream=None):
    """9 values per point (column-major 3x3) for device tensor xh (n, 3)."""
    import torch
    n = xh.shape[0]
    _shape(xh, (n, 3), "xh")
    A = torch.empty((n, 9), dtype=torch.float64, device=xh.device)
    _check(lib().pb_vR_batch(_dptr(A), _dptr(xh), n, float(a), float(d), float(e),
                             _stream_ptr(stream)))
    return A


def vF_batch(xh, k, a, d, stream=None):
    import torch
    n = xh.shape[0]
    _shape(xh, (n, 3), "xh")
    _shape(k, (n, 3), "k")
    A = torch.empty((n, 9), dtype=torch.float64, device=xh.device)
    _check(lib().pb_vF_batch(_dptr(A), _dptr(xh), _dptr(k), n, float(a), float(d),
                             _stream_ptr(stream)))
    return A


# ---- deterministic synthetic inputs --------------------------------------------
class Rng:
    """xoshiro256** seeded by SplitMix64 (pb_rng_*)."""

    def __init__(self, seed):
        self.state = (ctypes.c_ulonglong * 4)()
        lib().pb_rng_seed(self.state, ctypes.c_ulonglong(seed))

    def uniform(self, n, scale=1.0):
        out = np.empty(int(n))
        lib().pb_rng_uniform(self.state, int(n), _dp(out), float(scale))
        return out

    def normal(self, n):
        out = np.empty(int(n))
        lib().pb_rng_normal(self.state, int(n), _dp(out))
        return out


def config_seed(index):
    """Seed of workload `index` (0-based position in the configuration list)."""
    return 0x5EED0000 + int(index)


def uniform_system(seed, ns, nt, L=(1.0, 1.0, 1.0)):
    """Draw order: source positions, forces, target positions (all (n, 3) float64)."""
    r = Rng(seed)
    L = np.asarray(L, dtype=np.float64)
    x = r.uniform(3 * ns).reshape(ns, 3) * L
    f = r.normal(3 * ns).reshape(ns, 3)
    y = r.uniform(3 * nt).reshape(nt, 3) * L if nt > 0 else np.zeros((0, 3))
    return x, f, y


def clustered_system(seed, n, nclusters=64, sigma=0.02, L=(1.0, 1.0, 1.0)):
    """n points in Gaussian clusters (centres uniform), wrapped into the box."""
    r = Rng(seed)
    L = np.asarray(L, dtype=np.float64)
    centres = r.uniform(3 * nclusters).reshape(nclusters, 3) * L
    which = np.minimum((r.uniform(n) * nclusters).astype(np.int64), nclusters - 1)
    x = centres[which] + sigma * r.normal(3 * n).reshape(n, 3)
    x = np.mod(x, L)
    f = r.normal(3 * n).reshape(n, 3)
    return x, f


def batched_systems(seed, nsys=512, nmin=100, nmax=20000):
    """Independent periodic systems of uneven size, targets = sources.  One
    stream (seed) draws, per system in turn, N = floor(exp(U(ln nmin,
    ln nmax))), the cubic box side L ~ U[0.5, 2) and alpha = exp(U(0, ln 50));
    eps = 1e-3 L.  System i's points are uniform_system(seed + ((i + 1) << 32),
    N, 0, (L, L, L)).  Returns (systems, x, f, offsets): systems a list of
    dicts (n, L, alpha, eps), x and f the concatenated (sum N, 3) arrays."""
    r = Rng(seed)
    systems, xs, fs = [], [], []
    for i in range(nsys):
        u = r.uniform(3)
        n = int(np.floor(np.exp(np.log(nmin) + u[0] * (np.log(nmax) - np.log(nmin)))))
        L = 0.5 + 1.5 * float(u[1])
        alpha = float(np.exp(u[2] * np.log(50.0)))
        x, f, _ = uniform_system(seed + ((i + 1) << 32), n, 0, (L, L, L))
        systems.append({"n": n, "L": L, "alpha": alpha, "eps": 1e-3 * L})
        xs.append(x)
        fs.append(f)
    off = np.concatenate([[0], np.cumsum([s["n"] for s in systems])]).astype(np.int64)
    return systems, np.concatenate(xs), np.concatenate(fs), off


# ---- the reference driver ---------------------------------------------------------
def read_param_file(path):
    """Parameters of the reference driver's parameter file (pb_read_param_file):
    a dict with keys Lx, Ly, Lz, d, e, alpha, M, boxSize, toPrint; raises
    PBError if the file cannot be opened."""
    p = _DriverParams()
    _check(lib().pb_read_param_file(os.fsencode(path), ctypes.byref(p)))
    return {k: getattr(p, k) for k, _ in _DriverParams._fields_}


def driver_defaults():
    p = _DriverParams()
    lib().pb_driver_defaults(ctypes.byref(p))
    return {k: getattr(p, k) for k, _ in _DriverParams._fields_}


def split_sweep(params, x):
    """The driver's split sweep (pb_split_sweep) for targets x (nt, 3): returns
    (dval[8], err[8])."""
    p = _DriverParams(**{k: params[k] for k, _ in _DriverParams._fields_})
    x = np.ascontiguousarray(x, dtype=np.float64)
    dval, err = np.zeros(8), np.zeros(8)
    _check(lib().pb_split_sweep(ctypes.byref(p), _dp(x), int(x.shape[0]), _dp(dval), _dp(err)))
    return dval, err

