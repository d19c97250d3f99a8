/*
 * periodicbrinkman.h -- C ABI of the MI355X (gfx950) periodic Brinkman Ewald
 * evaluator (library: periodicbrinkman_amd/lib/libpbamd.so).
 *
 * Computes the velocity of a triply periodic Brinkman flow (viscosity 1,
 * Brinkman parameter alpha) driven by regularized point forces, by Ewald
 * summation: a real-space image sum over cell-listed neighbours plus a
 * k-space sum (Gaussian spreading, 3D FFT, Brinkman projection, inverse FFT,
 * gather).  All arithmetic runs on the GPU; nothing here falls back to the
 * host CPU.
 *
 * Conventions (as in the reference, hoangngan82/periodicBrinkman):
 *   - point arrays are 3 x n column-major: x[3*i + c];
 *   - 3x3 blocks are column-major (periodicBrinkman.h:16-17);
 *   - the dense matrix is 3*nt x 3*ns column-major with leading dimension lda;
 *   - the k-space sum carries the 1/(Lx*Ly*Lz) factor of fourierShells
 *     (periodicBrinkman.h:718,745) and excludes k = 0.
 *
 * Every function returns an int status (PB_OK = 0).  The entry points that
 * replace the reference's own functions keep its names and argument meaning;
 * the reference's asserts become status codes.
 */
#ifndef PERIODICBRINKMAN_AMD_H
#define PERIODICBRINKMAN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define PB_OK 0
#define PB_ERR_INVALID 1     /* bad argument (size, NaN parameter, null pointer) */
#define PB_ERR_HIP 2         /* HIP runtime error (no device, launch failure) */
#define PB_ERR_NOMEM 3       /* device allocation failed */
#define PB_ERR_UNSUPPORTED 4 /* valid but not supported (e.g. FFT length) */

#define PB_PREC_FP64 0  /* plain form: every step in FP64 */
#define PB_PREC_MIXED 1 /* FP32 k-space (spread, FFT, multiply, interpolation);
                           real space as chosen by the kernel field */

#define PB_KERNEL_TABLE 0  /* real-space radial functions from a piecewise-polynomial
                              table built from the closed forms, FP64 (default) */
#define PB_KERNEL_DIRECT 1 /* closed forms (exp, erfcx) evaluated per pair; with
                              PB_PREC_MIXED, pairs beyond max(3d, 8e) in FP32 */
#define PB_KERNEL_TABLE_FAR32 2 /* table; with PB_PREC_MIXED, pairs beyond
                              max(3d, 8e) from an FP32 copy of the table */

/* ---- entry points replacing the reference's functions -------------------- */

/* Replaces periodicBrinkman(MatrixOnDevice dA, dx, dx0, a, d, e, M, N, L,
 * maxBlockSize) -- cuda/periodicBrinkman.h:432-441.
 * dA (device, 3nt x 3ns col-major, lda >= 3nt) is ACCUMULATED:
 *   dA += sum_{l in real shells} vR(y_i - x_j - l L) + (1/tau) sum_{k in Fourier shells, k != 0} vF
 * with the reference's anisotropic shell counts (periodicBrinkman.h:457-487).
 * dx: nt targets, dx0: ns sources (device, 3 x n).  The real sum is skipped
 * for M < 0, the Fourier sum for N <= 0 (periodicBrinkman.h:499,510).
 * maxBlockSize is accepted for compatibility and ignored.  Launched on the
 * null stream, asynchronous, like the reference. */
int periodicBrinkman(double *dA, int lda, const double *dx, int nt, const double *dx0, int ns,
                     double a, double d, double e, int M, int N, const double *L,
                     int maxBlockSize);

/* Replaces fftBrinkman(U, y, f, xn, L, M, P, m, C, e, tol, d, alpha) --
 * cuda/fftBrinkman.h:34-50.  Host pointers.  U (3 x m) is OVERWRITTEN with
 * the k-space part of the velocity at the m points y, from N forces f at xn,
 * on an M[0] x M[1] x M[2] grid with P[i] window points per axis and window
 * half-width msd standard deviations.  C, e and tol are accepted and unused,
 * as in the reference.  Synchronous.  Computed in FP64 on the GPU. */
int fftBrinkman(double *U, const double *y, int m, const double *f, const double *xn, int N,
                const double *L, const int *M, const int *P, int msd, int C, double e, double tol,
                double d, double alpha);

/* Replaces realShells(A, absA, x, x0, newM, oldM, L, a, d, e) --
 * cuda/periodicBrinkman.h:564-573.  Host pointers, A and absA 3nt x 3ns
 * col-major.  If newM == oldM in components 0 and 1, A and absA are
 * OVERWRITTEN with the home-box term vR(x_i - x0_j) and its absolute value;
 * otherwise the image layers between the cubes [-oldM, oldM] and
 * [-newM, newM] are ADDED to A (and |.| to absA).  Synchronous. */
int realShells(double *A, double *absA, int lda, const double *x, int nt, const double *x0, int ns,
               const int *newM, const int *oldM, const double *L, double a, double d, double e);

/* Replaces fourierShells(...) -- cuda/periodicBrinkman.h:703-711.  Adds
 * vF/tau over the k layers between the cubes (k = 0 excluded) to A and |.| to
 * absA; no-op when newM == oldM in components 0 and 1.  Synchronous. */
int fourierShells(double *A, double *absA, int lda, const double *x, int nt, const double *x0,
                  int ns, const int *newM, const int *oldM, const double *L, double a, double d,
                  double e);

/* Replaces the pointwise vR(A, xh, a, d, e) -- periodicBrinkman.h:152-156.
 * Host pointers; A receives 9 doubles column-major.  Evaluated on the GPU. */
int vR(double *A, const double *xh, double a, double d, double e);

/* Replaces the pointwise vF(A, xh, a, d, k) -- periodicBrinkman.h:183-187
 * (no 1/tau, as in the reference).  Evaluated on the GPU. */
int vF(double *A, const double *xh, double a, double d, const double *k);

/* Batched pointwise forms on device pointers: A is 9 x n (column-major
 * 3x3 blocks), xh 3 x n.  vF_batch takes one k per point (3 x n). */
int pb_vR_batch(double *A, const double *xh, int n, double a, double d, double e, void *stream);
int pb_vF_batch(double *A, const double *xh, const double *k, int n, double a, double d,
                void *stream);

/* ---- velocity evaluator (no counterpart in the reference: the reference
 * returns the matrix, this is the velocity-from-forces composition) ------- */

typedef struct pb_context pb_context;

typedef struct pb_params {
    double L[3];   /* periodic box */
    double alpha;  /* Brinkman parameter a > 0 */
    double eps;    /* blob width e >= 0 (0: singular Brinkmanlet) */
    double d;      /* Ewald split (> 0; pb_tune fills it when <= 0) */
    double rc;     /* real-space cutoff (> 0; pb_tune fills it when <= 0) */
    int grid[3];   /* FFT grid (pb_tune fills zeros) */
    int P;         /* Gaussian window points per axis */
    double m;      /* window half-width in standard deviations (<= 0: sqrt(pi P - 5), P >= 4) */
    int precision; /* PB_PREC_FP64 or PB_PREC_MIXED */
    double tol;    /* target relative error for pb_tune */
    int kernel;    /* PB_KERNEL_TABLE, PB_KERNEL_DIRECT or PB_KERNEL_TABLE_FAR32 */
} pb_params;

/* Defaults: unit box, alpha 1, eps 0, P 12, FP64, tol 1e-6, the rest 0. */
void pb_params_default(pb_params *p);

/* Fill d, rc, grid and m of p (those <= 0) for ns sources and nt targets
 * from the error model; host only. */
int pb_tune(pb_params *p, int ns, int nt);

/* Error-model autotuner (replaces the reference's dead optimalNumShells,
 * periodicBrinkman.h:826-1016, with an error model of this evaluator).  For
 * the box, alpha, eps, tol and precision in p it scores every FFT grid (powers
 * of two per axis, 4..512) and window P in {4,6,8,10,12,16,20,24} whose
 * predicted errors meet tol, sets d and rc from the error model, and keeps
 * the combination with the smallest predicted time on MI355X; p->kernel is
 * kept.  The report (may be NULL) holds the prediction. */
typedef struct pb_tune_report {
    double err_real;   /* e^{-(rc/d)^2}: real-space truncation */
    double err_kspace; /* e^{-(X + 7.5)}, X = (d^2/4 - eta/2) kmax^2: dropped modes */
    double err_window; /* e^{-m^2/2}: window truncation (m = sqrt(pi P - 5)) */
    double pairs;      /* predicted (target, source image) pairs inside rc */
    double t_real_ms, t_kspace_ms, t_bin_ms, t_total_ms; /* predicted phase times */
    int candidates;    /* (grid, P) combinations scored */
} pb_tune_report;
int pb_autotune(pb_params *p, int ns, int nt, pb_tune_report *report);

/* One context per host thread; it owns the device scratch (cell lists,
 * grids, twiddles).  device: HIP device ordinal. */
int pb_create(pb_context **ctx, int device);
int pb_destroy(pb_context *ctx);
const char *pb_last_error(const pb_context *ctx);
const char *pb_status_string(int status);

/* u (device, 3 x nt) = velocity at the nt targets y from the ns forces f at
 * xn (device, 3 x ns): real space (cutoff p->rc) plus k space, OVERWRITTEN.
 * Positions may lie outside the box (they are wrapped).  Asynchronous on
 * `stream` (a hipStream_t, NULL for the null stream). */
int pb_velocity(pb_context *ctx, const pb_params *p, int ns, const double *xn, const double *f,
                int nt, const double *y, double *u, void *stream);

/* pb_velocity without allocating: no reference counterpart.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, result and
 * overwrite semantics as pb_velocity; it calls no hipMalloc, hipFree or other
 * allocating routine.  Every device buffer beyond xn, f, y and u comes from
 * one static device array of PB_NOALLOC_MAX_BYTES per device, held until the
 * process ends and shared by every call on that device (the last call's
 * contents stay in it): calls that overlap on one device -- on different
 * streams, or from several host threads -- race; calls on one stream are
 * ordered.  The pair kernel runs after the k space on `stream` (no internal
 * stream).  Rejected with PB_ERR_UNSUPPORTED, before any work is enqueued, when
 * the workload needs more scratch than the array holds (about 128 MiB: e.g.
 * 1e5 sources and targets on a 128^3 grid fit in both precisions, a 256^3
 * grid does not) or more than 2^18 cell and target keys.  The first call with
 * new parameters uploads the kernel constants and waits for that copy. */
#define PB_NOALLOC_MAX_BYTES (128u << 20)
int pb_velocity_noalloc(pb_context *ctx, const pb_params *p, int ns, const double *xn, const double *f,
                        int nt, const double *y, double *u, void *stream);

/* The two halves separately (same conventions; used for diagnostics). */
int pb_velocity_real(pb_context *ctx, const pb_params *p, int ns, const double *xn,
                     const double *f, int nt, const double *y, double *u, void *stream);
int pb_velocity_kspace(pb_context *ctx, const pb_params *p, int ns, const double *xn,
                       const double *f, int nt, const double *y, double *u, void *stream);

/* Many independent periodic systems in one call.  System i has the sources
 * src_off[i] .. src_off[i+1]-1 of xn and f and the targets tgt_off[i] ..
 * tgt_off[i+1]-1 of y and u (device arrays of 3 x n doubles; offsets are host
 * arrays of nsys + 1 points, starting at 0) and its own parameters params[i]
 * (box, alpha, eps, precision, kernel, tol); a system whose d or grid is
 * unset is tuned by pb_autotune, which writes the choice back into
 * params[i].  The kernel constants and tables of all systems are built and
 * uploaded once; systems with moderate cell grids then run together, one
 * launch per phase and kernel class for all of them, the others one by one
 * on internal streams, all ordered after `stream` without further host
 * synchronization.  The second of two identical calls (systems, parameters,
 * arrays, stream) captures a hipGraph that later identical calls replay.
 * u is OVERWRITTEN.  No reference counterpart: the reference evaluates one
 * system per call. */
int pb_velocity_batched(pb_context *ctx, int nsys, pb_params *params, const int *src_off, const double *xn,
                        const double *f, const int *tgt_off, const double *y, double *u, void *stream);

/* Many right-hand sides on one particle set: u (device, double, R blocks of
 * 3 x nt; block r at r*3*nt) = velocities of the R force sets f (device,
 * float, R blocks of 3 x ns; block r at r*3*ns) at the nt targets y, real
 * space plus k space, OVERWRITTEN.  Real space is the contraction
 * [3 T x 3 S] . [3 S x R] of the FP32 kernel matrix with the forces on the
 * matrix cores (FP32 accumulation); k space runs in FP32 per force set.
 * Always the mixed form (p->precision is not used); p->kernel must be a table
 * kernel.  No reference counterpart: the reference evaluates one force set
 * per call (periodicBrinkman.h:432, fftBrinkman.h:34). */
int pb_velocity_multirhs(pb_context *ctx, const pb_params *p, int ns, const double *xn, int R,
                         const float *f, int nt, const double *y, double *u, void *stream);

/* ---- the reference driver (periodicBrinkman.cu:32-219) -------------------- */

/* Parameters of the reference driver's parameter file: one "key: value" pair
 * per line, split at ':' and ' ' (periodicBrinkman.cu:44-70); keys Lx, Ly,
 * Lz, d, e, M, alpha, boxSize, toPrint.  Keys the file does not set keep the
 * reference defaults (periodicBrinkman.cu:79-87): 1, 1, 1, 0.6, 0, 4, 1, 1,
 * 0.  Other keys and lines without a value are ignored. */
typedef struct pb_driver_params {
    double Lx, Ly, Lz, d, e, alpha;
    int M, boxSize, toPrint;
} pb_driver_params;
void pb_driver_defaults(pb_driver_params *p);
int pb_read_param_file(const char *path, pb_driver_params *p); /* PB_ERR_INVALID: cannot open */

/* The driver's split sweep (periodicBrinkman.cu:141-203): the 3 nt x 3
 * Ewald matrix of the nt targets x (host, 3 x nt) for one source at
 * (0.5, 0.5, 0.5), by realShells (home cell, then layers to M) plus
 * fourierShells (to M), at d0 = p->d and at dval[i] = (0.1 + 0.95 i) d0,
 * i < 8; err[i] = max |A(dval[i]) - A(d0)|.  Evaluated on the GPU. */
int pb_split_sweep(const pb_driver_params *p, const double *x, int nt, double dval[8], double err[8]);

/* Statistics of the last pb_velocity call on ctx (device work counters read
 * back synchronously): pairs evaluated in FP64 and FP32. */
int pb_last_pair_counts(pb_context *ctx, long long *n_fp64, long long *n_fp32);

/* Work statistics of the last real-space pass (no counterpart in the
 * reference; for tuning): out[0] pairs evaluated in FP64, out[1] in FP32,
 * out[2] candidate sources broadcast to a target tile, out[3] candidate
 * sources read from the cell runs before the tile-box filter. */
int pb_last_pair_stats(pb_context *ctx, long long out[4]);

/* Phase timing: while enabled, every pb_velocity call records HIP events on
 * its stream around its phases (no host synchronization for up to 64 calls).
 * pb_last_timings (synchronizing) returns the average milliseconds over the
 * calls since pb_set_timing for [binning, spread, forward FFT, multiply,
 * inverse FFT, gather, pair kernel] into ms[0..n-1] (n <= 7); phases that
 * did not run report 0.  pb_set_timing restarts the averages.  Each event
 * costs the stream about 6 us: enable = 2 records only the two around the
 * pair kernel (the other phases report 0), enable = 1 every boundary. */
int pb_set_timing(pb_context *ctx, int enable);
int pb_last_timings(pb_context *ctx, float *ms, int n);

/* ---- deterministic synthetic inputs (host) ------------------------------- */
/* xoshiro256** seeded by SplitMix64.  uniform: (u >> 11) * 2^-53 scaled by
 * `scale`; normal: Box-Muller pairs.  State is 4 uint64 owned by the caller. */
void pb_rng_seed(unsigned long long *state, unsigned long long seed);
void pb_rng_uniform(unsigned long long *state, long n, double *out, double scale);
void pb_rng_normal(unsigned long long *state, long n, double *out);

#ifdef __cplusplus
}
#endif
#endif
