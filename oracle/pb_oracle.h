/*
 * pb_oracle.h -- FP64 CPU restatement of the periodic Brinkman Ewald evaluator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library never links or calls it.
 *
 * Reference: hoangngan82/periodicBrinkman, cuda/periodicBrinkman.h and
 * cuda/fftBrinkman.h (cited as file:line).  Layout conventions follow the
 * reference: points are 3 x n column-major (x[3*i + c]); 3x3 blocks are
 * column-major (IDX2D(row,col) = col*3 + row, periodicBrinkman.h:16-17);
 * the dense matrix is 3*nt x 3*ns column-major with leading dimension lda.
 *
 * Mathematics (viscosity 1, Brinkman parameter a = alpha, split d, blob e):
 *   S(x) = H1(r) I + H2(r) x x^T
 *   K_R(x) = S from [diffH(d) - diffH(e)]        (periodicBrinkman.h:152-176)
 *   K_F(k;x) = phi_d(k) (k^2 I - k k^T)/(k^2 (k^2 + a^2)) cos(k.x)   (:183-208)
 *   u(y) = sum_n [ sum_l K_R(y - x_n - l L) + (1/tau) sum_{k != 0} K_F(k; y - x_n) ] f_n
 * The 1/tau factor follows fourierShells (:718,745); the k = 0 mode is
 * excluded as in :377-380 and fftBrinkman.h:201-207.
 */
#ifndef PB_ORACLE_H
#define PB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- pointwise functions ------------------------------------------------ */

/* e^{c x z} erfc(z + x) with the reference's underflow guard
 * (periodicBrinkman.h:19-25). */
double pbo_experf(double x, double z, double c);

/* Literal restatements of singH/regH/diffH (periodicBrinkman.h:31-112): the
 * reference formulas exactly as written, with the small-r switch `eps`
 * supplied by the caller (the reference's eps is not vendored). */
void pbo_singH_literal(double r2, double a, double eps, double *H1, double *H2);
void pbo_regH_literal(double r2, double a, double d, double eps, double *H1, double *H2);
void pbo_diffH_literal(double r2, double a, double d, double eps, double *H1, double *H2);

/* Accurate evaluations of the same functions: direct formulas where they are
 * well conditioned, Taylor series near r = 0 (regularized part: series in
 * (r/d)^2 whose coefficients are moments of the blob transform; singular
 * part: series in a*r).  singH returns the reference's sentinel 1.79e308 at
 * r = 0 (periodicBrinkman.h:43-45). */
void pbo_singH(double r2, double a, double *H1, double *H2);
void pbo_regH(double r2, double a, double d, double *H1, double *H2);
void pbo_diffH(double r2, double a, double d, double *H1, double *H2);

/* Series coefficients of regH in z = r/d:  H1 = sum_m c1[m] z^(2m),
 * H2 = sum_m c2[m] z^(2m) / d^2 ... see pb_oracle.c for the exact scaling. */
void pbo_regH_series_coeffs(double a, double d, int nterms, double *c1, double *c2);

/* K_R tensor (periodicBrinkman.h:152-176), output 9 doubles column-major. */
void pbo_vR(double *A, const double *xh, double a, double d, double e);
/* K_F tensor (periodicBrinkman.h:183-208), no 1/tau (as in the reference). */
void pbo_vF(double *A, const double *xh, double a, double d, const double *k);
/* Fourier symbol of the split blob: phi_d(k) = [1 + (1-e^{-a^2 d^2/4}) k^2/a^2] e^{-d^2 k^2/4}
 * (periodicBrinkman.h:191-195). */
double pbo_blob_hat(double k2, double a, double d);

/* ---- shell sums: reference semantics (periodicBrinkman.h:557-824) ------- */

/* realShells: if newM == oldM in components 0 and 1, A and absA are
 * OVERWRITTEN with the home-box term; otherwise the image layers between the
 * cubes [-oldM,oldM] and [-newM,newM] are ADDED.  nt targets x, ns sources x0. */
void pbo_realShells(double *A, double *absA, int lda, const double *x, int nt,
                    const double *x0, int ns, const int *newM, const int *oldM,
                    const double *L, double a, double d, double e);
/* fourierShells: adds vF/tau over the k layers between the cubes (k = 0
 * excluded); no-op when newM == oldM in components 0 and 1. */
void pbo_fourierShells(double *A, double *absA, int lda, const double *x, int nt,
                       const double *x0, int ns, const int *newM, const int *oldM,
                       const double *L, double a, double d, double e);

/* Dense Ewald matrix over full cubes: A = sum_{l in [-M,M]} vR + (1/tau) sum_{k in [-K,K] \ 0} vF
 * (overwrite).  Equivalent to main()'s realShells(0), realShells(M),
 * fourierShells(M) sequence (periodicBrinkman.cu:176-182). */
void pbo_ewald_matrix(double *A, int lda, const double *y, int nt, const double *x, int ns,
                      const double *L, double a, double d, double e, const int *M, const int *K);

/* ---- velocity forms ------------------------------------------------------ */

/* Direct Ewald velocity over full cubes (no matrix): real images [-M,M]^3,
 * k modes [-K,K]^3 \ 0 via structure factors.  u is overwritten (3 x nt). */
void pbo_velocity_direct(double *u, const double *y, int nt, const double *x, const double *f,
                         int ns, const double *L, double a, double d, double e,
                         const int *M, const int *K);
/* Real-space part only over the image cube. */
void pbo_real_velocity_cube(double *u, const double *y, int nt, const double *x,
                            const double *f, int ns, const double *L, double a, double d,
                            double e, const int *M);
/* k-space part only, direct sum over the mode cube. */
void pbo_kspace_direct(double *u, const double *y, int nt, const double *x, const double *f,
                       int ns, const double *L, double a, double d, const int *K);
/* Real-space part with a spherical cutoff rc: every (target, source image)
 * pair with |y - x - lL| < rc, via a cell list.  u overwritten. */
void pbo_real_velocity_cutoff(double *u, const double *y, int nt, const double *x,
                              const double *f, int ns, const double *L, double a, double d,
                              double e, double rc);
/* Spectral k-space part: the fftBrinkman algorithm (fftBrinkman.h:34-287)
 * with a centred window, no conjugation and a window-only gather.
 * grid[3] = FFT grid, P[3] = window points per axis, m = window half-width in
 * standard deviations.  u overwritten. */
void pbo_kspace_spectral(double *u, const double *y, int nt, const double *x, const double *f,
                         int ns, const double *L, double a, double d, const int *grid,
                         const int *P, double m);
/* Reference-signature restatement of fftBrinkman (fftBrinkman.h:34-50):
 * U = k-space velocity at y.  C, e and tol are accepted and unused, as in the
 * reference.  Returns 0. */
int pbo_fftBrinkman(double *U, const double *y, int m, const double *f, const double *xn, int N,
                    const double *L, const int *M, const int *P, int msd, int C, double e,
                    double tol, double d, double alpha);

/* 3D complex FFT used by the spectral form (mixed radix, any size).
 * data: interleaved complex, index (i0*n1 + i1)*n2 + i2; sign -1 forward. */
void pbo_fft3d(double *data, const int *n, int sign);

/* Threads used by the parallel loops (0 = all available); returns the count. */
int pbo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
