/*
 * pb_oracle.c -- FP64 CPU restatement of the periodic Brinkman Ewald evaluator.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
 * it.  It is written for clarity and numerical care, not speed, apart from
 * OpenMP over targets.
 *
 * Reference: hoangngan82/periodicBrinkman cuda/periodicBrinkman.h (pointwise
 * functions :19-209, shell sums :557-824) and cuda/fftBrinkman.h (spectral
 * k-space :34-287).  Departures, each following the mathematics the
 * reference code implements:
 *   - 1/tau on the k-space sum everywhere (fourierShells :718,745);
 *   - 1 - e^{-a^2 d^2/4} is evaluated with expm1 (the reference's
 *     1 - exp(...) loses digits when a*d is small);
 *   - near r = 0 the regularized functions are evaluated from their Taylor
 *     series and the singular function from its series in a*r, where the
 *     closed forms cancel catastrophically (the reference switches at an
 *     un-vendored eps, periodicBrinkman.h:40,63,95);
 *   - the spectral k-space uses a centred window, no conjugation and a
 *     window-only gather (fftBrinkman.h:78-81,209-216,253-271).
 */
#include "pb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793238462643383279502884
#define SPI 1.772453850905516027298167483341145183 /* sqrt(pi) */

/* Series regions.  regH and diffH use the (r/d)^2 series of the regularized
 * part for r/d < zswitch(a d/2): the closed forms lose about
 * 10/(x^2 z^4) ulps (x = a d/2, z = r/d), so the series region widens as
 * a*d shrinks, up to r/d = 2 where 36 terms still converge to 1e-16.
 * singH uses its a*r series for a*r < SIGMA. */
#define SIGMA 0.2
#define NREG 36 /* terms of the regH series (m = 0..NREG-1) */
#define NSING 14 /* terms of the singH series (j = 0..NSING-1) */

/* ------------------------------------------------------------------------ */
/* experf: periodicBrinkman.h:19-25                                         */
/* ------------------------------------------------------------------------ */
double pbo_experf(double x, double z, double c)
{
    double a = 0;
    if ((x * z < 0) || (x * x + z * z < 700)) a = exp(c * x * z) * erfc(z + x);
    return a;
}

/* ------------------------------------------------------------------------ */
/* Literal restatements (periodicBrinkman.h:31-46, :53-78, :85-112)          */
/* ------------------------------------------------------------------------ */
void pbo_singH_literal(double r2, double a, double eps, double *H1, double *H2)
{
    double r = sqrt(r2);
    if (r > eps) {
        *H1 = -1 / (4 * PI * a * a * r * r2) + exp(-a * r) / (4 * PI * r) * (1 + 1 / (a * r) + 1 / (a * a * r2));
        *H2 = 3 / (4 * PI * a * a * r * r2 * r2) - exp(-a * r) / (4 * PI * r2 * r) * (1 + 3 / (a * r) + 3 / (a * a * r2));
    } else {
        *H1 = *H2 = 1.79e+308;
    }
}

void pbo_regH_literal(double r2, double a, double d, double eps, double *H1, double *H2)
{
    double expx2 = exp(-r2 / d / d);
    double r = sqrt(r2);
    double Erf = erf(r / d);
    if (r > eps) {
        double h1, h2;
        h1 = -1 / (4 * PI * a * a * r2 * r) * Erf;
        h1 += 1 / (2 * a * a * d * d * d * PI * SPI * r2) * expx2 * (d * d + 2 * r2) * (1 - exp(-a * a * d * d / 4));
        h1 += 1 / (8 * a * a * PI * r2 * r) * (1 + a * r + a * a * r2) * pbo_experf(a * d / 2, -r / d, 2);
        h1 -= 1 / (8 * a * a * PI * r2 * r) * (1 - a * r + a * a * r2) * pbo_experf(a * d / 2, r / d, 2);
        h2 = 3 / (4 * PI * a * a * r2 * r2 * r) * Erf;
        h2 -= (1 - exp(-a * a * d * d / 4)) / (2 * a * a * d * d * d * PI * SPI * r2 * r2) * expx2 * (3 * d * d + 2 * r2);
        h2 -= 1 / (8 * a * a * PI * r2 * r2 * r) * (3 + 3 * a * r + a * a * r2) * pbo_experf(a * d / 2, -r / d, 2);
        h2 += 1 / (8 * a * a * PI * r2 * r2 * r) * (3 - 3 * a * r + a * a * r2) * pbo_experf(a * d / 2, r / d, 2);
        *H1 = h1;
        *H2 = h2;
    } else {
        double h2;
        *H1 = -a * erfc(a * d / 2) / (6 * PI) + (2 + exp(-a * a * d * d / 4) * (a * a * d * d - 2)) / (3 * a * a * d * d * d * PI * SPI);
        h2 = (-24 + 2 * a * a * d * d * (2 - a * a * d * d)) * exp(-a * a * d * d / 4) + 24;
        *H2 = h2 / (60 * a * a * d * d * d * d * d * PI * SPI) + a * a * a / (60 * PI) * erfc(a * d / 2);
    }
}

void pbo_diffH_literal(double r2, double a, double d, double eps, double *H1, double *H2)
{
    double expx2 = exp(-r2 / d / d) * (1 - exp(-a * a * d * d / 4));
    double r = sqrt(r2);
    double Erfc = erfc(r / d);
    if (r > eps) {
        double h1, h2;
        h1 = -1 / (4 * PI * a * a) / (r2 * r) * Erfc;
        h1 -= (d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2) * expx2;
        h1 += 1 / (8 * PI * a * a * r2 * r) * (1 + a * r + a * a * r2) * pbo_experf(-a * d / 2, r / d, 2);
        h1 += 1 / (8 * PI * a * a * r2 * r) * (1 - a * r + a * a * r2) * pbo_experf(a * d / 2, r / d, 2);
        h2 = 3 / (4 * PI * a * a) / (r2 * r2 * r) * Erfc;
        h2 += (3 * d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2 * r2) * expx2;
        h2 -= 1 / (8 * PI * a * a * r2 * r2 * r) * (3 + 3 * a * r + a * a * r2) * pbo_experf(-a * d / 2, r / d, 2);
        h2 -= 1 / (8 * PI * a * a * r2 * r2 * r) * (3 - 3 * a * r + a * a * r2) * pbo_experf(a * d / 2, r / d, 2);
        *H1 = h1;
        *H2 = h2;
    } else {
        *H1 = -(4 + 2 * (-2 + a * a * d * d) * exp(-a * a * d * d / 4)) / (a * a * d * d * d * 6 * PI * SPI) + a * erfc(a * d / 2) / (6 * PI);
        *H2 = -(2 - a * a * d * d) * exp(-a * a * d * d / 4) / (30 * d * d * d * PI * SPI) - 2 * (1 - exp(-a * a * d * d / 4)) / (5 * a * a * d * d * d * d * d * PI * SPI) - a * a * a * erfc(a * d / 2) / (60 * PI);
    }
}

/* ------------------------------------------------------------------------ */
/* Series of the regularized functions                                      */
/*                                                                          */
/* With phi_d(k) = (1 + c k^2) e^{-b k^2}, b = d^2/4, c = (1-e^{-a^2 b})/a^2, */
/* the regularized tensor is S = G I + grad grad^T B where                  */
/* G^ = phi/(k^2+a^2), B^ = phi/(k^2 (k^2+a^2)).  Expanding the radial       */
/* inverse transforms f(r) = (1/(2 pi^2 r)) int f^(k) k sin(kr) dk in r gives */
/*   H1 = sum_m (-1)^m (2m+2) / ((2m+3)(2m+1)!) I_{m+1} r^{2m} / (2 pi^2)   */
/*   H2 = sum_m (-1)^m 4(m+1)(m+2) / (2m+5)!     I_{m+2} r^{2m} / (2 pi^2)   */
/* with I_n = int_0^inf phi(k) k^{2n}/(k^2+a^2) dk                           */
/*          = b^{1/2-n} [A_n(x) + gam A_{n+1}(x)],  x = a d/2,               */
/*   gam = (1 - e^{-x^2})/x^2,  A_n(x) = int_0^inf t^{2n} e^{-t^2}/(t^2+x^2) dt.*/
/* A_n is computed through a_n = A_n / (Gamma(n+1/2)/2), which obeys          */
/* a_{n+1} = (1 - x^2 a_n)/(n + 1/2), a_0 = sqrt(pi) erfcx(x)/x (forward,     */
/* stable for x <= 2), or by Gauss-Legendre quadrature for x > 2.            */
/* ------------------------------------------------------------------------ */
#define NA (NREG + 4)

static void gauss_legendre20(long double *xn, long double *wn)
{
    const int n = 20;
    for (int i = 0; i < n; i++) {
        long double x = cosl(PI * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = x;
            for (int k = 2; k <= n; k++) {
                long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            long double dp = n * (x * p1 - p0) / (x * x - 1);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) {
                p0 = 1;
                p1 = x;
                for (int k = 2; k <= n; k++) {
                    long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (x * p1 - p0) / (x * x - 1);
                xn[i] = x;
                wn[i] = 2 / ((1 - x * x) * dp * dp);
                break;
            }
        }
    }
}

static void scaled_moments(long double x, long double *an)
{
    long double x2 = x * x;
    if (x <= 2) {
        long double erfcx0 = expl(x2) * erfcl(x);
        an[0] = sqrtl(PI) * erfcx0 / x;
        for (int n = 0; n + 1 < NA; n++) an[n + 1] = (1 - x2 * an[n]) / (n + 0.5L);
        return;
    }
    long double gx[20], gw[20];
    gauss_legendre20(gx, gw);
    long double G = sqrtl(PI) / 2; /* Gamma(n+1/2)/2 at n = 0 */
    for (int n = 0; n < NA; n++) {
        long double T = sqrtl((long double)n + 1) + 10;
        int panels = (int)(T / 0.5L) + 1;
        long double hp = T / panels, sum = 0;
        for (int p = 0; p < panels; p++) {
            long double lo = p * hp;
            for (int i = 0; i < 20; i++) {
                long double t = lo + 0.5L * hp * (gx[i] + 1);
                long double t2 = t * t;
                sum += 0.5L * hp * gw[i] * powl(t2, n) * expl(-t2) / (t2 + x2);
            }
        }
        an[n] = sum / G;
        G *= (n + 0.5L);
    }
}

/* c1, c2 such that H1 = sum_m c1[m] z^{2m}, H2 = sum_m c2[m] z^{2m}, z = r/d. */
void pbo_regH_series_coeffs(double a, double d, int nterms, double *c1, double *c2)
{
    long double x = (long double)a * d / 2, x2 = x * x;
    long double gam = (x2 > 0) ? -expm1l(-x2) / x2 : 1;
    long double an[NA];
    scaled_moments(x, an);
    long double Gn[NA]; /* Gamma(n+1/2)/2 */
    Gn[0] = sqrtl(PI) / 2;
    for (int n = 1; n < NA; n++) Gn[n] = Gn[n - 1] * (n - 0.5L);
    long double pi2 = 2 * (long double)PI * PI;
    long double fact = 1; /* (2m+1)! */
    for (int m = 0; m < nterms && m < NREG; m++) {
        if (m > 0) fact *= (2 * m) * (2 * m + 1);
        long double sgn = (m % 2) ? -1 : 1;
        long double p4 = ldexpl(1, 2 * m + 1); /* 2^{2m+1} */
        long double In1 = Gn[m + 1] * (an[m + 1] + gam * (m + 1.5L) * an[m + 2]);
        c1[m] = (double)(sgn * (2 * m + 2) * p4 * In1 / ((2 * m + 3) * fact * pi2 * d));
        long double f5 = fact * (2 * m + 2) * (2 * m + 3) * (2 * m + 4) * (2 * m + 5); /* (2m+5)! */
        long double In2 = Gn[m + 2] * (an[m + 2] + gam * (m + 2.5L) * an[m + 3]);
        c2[m] = (double)(sgn * 4 * (m + 1) * (m + 2) * 4 * p4 * In2 / (f5 * pi2 * d * d * d));
    }
}

/* Per-thread cache of series coefficients for the last (a, d) pairs. */
typedef struct {
    double a, d;
    double c1[NREG], c2[NREG];
    int valid;
} series_cache_t;
static _Thread_local series_cache_t g_cache[4];
static _Thread_local int g_cache_next;

static const series_cache_t *series_for(double a, double d)
{
    for (int i = 0; i < 4; i++)
        if (g_cache[i].valid && g_cache[i].a == a && g_cache[i].d == d) return &g_cache[i];
    series_cache_t *c = &g_cache[g_cache_next];
    g_cache_next = (g_cache_next + 1) & 3;
    pbo_regH_series_coeffs(a, d, NREG, c->c1, c->c2);
    c->a = a;
    c->d = d;
    c->valid = 1;
    return c;
}

static void regH_series(double r, double a, double d, double *H1, double *H2)
{
    const series_cache_t *c = series_for(a, d);
    double z2 = (r / d) * (r / d);
    double h1 = 0, h2 = 0;
    for (int m = NREG - 1; m >= 0; m--) {
        h1 = h1 * z2 + c->c1[m];
        h2 = h2 * z2 + c->c2[m];
    }
    *H1 = h1;
    *H2 = h2;
}

static void regH_direct(double r, double a, double d, double *H1, double *H2)
{
    double r2 = r * r, z = r / d, x = a * d / 2, s = a * r;
    double E = exp(-z * z);
    double om = -expm1(-x * x);
    double Erf = erf(z);
    double ep = pbo_experf(x, -z, 2); /* e^{-ar} erfc(ad/2 - r/d) */
    double em = pbo_experf(x, z, 2);  /* e^{ar} erfc(ad/2 + r/d) */
    double c3 = 1 / (8 * PI * a * a * r2 * r);
    double c5 = c3 / r2;
    *H1 = -Erf / (4 * PI * a * a * r2 * r) + om * E * (d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2)
          + c3 * (1 + s + s * s) * ep - c3 * (1 - s + s * s) * em;
    *H2 = 3 * Erf / (4 * PI * a * a * r2 * r2 * r) - om * E * (3 * d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2 * r2)
          - c5 * (3 + 3 * s + s * s) * ep + c5 * (3 - 3 * s + s * s) * em;
}

static void singH_series(double r, double a, double *H1, double *H2)
{
    /* H1 = a/(4 pi s) sum_j (-1)^j (j+1)^2/(j+2)! s^j
     * H2 = a^3/(4 pi s^3) sum_j (-1)^{j+1} (j+1)(j-1)/(j+2)! s^j,  s = a r */
    double s = a * r;
    double p1 = 0, p2 = 0;
    for (int j = NSING - 1; j >= 0; j--) {
        double f = 1;
        for (int k = 2; k <= j + 2; k++) f *= k;
        double sg = (j % 2) ? -1.0 : 1.0;
        p1 = p1 * s + sg * (j + 1) * (j + 1) / f;
        p2 = p2 * s - sg * (j + 1) * (j - 1) / f;
    }
    *H1 = a * p1 / (4 * PI * s);
    *H2 = a * a * a * p2 / (4 * PI * s * s * s);
}

static void singH_direct(double r, double a, double *H1, double *H2)
{
    double r2 = r * r, s = a * r, es = exp(-s);
    *H1 = -1 / (4 * PI * a * a * r * r2) + es / (4 * PI * r) * (1 + 1 / s + 1 / (s * s));
    *H2 = 3 / (4 * PI * a * a * r * r2 * r2) - es / (4 * PI * r2 * r) * (1 + 3 / s + 3 / (s * s));
}

static void diffH_direct(double r, double a, double d, double *H1, double *H2)
{
    double r2 = r * r, z = r / d, x = a * d / 2, s = a * r;
    double E = exp(-z * z) * (-expm1(-x * x));
    double Erfc = erfc(z);
    double tm = pbo_experf(-x, z, 2); /* e^{-ar} erfc(r/d - ad/2) */
    double tp = pbo_experf(x, z, 2);  /* e^{ar}  erfc(r/d + ad/2) */
    double c3 = 1 / (8 * PI * a * a * r2 * r);
    double c5 = c3 / r2;
    *H1 = -Erfc / (4 * PI * a * a * r2 * r) - (d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2) * E
          + c3 * (1 + s + s * s) * tm + c3 * (1 - s + s * s) * tp;
    *H2 = 3 * Erfc / (4 * PI * a * a * r2 * r2 * r) + (3 * d * d + 2 * r2) / (2 * a * a * d * d * d * PI * SPI * r2 * r2) * E
          - c5 * (3 + 3 * s + s * s) * tm - c5 * (3 - 3 * s + s * s) * tp;
}

void pbo_singH(double r2, double a, double *H1, double *H2)
{
    double r = sqrt(r2);
    if (r <= 0) {
        *H1 = *H2 = 1.79e+308;
    } else if (a * r < SIGMA) {
        singH_series(r, a, H1, H2);
    } else {
        singH_direct(r, a, H1, H2);
    }
}

static double zswitch(double a, double d)
{
    double zs = 0.32 / sqrt(a * d / 2);
    return zs < 0.2 ? 0.2 : (zs > 2.0 ? 2.0 : zs);
}

void pbo_regH(double r2, double a, double d, double *H1, double *H2)
{
    double r = sqrt(r2);
    if (r < zswitch(a, d) * d)
        regH_series(r, a, d, H1, H2);
    else
        regH_direct(r, a, d, H1, H2);
}

void pbo_diffH(double r2, double a, double d, double *H1, double *H2)
{
    double r = sqrt(r2);
    if (r <= 0) { /* periodicBrinkman.h:105-111: the singular self term is dropped */
        regH_series(0, a, d, H1, H2);
        *H1 = -*H1;
        *H2 = -*H2;
    } else if (r < zswitch(a, d) * d) {
        double s1, s2, g1, g2;
        pbo_singH(r2, a, &s1, &s2);
        regH_series(r, a, d, &g1, &g2);
        *H1 = s1 - g1;
        *H2 = s2 - g2;
    } else {
        diffH_direct(r, a, d, H1, H2);
    }
}

/* Radial functions of K_R = diffH(d) - diffH(e) (periodicBrinkman.h:152-168).
 * For e > 0 close to r = 0 the difference is the difference of the two
 * regularized functions (the singular parts cancel exactly). */
static void vR_radial(double r2, double a, double d, double e, double *H1, double *H2)
{
    double r = sqrt(r2);
    if (e > 0 && r < 8 * e && r < 3 * d) {
        double p1, p2, q1, q2;
        pbo_regH(r2, a, e, &p1, &p2);
        pbo_regH(r2, a, d, &q1, &q2);
        *H1 = p1 - q1;
        *H2 = p2 - q2;
        return;
    }
    pbo_diffH(r2, a, d, H1, H2);
    if (e > 0) {
        double t1, t2;
        pbo_diffH(r2, a, e, &t1, &t2);
        *H1 -= t1;
        *H2 -= t2;
    }
}

void pbo_vR(double *A, const double *xh, double a, double d, double e)
{
    double H1, H2;
    double r2 = xh[0] * xh[0] + xh[1] * xh[1] + xh[2] * xh[2];
    vR_radial(r2, a, d, e, &H1, &H2);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A[j * 3 + i] = xh[i] * xh[j] * H2;
        A[i * 3 + i] += H1;
    }
}

double pbo_blob_hat(double k2, double a, double d)
{
    double x = a * d * a * d;
    double a1 = -expm1(-x / 4);
    return (1 + a1 * d * d * k2 / x) * exp(-d * d * k2 / 4);
}

void pbo_vF(double *A, const double *xh, double a, double d, const double *k)
{
    double kx = xh[0] * k[0] + xh[1] * k[1] + xh[2] * k[2];
    double kn = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double bhat = pbo_blob_hat(kn, a, d) * cos(kx) / (kn * (kn + a * a));
    for (int i = 0; i < 3; i++) {
        A[i * 3 + i] = 0;
        for (int j = 0; j < 3; j++) {
            if (i != j) {
                A[j * 3 + i] = -k[i] * k[j] * bhat;
                A[i * 3 + i] += k[j] * k[j] * bhat;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Shell sums (periodicBrinkman.h:564-693, :703-823)                        */
/* ------------------------------------------------------------------------ */
#define AIJ(A, lda, r, c) ((A)[(size_t)(c) * (size_t)(lda) + (size_t)(r)])

static int in_cube(int l0, int l1, int l2, const int *M)
{
    return abs(l0) <= M[0] && abs(l1) <= M[1] && abs(l2) <= M[2];
}

void pbo_realShells(double *A, double *absA, int lda, const double *x, int nt,
                    const double *x0, int ns, const int *newM, const int *oldM,
                    const double *L, double a, double d, double e)
{
    int home = (newM[0] == oldM[0]) && (newM[1] == oldM[1]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nt; i++) {
        double xh[3], tA[9];
        for (int j = 0; j < ns; j++) {
            if (home) {
                for (int k = 0; k < 3; k++) xh[k] = x[3 * i + k] - x0[3 * j + k];
                pbo_vR(tA, xh, a, d, e);
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++) {
                        AIJ(A, lda, 3 * i + p, 3 * j + q) = tA[q * 3 + p];
                        AIJ(absA, lda, 3 * i + p, 3 * j + q) = fabs(tA[q * 3 + p]);
                    }
                continue;
            }
            /* the layer between the cubes [-oldM,oldM] and [-newM,newM] */
            for (int l2 = -newM[2]; l2 <= newM[2]; l2++)
                for (int l1 = -newM[1]; l1 <= newM[1]; l1++)
                    for (int l0 = -newM[0]; l0 <= newM[0]; l0++) {
                        if (in_cube(l0, l1, l2, oldM)) continue;
                        xh[0] = x[3 * i + 0] - x0[3 * j + 0] - l0 * L[0];
                        xh[1] = x[3 * i + 1] - x0[3 * j + 1] - l1 * L[1];
                        xh[2] = x[3 * i + 2] - x0[3 * j + 2] - l2 * L[2];
                        pbo_vR(tA, xh, a, d, e);
                        for (int p = 0; p < 3; p++)
                            for (int q = 0; q < 3; q++) {
                                AIJ(A, lda, 3 * i + p, 3 * j + q) += tA[q * 3 + p];
                                AIJ(absA, lda, 3 * i + p, 3 * j + q) += fabs(tA[q * 3 + p]);
                            }
                    }
        }
    }
}

void pbo_fourierShells(double *A, double *absA, int lda, const double *x, int nt,
                       const double *x0, int ns, const int *newM, const int *oldM,
                       const double *L, double a, double d, double e)
{
    (void)e;
    if ((newM[0] == oldM[0]) && (newM[1] == oldM[1])) return;
    double tau = L[0] * L[1] * L[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nt; i++) {
        double xh[3], k[3], tA[9];
        for (int j = 0; j < ns; j++) {
            for (int c = 0; c < 3; c++) xh[c] = x[3 * i + c] - x0[3 * j + c];
            for (int l2 = -newM[2]; l2 <= newM[2]; l2++)
                for (int l1 = -newM[1]; l1 <= newM[1]; l1++)
                    for (int l0 = -newM[0]; l0 <= newM[0]; l0++) {
                        if (in_cube(l0, l1, l2, oldM)) continue; /* also excludes k = 0 */
                        k[0] = 2 * PI * l0 / L[0];
                        k[1] = 2 * PI * l1 / L[1];
                        k[2] = 2 * PI * l2 / L[2];
                        pbo_vF(tA, xh, a, d, k);
                        for (int p = 0; p < 3; p++)
                            for (int q = 0; q < 3; q++) {
                                double v = tA[q * 3 + p] / tau;
                                AIJ(A, lda, 3 * i + p, 3 * j + q) += v;
                                AIJ(absA, lda, 3 * i + p, 3 * j + q) += fabs(v);
                            }
                    }
        }
    }
}

void pbo_ewald_matrix(double *A, int lda, const double *y, int nt, const double *x, int ns,
                      const double *L, double a, double d, double e, const int *M, const int *K)
{
    double tau = L[0] * L[1] * L[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nt; i++) {
        double xh[3], k[3], tA[9];
        for (int j = 0; j < ns; j++) {
            double acc[9] = {0};
            for (int l2 = -M[2]; l2 <= M[2]; l2++)
                for (int l1 = -M[1]; l1 <= M[1]; l1++)
                    for (int l0 = -M[0]; l0 <= M[0]; l0++) {
                        xh[0] = y[3 * i + 0] - x[3 * j + 0] - l0 * L[0];
                        xh[1] = y[3 * i + 1] - x[3 * j + 1] - l1 * L[1];
                        xh[2] = y[3 * i + 2] - x[3 * j + 2] - l2 * L[2];
                        pbo_vR(tA, xh, a, d, e);
                        for (int q = 0; q < 9; q++) acc[q] += tA[q];
                    }
            for (int c = 0; c < 3; c++) xh[c] = y[3 * i + c] - x[3 * j + c];
            for (int l2 = -K[2]; l2 <= K[2]; l2++)
                for (int l1 = -K[1]; l1 <= K[1]; l1++)
                    for (int l0 = -K[0]; l0 <= K[0]; l0++) {
                        if (l0 == 0 && l1 == 0 && l2 == 0) continue;
                        k[0] = 2 * PI * l0 / L[0];
                        k[1] = 2 * PI * l1 / L[1];
                        k[2] = 2 * PI * l2 / L[2];
                        pbo_vF(tA, xh, a, d, k);
                        for (int q = 0; q < 9; q++) acc[q] += tA[q] / tau;
                    }
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) AIJ(A, lda, 3 * i + p, 3 * j + q) = acc[q * 3 + p];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Velocity forms                                                           */
/* ------------------------------------------------------------------------ */
static void apply_tensor(double *u, const double *A, const double *f)
{
    for (int p = 0; p < 3; p++) u[p] += A[p] * f[0] + A[3 + p] * f[1] + A[6 + p] * f[2];
}

void pbo_real_velocity_cube(double *u, const double *y, int nt, const double *x,
                            const double *f, int ns, const double *L, double a, double d,
                            double e, const int *M)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < nt; i++) {
        double acc[3] = {0, 0, 0}, xh[3], tA[9];
        for (int j = 0; j < ns; j++)
            for (int l2 = -M[2]; l2 <= M[2]; l2++)
                for (int l1 = -M[1]; l1 <= M[1]; l1++)
                    for (int l0 = -M[0]; l0 <= M[0]; l0++) {
                        xh[0] = y[3 * i + 0] - x[3 * j + 0] - l0 * L[0];
                        xh[1] = y[3 * i + 1] - x[3 * j + 1] - l1 * L[1];
                        xh[2] = y[3 * i + 2] - x[3 * j + 2] - l2 * L[2];
                        pbo_vR(tA, xh, a, d, e);
                        apply_tensor(acc, tA, f + 3 * j);
                    }
        for (int c = 0; c < 3; c++) u[3 * i + c] = acc[c];
    }
}

void pbo_kspace_direct(double *u, const double *y, int nt, const double *x, const double *f,
                       int ns, const double *L, double a, double d, const int *K)
{
    /* u(y) = (1/tau) sum_{k != 0} B(k) [cos(k.y) C(k) + sin(k.y) S(k)],
     * C(k) = sum_j cos(k.x_j) f_j, S(k) = sum_j sin(k.x_j) f_j,
     * B(k) = phi_d(k) (k^2 I - k k^T) / (k^2 (k^2 + a^2)).  Only half of the
     * modes are visited (k and -k contribute equally). */
    double tau = L[0] * L[1] * L[2];
    int n0 = 2 * K[0] + 1, n1 = 2 * K[1] + 1, n2 = 2 * K[2] + 1;
    long nk = (long)n0 * n1 * n2;
    for (int i = 0; i < nt; i++) u[3 * i] = u[3 * i + 1] = u[3 * i + 2] = 0;
    /* half space: first nonzero index positive */
#pragma omp parallel
    {
        double *uloc = calloc((size_t)3 * (nt > 0 ? nt : 1), sizeof(double));
#pragma omp for schedule(dynamic, 16)
        for (long q = 0; q < nk; q++) {
            int l0 = (int)(q / ((long)n1 * n2)) - K[0];
            int l1 = (int)((q / n2) % n1) - K[1];
            int l2 = (int)(q % n2) - K[2];
            int first = l0 != 0 ? l0 : (l1 != 0 ? l1 : l2);
            if (first <= 0) continue;
            double k[3] = {2 * PI * l0 / L[0], 2 * PI * l1 / L[1], 2 * PI * l2 / L[2]};
            double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
            double b = 2 * pbo_blob_hat(k2, a, d) / (k2 * (k2 + a * a)) / tau;
            double C[3] = {0, 0, 0}, S[3] = {0, 0, 0};
            for (int j = 0; j < ns; j++) {
                double ph = k[0] * x[3 * j] + k[1] * x[3 * j + 1] + k[2] * x[3 * j + 2];
                double cs = cos(ph), sn = sin(ph);
                for (int c = 0; c < 3; c++) {
                    C[c] += cs * f[3 * j + c];
                    S[c] += sn * f[3 * j + c];
                }
            }
            /* project: (k^2 I - k k^T) v */
            double kC = k[0] * C[0] + k[1] * C[1] + k[2] * C[2];
            double kS = k[0] * S[0] + k[1] * S[1] + k[2] * S[2];
            double PC[3], PS[3];
            for (int c = 0; c < 3; c++) {
                PC[c] = b * (k2 * C[c] - k[c] * kC);
                PS[c] = b * (k2 * S[c] - k[c] * kS);
            }
            for (int i = 0; i < nt; i++) {
                double ph = k[0] * y[3 * i] + k[1] * y[3 * i + 1] + k[2] * y[3 * i + 2];
                double cs = cos(ph), sn = sin(ph);
                for (int c = 0; c < 3; c++) uloc[3 * i + c] += cs * PC[c] + sn * PS[c];
            }
        }
#pragma omp critical
        for (int i = 0; i < 3 * nt; i++) u[i] += uloc[i];
        free(uloc);
    }
}

void pbo_velocity_direct(double *u, const double *y, int nt, const double *x, const double *f,
                         int ns, const double *L, double a, double d, double e,
                         const int *M, const int *K)
{
    double *uk = malloc((size_t)3 * (nt > 0 ? nt : 1) * sizeof(double));
    pbo_real_velocity_cube(u, y, nt, x, f, ns, L, a, d, e, M);
    pbo_kspace_direct(uk, y, nt, x, f, ns, L, a, d, K);
    for (int i = 0; i < 3 * nt; i++) u[i] += uk[i];
    free(uk);
}

static double wrap_coord(double v, double L)
{
    double w = v - floor(v / L) * L;
    if (w >= L) w -= L;
    if (w < 0) w = 0;
    return w;
}

void pbo_real_velocity_cutoff(double *u, const double *y, int nt, const double *x,
                              const double *f, int ns, const double *L, double a, double d,
                              double e, double rc)
{
    int nc[3], st[3];
    double h[3];
    for (int c = 0; c < 3; c++) {
        nc[c] = (int)floor(L[c] / rc);
        if (nc[c] < 1) nc[c] = 1;
        if (nc[c] > 256) nc[c] = 256;
        h[c] = L[c] / nc[c];
        st[c] = (int)ceil(rc / h[c]);
    }
    long ncell = (long)nc[0] * nc[1] * nc[2];
    int *head = malloc(sizeof(int) * ncell), *next = malloc(sizeof(int) * (ns > 0 ? ns : 1));
    double *xw = malloc(sizeof(double) * 3 * (ns > 0 ? ns : 1));
    for (long c = 0; c < ncell; c++) head[c] = -1;
    for (int j = ns - 1; j >= 0; j--) {
        int ci[3];
        for (int c = 0; c < 3; c++) {
            xw[3 * j + c] = wrap_coord(x[3 * j + c], L[c]);
            ci[c] = (int)(xw[3 * j + c] / h[c]);
            if (ci[c] >= nc[c]) ci[c] = nc[c] - 1;
        }
        long cell = ((long)ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        next[j] = head[cell];
        head[cell] = j;
    }
    double rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nt; i++) {
        double yw[3], acc[3] = {0, 0, 0}, xh[3], tA[9];
        int ci[3];
        for (int c = 0; c < 3; c++) {
            yw[c] = wrap_coord(y[3 * i + c], L[c]);
            ci[c] = (int)(yw[c] / h[c]);
            if (ci[c] >= nc[c]) ci[c] = nc[c] - 1;
        }
        for (int o0 = -st[0]; o0 <= st[0]; o0++) {
            int u0 = ci[0] + o0, w0 = ((u0 % nc[0]) + nc[0]) % nc[0];
            double s0 = (double)((u0 - w0) / nc[0]) * L[0];
            for (int o1 = -st[1]; o1 <= st[1]; o1++) {
                int u1 = ci[1] + o1, w1 = ((u1 % nc[1]) + nc[1]) % nc[1];
                double s1 = (double)((u1 - w1) / nc[1]) * L[1];
                for (int o2 = -st[2]; o2 <= st[2]; o2++) {
                    int u2 = ci[2] + o2, w2 = ((u2 % nc[2]) + nc[2]) % nc[2];
                    double s2 = (double)((u2 - w2) / nc[2]) * L[2];
                    long cell = ((long)w0 * nc[1] + w1) * nc[2] + w2;
                    for (int j = head[cell]; j >= 0; j = next[j]) {
                        xh[0] = yw[0] - (xw[3 * j + 0] + s0);
                        xh[1] = yw[1] - (xw[3 * j + 1] + s1);
                        xh[2] = yw[2] - (xw[3 * j + 2] + s2);
                        double r2 = xh[0] * xh[0] + xh[1] * xh[1] + xh[2] * xh[2];
                        if (r2 >= rc2) continue;
                        pbo_vR(tA, xh, a, d, e);
                        apply_tensor(acc, tA, f + 3 * j);
                    }
                }
            }
        }
        for (int c = 0; c < 3; c++) u[3 * i + c] = acc[c];
    }
    free(head);
    free(next);
    free(xw);
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, exact twiddle table.      */
/* ------------------------------------------------------------------------ */
typedef struct {
    double re, im;
} cplx;

static void fft_rec(const cplx *in, cplx *out, int n, int istride, const cplx *w, int wstride,
                    cplx *tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = 2;
    while (n % p) p++;
    int m = n / p;
    for (int r = 0; r < p; r++) fft_rec(in + (size_t)r * istride, tmp + (size_t)r * m, m, istride * p, w, wstride * p, out);
    /* out[k + q m] = sum_r w^{r (k + q m)} tmp[r m + k]; w has period n*wstride */
    int N = n * wstride;
    for (int k = 0; k < m; k++)
        for (int q = 0; q < p; q++) {
            double sr = 0, si = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * (k + q * m)) % n;
                const cplx tw = w[(e * wstride) % N];
                const cplx v = tmp[(size_t)r * m + k];
                sr += tw.re * v.re - tw.im * v.im;
                si += tw.re * v.im + tw.im * v.re;
            }
            out[k + q * m].re = sr;
            out[k + q * m].im = si;
        }
}

static void fft1d(cplx *data, int n, int stride, const cplx *w, cplx *line, cplx *res, cplx *tmp)
{
    for (int i = 0; i < n; i++) line[i] = data[(size_t)i * stride];
    fft_rec(line, res, n, 1, w, 1, tmp);
    for (int i = 0; i < n; i++) data[(size_t)i * stride] = res[i];
}

void pbo_fft3d(double *data_, const int *n, int sign)
{
    cplx *data = (cplx *)data_;
    for (int ax = 0; ax < 3; ax++) {
        int len = n[ax];
        cplx *w = malloc(sizeof(cplx) * len);
        for (int j = 0; j < len; j++) {
            long double ang = 2 * (long double)PI * j / len;
            w[j].re = (double)cosl(ang);
            w[j].im = (double)(sign * sinl(ang));
        }
        size_t stride = ax == 0 ? (size_t)n[1] * n[2] : (ax == 1 ? (size_t)n[2] : 1);
        size_t nlines = (size_t)n[0] * n[1] * n[2] / len;
#pragma omp parallel
        {
            cplx *line = malloc(sizeof(cplx) * len * 3);
            cplx *res = line + len, *tmp = line + 2 * len;
#pragma omp for schedule(static)
            for (size_t li = 0; li < nlines; li++) {
                size_t base;
                if (ax == 0) base = li;                                          /* li in [0, n1*n2) */
                else if (ax == 1) base = (li / n[2]) * (size_t)n[1] * n[2] + li % n[2];
                else base = li * (size_t)n[2];
                fft1d(data + base, len, (int)stride, w, line, res, tmp);
            }
            free(line);
        }
        free(w);
    }
}

/* ------------------------------------------------------------------------ */
/* Spectral k-space (fftBrinkman.h:34-287, with the fixes listed above)     */
/* ------------------------------------------------------------------------ */
static void window_weights(double xc, double h, double inv_h, int P, double eta, int *first,
                           double *w)
{
    /* centred window: P grid points g = first .. first+P-1 with
     * first = floor(x*(M/L) - P/2) + 1 (fftBrinkman.h:122-138 with D3 fixed).
     * The product computes the identical expression so both pick the same
     * window for points on a cell boundary. */
    double s = xc * inv_h;
    int f0 = (int)floor(s - 0.5 * P) + 1;
    *first = f0;
    for (int p = 0; p < P; p++) {
        double dx = (f0 + p) * h - xc;
        w[p] = exp(-dx * dx / eta);
    }
}

void pbo_kspace_spectral(double *u, const double *y, int nt, const double *x, const double *f,
                         int ns, const double *L, double a, double d, const int *grid,
                         const int *P, double m)
{
    int M0 = grid[0], M1 = grid[1], M2 = grid[2];
    size_t Ng = (size_t)M0 * M1 * M2;
    double h[3] = {L[0] / M0, L[1] / M1, L[2] / M2};
    double ih[3] = {M0 / L[0], M1 / L[1], M2 / L[2]};
    double eta = P[0] * h[0] / m;
    eta = eta * eta / 2; /* fftBrinkman.h:113-114 */
    double scale = 1.0 / (eta * PI);
    scale = scale * sqrt(scale);
    double *H[3];
    for (int c = 0; c < 3; c++) H[c] = calloc(2 * Ng, sizeof(double));
    int Pmax = P[0] > P[1] ? (P[0] > P[2] ? P[0] : P[2]) : (P[1] > P[2] ? P[1] : P[2]);
    /* spread (fftBrinkman.h:120-155).  Each thread owns a slab of x planes and
     * adds, source by source in order, the window points in its slab: every
     * grid point receives its sources in source order, so the sums are the
     * same as those of one sequential pass, whatever the thread count. */
#pragma omp parallel
    {
        const int nth = omp_get_num_threads(), th = omp_get_thread_num();
        const int x0 = (int)((long)M0 * th / nth), x1 = (int)((long)M0 * (th + 1) / nth);
        double *wx = malloc(sizeof(double) * 3 * Pmax), *wy = wx + Pmax, *wz = wx + 2 * Pmax;
        for (int n = 0; n < ns && x0 < x1; n++) {
            double xc[3];
            int f0[3];
            xc[0] = wrap_coord(x[3 * n], L[0]);
            /* the window's planes f0 .. f0 + P - 1 (mod M0): skip sources whose
             * window misses the slab before evaluating any weight */
            int lo = (int)floor(xc[0] * ih[0] - 0.5 * P[0]) + 1, hit = 0;
            for (int px = 0; px < P[0] && !hit; px++) {
                int jx = ((lo + px) % M0 + M0) % M0;
                hit = jx >= x0 && jx < x1;
            }
            if (!hit) continue;
            for (int c = 1; c < 3; c++) xc[c] = wrap_coord(x[3 * n + c], L[c]);
            window_weights(xc[0], h[0], ih[0], P[0], eta, &f0[0], wx);
            window_weights(xc[1], h[1], ih[1], P[1], eta, &f0[1], wy);
            window_weights(xc[2], h[2], ih[2], P[2], eta, &f0[2], wz);
            for (int px = 0; px < P[0]; px++) {
                int jx = ((f0[0] + px) % M0 + M0) % M0;
                if (jx < x0 || jx >= x1) continue;
                for (int py = 0; py < P[1]; py++) {
                    int jy = ((f0[1] + py) % M1 + M1) % M1;
                    double wxy = scale * wx[px] * wy[py];
                    for (int pz = 0; pz < P[2]; pz++) {
                        int jz = ((f0[2] + pz) % M2 + M2) % M2;
                        double wt = wxy * wz[pz];
                        size_t idx = ((size_t)jx * M1 + jy) * M2 + jz;
                        for (int c = 0; c < 3; c++) H[c][2 * idx] += f[3 * n + c] * wt;
                    }
                }
            }
        }
        free(wx);
    }
    for (int c = 0; c < 3; c++) pbo_fft3d(H[c], grid, -1);
    /* multiply by the Brinkman projection with the deconvolution e^{eta k^2/2}
     * (fftBrinkman.h:185-219; the 1/N_g of :167-172 is folded in). */
#pragma omp parallel for schedule(static)
    for (int q0 = 0; q0 < M0; q0++) {
        double k[3];
        int f0i = q0 < (M0 + 1) / 2 ? q0 : q0 - M0;
        if (2 * q0 == M0) f0i = -M0 / 2;
        k[0] = 2 * PI * f0i / L[0];
        for (int q1 = 0; q1 < M1; q1++) {
            int f1i = q1 < (M1 + 1) / 2 ? q1 : q1 - M1;
            if (2 * q1 == M1) f1i = -M1 / 2;
            k[1] = 2 * PI * f1i / L[1];
            for (int q2 = 0; q2 < M2; q2++) {
                int f2i = q2 < (M2 + 1) / 2 ? q2 : q2 - M2;
                if (2 * q2 == M2) f2i = -M2 / 2;
                k[2] = 2 * PI * f2i / L[2];
                size_t idx = ((size_t)q0 * M1 + q1) * M2 + q2;
                double kn = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                if (kn == 0) {
                    for (int c = 0; c < 3; c++) H[c][2 * idx] = H[c][2 * idx + 1] = 0;
                    continue;
                }
                double b = pbo_blob_hat(kn, a, d) * exp(eta * kn / 2) / (kn * (kn + a * a)) / (double)Ng;
                for (int ri = 0; ri < 2; ri++) {
                    double v[3] = {H[0][2 * idx + ri], H[1][2 * idx + ri], H[2][2 * idx + ri]};
                    double kv = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
                    for (int c = 0; c < 3; c++) H[c][2 * idx + ri] = b * (kn * v[c] - k[c] * kv);
                }
            }
        }
    }
    for (int c = 0; c < 3; c++) pbo_fft3d(H[c], grid, +1);
    /* gather over the window only (fftBrinkman.h:231-273 with D4 fixed) */
    double vol = h[0] * h[1] * h[2];
#pragma omp parallel
    {
        double *gx = malloc(sizeof(double) * 3 * Pmax), *gy = gx + Pmax, *gz = gx + 2 * Pmax;
#pragma omp for schedule(static)
        for (int i = 0; i < nt; i++) {
            double yc[3], acc[3] = {0, 0, 0};
            int f0[3];
            for (int c = 0; c < 3; c++) yc[c] = wrap_coord(y[3 * i + c], L[c]);
            window_weights(yc[0], h[0], ih[0], P[0], eta, &f0[0], gx);
            window_weights(yc[1], h[1], ih[1], P[1], eta, &f0[1], gy);
            window_weights(yc[2], h[2], ih[2], P[2], eta, &f0[2], gz);
            for (int px = 0; px < P[0]; px++) {
                int jx = ((f0[0] + px) % M0 + M0) % M0;
                for (int py = 0; py < P[1]; py++) {
                    int jy = ((f0[1] + py) % M1 + M1) % M1;
                    double wxy = scale * gx[px] * gy[py];
                    for (int pz = 0; pz < P[2]; pz++) {
                        int jz = ((f0[2] + pz) % M2 + M2) % M2;
                        double wt = wxy * gz[pz];
                        size_t idx = ((size_t)jx * M1 + jy) * M2 + jz;
                        for (int c = 0; c < 3; c++) acc[c] += wt * H[c][2 * idx];
                    }
                }
            }
            for (int c = 0; c < 3; c++) u[3 * i + c] = acc[c] * vol;
        }
        free(gx);
    }
    for (int c = 0; c < 3; c++) free(H[c]);
}

int pbo_fftBrinkman(double *U, const double *y, int m, const double *f, const double *xn, int N,
                    const double *L, const int *M, const int *P, int msd, int C, double e,
                    double tol, double d, double alpha)
{
    (void)C;
    (void)e;
    (void)tol;
    pbo_kspace_spectral(U, y, m, xn, f, N, L, alpha, d, M, P, (double)msd);
    return 0;
}

int pbo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
