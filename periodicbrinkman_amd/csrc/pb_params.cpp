// pb_params.cpp -- host-side constants for the device kernels, the parameter
// tuner and the deterministic input generator.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "pb_internal.h"

namespace {

const long double kPi = 3.141592653589793238462643383279502884L;

// a_n = A_n / (Gamma(n+1/2)/2), A_n(x) = int_0^inf t^{2n} e^{-t^2}/(t^2+x^2) dt.
// Forward recursion a_{n+1} = (1 - x^2 a_n)/(n + 1/2) from a_0 = sqrt(pi) erfcx(x)/x
// is stable while x^2 stays small against n; for x > 2 the integrals are
// evaluated by 16-point Gauss-Legendre on panels of width 1/4.
void scaled_moments(long double x, int count, long double *an)
{
    long double x2 = x * x;
    if (x <= 2) {
        an[0] = sqrtl(kPi) * expl(x2) * erfcl(x) / x;
        for (int n = 0; n + 1 < count; n++) an[n + 1] = (1 - x2 * an[n]) / (n + 0.5L);
        return;
    }
    static long double gx[16], gw[16];
    static bool init = false;
    if (!init) {
        const int q = 16;
        for (int i = 0; i < q; i++) {
            long double t = cosl(kPi * (i + 0.75L) / (q + 0.5L));
            for (int it = 0; it < 60; it++) {
                long double p0 = 1, p1 = t;
                for (int k = 2; k <= q; k++) {
                    long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                long double dp = q * (t * p1 - p0) / (t * t - 1);
                t -= p1 / dp;
                if (it > 50) {
                    gx[i] = t;
                    gw[i] = 2 / ((1 - t * t) * dp * dp);
                }
            }
        }
        init = true;
    }
    long double G = sqrtl(kPi) / 2;
    for (int n = 0; n < count; n++) {
        long double T = sqrtl((long double)n + 1) + 10;
        int panels = (int)(T * 4) + 1;
        long double hp = T / panels, sum = 0;
        for (int p = 0; p < panels; p++)
            for (int i = 0; i < 16; i++) {
                long double t = p * hp + 0.5L * hp * (gx[i] + 1), t2 = t * t;
                sum += 0.5L * hp * gw[i] * powl(t2, n) * expl(-t2) / (t2 + x2);
            }
        an[n] = sum / G;
        G *= (n + 0.5L);
    }
}

} // namespace

// Taylor coefficients of regH in z = r/d (H1 = sum c1 z^{2m}, H2 = sum c2 z^{2m}):
// with I_n = int_0^inf phi_d(k) k^{2n}/(k^2 + a^2) dk,
//   H1 = sum_m (-1)^m (2m+2)/((2m+3)(2m+1)!) I_{m+1} r^{2m}/(2 pi^2),
//   H2 = sum_m (-1)^m 4(m+1)(m+2)/(2m+5)!  I_{m+2} r^{2m}/(2 pi^2),
// and I_n = (d/2)^{1-2n} [A_n + gam A_{n+1}], gam = (1 - e^{-x^2})/x^2.
void pb_fill_radial(pb_radial *r, double a, double d)
{
    std::memset(r, 0, sizeof(*r));
    r->a = a;
    r->d = d;
    if (!(d > 0)) return;
    r->inv_d = 1.0 / d;
    r->x = a * d / 2;
    r->ex2 = std::exp(-r->x * r->x);
    double om = -std::expm1(-r->x * r->x);
    r->k4pa2 = 1.0 / (4 * M_PI * a * a);
    r->kg = om / (2 * a * a * d * d * d * M_PI * std::sqrt(M_PI));
    double zs = 0.32 / std::sqrt(r->x);
    r->zs = zs < 0.2 ? 0.2 : (zs > 2.0 ? 2.0 : zs);

    const int NA = PB_NSER + 4;
    long double x = (long double)a * d / 2, x2 = x * x;
    long double gam = -expm1l(-x2) / x2;
    long double an[PB_NSER + 4], Gn[PB_NSER + 4];
    scaled_moments(x, NA, an);
    Gn[0] = sqrtl(kPi) / 2;
    for (int n = 1; n < NA; n++) Gn[n] = Gn[n - 1] * (n - 0.5L);
    long double pi2 = 2 * kPi * kPi, fact = 1;
    for (int m = 0; m < PB_NSER; m++) {
        if (m > 0) fact *= (2 * m) * (2 * m + 1);
        long double sgn = (m & 1) ? -1 : 1;
        long double p2m1 = ldexpl(1, 2 * m + 1);
        long double I1 = Gn[m + 1] * (an[m + 1] + gam * (m + 1.5L) * an[m + 2]);
        r->c1[m] = (double)(sgn * (2 * m + 2) * p2m1 * I1 / ((2 * m + 3) * fact * pi2 * d));
        long double f5 = fact * (2 * m + 2) * (2 * m + 3) * (2 * m + 4) * (2 * m + 5);
        long double I2 = Gn[m + 2] * (an[m + 2] + gam * (m + 2.5L) * an[m + 3]);
        r->c2[m] = (double)(sgn * 16 * (m + 1) * (m + 2) * p2m1 * I2 / (f5 * pi2 * d * d * d));
    }
}

// ---- radial table ------------------------------------------------------------
namespace {

// diffH(d) closed form (periodicBrinkman.h:95-104) in long double; used only
// to build the table, on z >= z0 where the closed form loses at most a few
// digits of the 64-bit mantissa.
void diffH_ld(long double r, long double a, long double d, long double &H1, long double &H2)
{
    const long double pi = kPi, spi = sqrtl(kPi);
    long double r2 = r * r, z = r / d, x = a * d / 2, s = a * r;
    long double om = -expm1l(-x * x);
    long double E = expl(-z * z) * om;
    long double Erfc = erfcl(z);
    long double tm = expl(-s) * erfcl(z - x);
    long double tp = (z + x) < 27 ? expl(s) * erfcl(z + x) : 0.0L;
    long double c3 = 1 / (8 * pi * a * a * r2 * r), c5 = c3 / r2;
    H1 = -Erfc / (4 * pi * a * a * r2 * r) - (d * d + 2 * r2) / (2 * a * a * d * d * d * pi * spi * r2) * E +
         c3 * (1 + s + s * s) * tm + c3 * (1 - s + s * s) * tp;
    H2 = 3 * Erfc / (4 * pi * a * a * r2 * r2 * r) + (3 * d * d + 2 * r2) / (2 * a * a * d * d * d * pi * spi * r2 * r2) * E -
         c5 * (3 + 3 * s + s * s) * tm - c5 * (3 - 3 * s + s * s) * tp;
}

// Chebyshev interpolation of f on [-1, 1] at the deg+1 Chebyshev points, then
// conversion to monomial coefficients in t (well conditioned here: the
// coefficients decay geometrically on the short intervals used).
void cheb_to_monomial(const long double *cheb, int deg, long double *mono)
{
    long double Tprev[PB_TAB_MAXDEG + 1] = {0}, Tcur[PB_TAB_MAXDEG + 1] = {0}, Tnext[PB_TAB_MAXDEG + 1];
    for (int k = 0; k <= deg; k++) mono[k] = 0;
    Tprev[0] = 1; // T0
    mono[0] += cheb[0];
    if (deg >= 1) {
        Tcur[1] = 1; // T1 = t
        mono[1] += cheb[1];
    }
    for (int n = 2; n <= deg; n++) {
        for (int k = 0; k <= deg; k++) Tnext[k] = -Tprev[k] + (k > 0 ? 2 * Tcur[k - 1] : 0);
        for (int k = 0; k <= deg; k++) {
            mono[k] += cheb[n] * Tnext[k];
            Tprev[k] = Tcur[k];
            Tcur[k] = Tnext[k];
        }
    }
}

} // namespace

double pb_build_table(pb_kparams *k, double z0, double z1, std::vector<double> &coef, int deg, double tol)
{
    const long double a = k->pd.a, d = k->pd.d;
    if (deg < 1 || deg > PB_TAB_MAXDEG) deg = PB_TAB_DEG;
    // Chebyshev nodes and the cosine matrix of the coefficient transform, per degree
    struct cheb_consts {
        long double node[PB_TAB_MAXDEG + 1], cosm[PB_TAB_MAXDEG + 1][PB_TAB_MAXDEG + 1];
        explicit cheb_consts(int n)
        {
            for (int j = 0; j <= n; j++) node[j] = cosl(kPi * (j + 0.5L) / (n + 1));
            for (int q = 0; q <= n; q++)
                for (int j = 0; j <= n; j++) cosm[q][j] = cosl(kPi * q * (j + 0.5L) / (n + 1));
        }
    };
    static const cheb_consts *const CCs[PB_TAB_MAXDEG + 1] = {
        nullptr, new cheb_consts(1), new cheb_consts(2), new cheb_consts(3), new cheb_consts(4),
        new cheb_consts(5), new cheb_consts(6), new cheb_consts(7), new cheb_consts(8), new cheb_consts(9),
        new cheb_consts(10), new cheb_consts(11), new cheb_consts(12), new cheb_consts(13), new cheb_consts(14),
        new cheb_consts(15)};
    const cheb_consts &CC = *CCs[deg];
    double err = 1;
    for (int per_unit = 8; per_unit <= 256; per_unit *= 2) {
        int n = (int)std::ceil((z1 - z0) * per_unit);
        if (n < 1) n = 1;
        coef.assign((size_t)n * (deg + 1) * 2, 0.0);
        double worst = 0;
        for (int i = 0; i < n; i++) {
            long double zl = z0 + (long double)i / per_unit, hw = 0.5L / per_unit, zc = zl + hw;
            long double f1[PB_TAB_MAXDEG + 1], f2[PB_TAB_MAXDEG + 1];
            for (int j = 0; j <= deg; j++) diffH_ld((zc + hw * CC.node[j]) * d, a, d, f1[j], f2[j]);
            long double c1[PB_TAB_MAXDEG + 1], c2[PB_TAB_MAXDEG + 1], m1[PB_TAB_MAXDEG + 1], m2[PB_TAB_MAXDEG + 1];
            for (int q = 0; q <= deg; q++) {
                long double s1 = 0, s2 = 0;
                for (int j = 0; j <= deg; j++) {
                    s1 += f1[j] * CC.cosm[q][j];
                    s2 += f2[j] * CC.cosm[q][j];
                }
                c1[q] = s1 * (q == 0 ? 1.0L : 2.0L) / (deg + 1);
                c2[q] = s2 * (q == 0 ? 1.0L : 2.0L) / (deg + 1);
            }
            cheb_to_monomial(c1, deg, m1);
            cheb_to_monomial(c2, deg, m2);
            for (int q = 0; q <= deg; q++) {
                coef[((size_t)i * (deg + 1) + q) * 2 + 0] = (double)m1[q];
                coef[((size_t)i * (deg + 1) + q) * 2 + 1] = (double)m2[q];
            }
            // verify in double Horner (what the kernel does) against long double
            long double mx1 = 0, mx2 = 0, e1 = 0, e2 = 0;
            for (int j = 0; j <= 16; j++) {
                double t = -1.0 + 2.0 * j / 16;
                double h1 = 0, h2 = 0;
                for (int q = deg; q >= 0; q--) {
                    h1 = h1 * t + coef[((size_t)i * (deg + 1) + q) * 2 + 0];
                    h2 = h2 * t + coef[((size_t)i * (deg + 1) + q) * 2 + 1];
                }
                long double g1, g2;
                diffH_ld((zc + hw * t) * d, a, d, g1, g2);
                mx1 = fmaxl(mx1, fabsl(g1));
                mx2 = fmaxl(mx2, fabsl(g2));
                e1 = fmaxl(e1, fabsl(h1 - g1));
                e2 = fmaxl(e2, fabsl(h2 - g2));
            }
            double ei = (double)fmaxl(mx1 > 0 ? e1 / mx1 : 0, mx2 > 0 ? e2 / mx2 : 0);
            worst = std::fmax(worst, ei);
        }
        err = worst;
        if ((size_t)n * (deg + 1) * 2 > PB_TAB_LDS_DOUBLES) break; // too large: no table
        if (worst <= tol) {
            k->tz0 = z0;
            k->tinv = per_unit;
            k->tn = n;
            k->tdeg = deg;
            return err;
        }
    }
    k->tn = 0;
    coef.clear();
    return err;
}

void pb_fill_kparams(pb_kparams *k, double a, double d, double e, double rfar)
{
    std::memset(k, 0, sizeof(*k));
    pb_fill_radial(&k->pd, a, d);
    k->has_e = e > 0;
    if (k->has_e) pb_fill_radial(&k->pe, a, e);
    k->e8 = 8 * e;
    k->d3 = 3 * d;
    k->rfar2 = rfar * rfar;
    k->rcommon = k->pd.zs * d;
    if (k->has_e && 8 * e > k->rcommon) k->rcommon = 8 * e;
    k->pf.a = (float)a;
    k->pf.inv_d = (float)(1.0 / d);
    k->pf.x = (float)k->pd.x;
    k->pf.ex2 = (float)k->pd.ex2;
    k->pf.k4pa2 = (float)k->pd.k4pa2;
    k->pf.kg = (float)k->pd.kg;
    k->pf.d2 = (float)(d * d);
}

void pb_fill_box(pb_box *b, const double *L)
{
    for (int i = 0; i < 3; i++) {
        b->L[i] = L[i];
        b->inv_L[i] = 1.0 / L[i];
    }
}

double pb_window_m(int P)
{
    return std::sqrt(std::fmax(M_PI * P - 5.0, 0.5 * M_PI * P));
}

namespace {
double window_m2(int P)
{
    const double m = pb_window_m(P);
    return m * m;
}
double kspace_X(double lt)
{
    return std::fmax(lt - 7.5, 0.4 * lt);
}
} // namespace

extern "C" {

void pb_params_default(pb_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->L[0] = p->L[1] = p->L[2] = 1.0;
    p->alpha = 1.0;
    p->P = 12;
    p->precision = PB_PREC_FP64;
    p->tol = 1e-6;
}

// Error model (relative L2 of the velocity), calibrated against converged
// sums of the FP64 oracle (1e5 uniform points on 128^3 with P 12, 512 uneven
// systems, 1e6 clustered points on 256^3 with P 16, tolerances 4e-6 to 1e-11).
// Real space: truncation ~ e^{-(rc/d)^2}, rc = d sqrt(ln(1/tol) + 1).  k space:
// spreading and gathering each multiply mode k by e^{-eta k^2/4} and the
// multiply divides both out, so a dropped mode carries e^{-(d^2/4 - eta/2) k^2};
// at the Nyquist wavenumber that is e^{-X}, X = (d^2/4 - eta/2) k_max^2.  The
// velocity sits mostly at small k: the dropped modes' share of it measured
// e^{-(X + 7.5)} or less, so X = ln(1/tol) - 7.5, and never below 0.4 ln(1/tol)
// (at loose tolerances the error falls off slower than e^{-X}).  Window: the
// Gaussian cut at the window's edge leaves e^{-m^2/2}; aliasing, which the
// deconvolution amplifies, falls with smaller m, and the measured optimum is
// m^2 = pi P - 5 (the unamplified balance would be m^2 = pi P).  The grid
// (when not given) gives about two grid points per particle, rounded up to a
// power of two per axis.

int pb_tune(pb_params *p, int ns, int nt)
{
    if (!p || !(p->alpha > 0) || !(p->L[0] > 0) || !(p->L[1] > 0) || !(p->L[2] > 0) || p->P < 2)
        return PB_ERR_INVALID;
    double tol = p->tol > 0 ? p->tol : 1e-6;
    double lt = std::log(1.0 / tol);
    double Lmin = std::fmin(p->L[0], std::fmin(p->L[1], p->L[2]));
    if (p->grid[0] <= 0 || p->grid[1] <= 0 || p->grid[2] <= 0) {
        double vol = p->L[0] * p->L[1] * p->L[2];
        double npart = ns > nt ? ns : nt;
        double hh = std::cbrt(vol / (2.0 * (npart > 1 ? npart : 1)));
        for (int i = 0; i < 3; i++) {
            int g = 16;
            while (g < 512 && p->L[i] / g > hh) g *= 2;
            p->grid[i] = g;
        }
    }
    if (!(p->m > 0)) p->m = std::sqrt(window_m2(p->P));
    if (!(p->d > 0)) {
        double kmax = 1e300;
        for (int i = 0; i < 3; i++) kmax = std::fmin(kmax, M_PI * p->grid[i] / p->L[i]);
        double w = p->P * (p->L[0] / p->grid[0]) / p->m;
        double eta = w * w / 2;
        p->d = 2.0 * std::sqrt(kspace_X(lt) / (kmax * kmax) + eta / 2);
    }
    if (!(p->rc > 0)) p->rc = p->d * std::sqrt(lt + 1.0);
    if (p->rc > Lmin * 3) p->rc = Lmin * 3;
    return PB_OK;
}

// ---- autotuner -----------------------------------------------------------------
// Cost model of this build on MI355X (measured at 1e5/1e5, 128^3, P 12; per
// unit of work): pair kernel per (target, source image) inside rc, spreading
// and gathering per window point, FFT per grid point of the five passes, and
// fixed costs (binning, launches).  FP64 k space costs about twice FP32.
namespace {
struct cost_model {
    double pair_s;    // per pair-image
    double spread_s;  // per source window point
    double gather_s;  // per target window point
    double fft_s;     // per grid point (five passes)
    double fixed_s;   // binning and launches
    double bin_s;     // per point binned
};
cost_model model_for(int precision)
{
    if (precision == PB_PREC_MIXED) return {9.9e-12, 1.23e-12, 0.83e-12, 5.8e-11, 6.0e-5, 3.0e-10};
    return {9.8e-12, 3.1e-12, 1.8e-12, 8.6e-11, 6.0e-5, 3.0e-10};
}
} // namespace

int pb_autotune(pb_params *p, int ns, int nt, pb_tune_report *rep)
{
    if (!p || !(p->alpha > 0) || !(p->L[0] > 0) || !(p->L[1] > 0) || !(p->L[2] > 0) || ns < 0 || nt < 0)
        return PB_ERR_INVALID;
    const double tol = p->tol > 0 ? p->tol : 1e-6;
    const double lt = std::log(1.0 / tol);
    const double Lmin = std::fmin(p->L[0], std::fmin(p->L[1], p->L[2]));
    const double vol = p->L[0] * p->L[1] * p->L[2];
    const cost_model cm = model_for(p->precision);
    static const int Ps[] = {4, 6, 8, 10, 12, 16, 20, 24};
    double best = 1e300;
    pb_params bestp = *p;
    pb_tune_report br{};
    int scored = 0;
    for (int P : Ps) {
        const double m = std::sqrt(window_m2(P));
        const double ew = std::exp(-m * m / 2);
        if (ew > tol) continue;
        // grids: one spacing h for all axes, each axis rounded up to a power of two
        for (int k = 2; k <= 9; k++) {
            const double h = Lmin / (1 << k);
            int g[3];
            bool ok = true;
            for (int i = 0; i < 3; i++) {
                int n = 4;
                while (n < 512 && p->L[i] / n > h * (1 + 1e-12)) n *= 2;
                if (p->L[i] / n > h * (1 + 1e-12) || P > n) ok = false;
                g[i] = n;
            }
            if (!ok) continue;
            double kmax = 1e300, hx = p->L[0] / g[0];
            for (int i = 0; i < 3; i++) kmax = std::fmin(kmax, M_PI * g[i] / p->L[i]);
            const double w = P * hx / m, eta = w * w / 2;
            const double d = 2.0 * std::sqrt(kspace_X(lt) / (kmax * kmax) + eta / 2);
            const double rc = d * std::sqrt(lt + 1.0);
            if (rc > 3 * Lmin) continue;
            const double pairs = (double)nt * ns / vol * (4.0 / 3.0) * M_PI * rc * rc * rc;
            const double ng = (double)g[0] * g[1] * g[2];
            const double P3 = (double)P * P * P;
            const double t_real = pairs * cm.pair_s;
            const double t_k = ns * P3 * cm.spread_s + nt * P3 * cm.gather_s + ng * cm.fft_s;
            const double t_bin = cm.fixed_s + (ns + nt) * cm.bin_s;
            const double t = t_real + t_k + t_bin;
            scored++;
            if (t < best) {
                best = t;
                bestp = *p;
                bestp.grid[0] = g[0];
                bestp.grid[1] = g[1];
                bestp.grid[2] = g[2];
                bestp.P = P;
                bestp.m = m;
                bestp.d = d;
                bestp.rc = rc;
                br.err_real = std::exp(-(rc / d) * (rc / d));
                br.err_kspace = std::exp(-((d * d / 4 - eta / 2) * kmax * kmax + 7.5));
                br.err_window = ew;
                br.pairs = pairs;
                br.t_real_ms = t_real * 1e3;
                br.t_kspace_ms = t_k * 1e3;
                br.t_bin_ms = t_bin * 1e3;
                br.t_total_ms = t * 1e3;
            }
        }
    }
    if (!scored) return PB_ERR_UNSUPPORTED;
    br.candidates = scored;
    *p = bestp;
    if (rep) *rep = br;
    return PB_OK;
}

// ---- the reference driver's parameter file ----------------------------------
void pb_driver_defaults(pb_driver_params *p)
{
    p->Lx = p->Ly = p->Lz = 1.0;
    p->d = 0.6;
    p->e = 0.0;
    p->alpha = 1.0;
    p->M = 4;
    p->boxSize = 1;
    p->toPrint = 0;
}

int pb_read_param_file(const char *path, pb_driver_params *p)
{
    if (!path || !p) return PB_ERR_INVALID;
    pb_driver_defaults(p);
    FILE *fh = std::fopen(path, "r");
    if (!fh) return PB_ERR_INVALID;
    char line[1024];
    while (std::fgets(line, sizeof line, fh)) {
        // tokens separated by ':' , ' ' and line ends; the first is the key,
        // the second the value
        const char *sep = ": \t\r\n";
        char *save = nullptr;
        char *key = strtok_r(line, sep, &save);
        if (!key) continue;
        char *val = strtok_r(nullptr, sep, &save);
        if (!val) continue;
        const double v = std::atof(val);
        if (!std::strcmp(key, "Lx")) p->Lx = v;
        else if (!std::strcmp(key, "Ly")) p->Ly = v;
        else if (!std::strcmp(key, "Lz")) p->Lz = v;
        else if (!std::strcmp(key, "d")) p->d = v;
        else if (!std::strcmp(key, "e")) p->e = v;
        else if (!std::strcmp(key, "alpha")) p->alpha = v;
        else if (!std::strcmp(key, "M")) p->M = (int)v;
        else if (!std::strcmp(key, "boxSize")) p->boxSize = (int)v;
        else if (!std::strcmp(key, "toPrint")) p->toPrint = (int)v;
    }
    std::fclose(fh);
    return PB_OK;
}

// ---- xoshiro256** seeded by SplitMix64 -------------------------------------
static inline unsigned long long rotl(unsigned long long x, int k)
{
    return (x << k) | (x >> (64 - k));
}

void pb_rng_seed(unsigned long long *s, unsigned long long seed)
{
    unsigned long long z = seed;
    for (int i = 0; i < 4; i++) {
        z += 0x9E3779B97F4A7C15ULL;
        unsigned long long t = z;
        t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ULL;
        t = (t ^ (t >> 27)) * 0x94D049BB133111EBULL;
        s[i] = t ^ (t >> 31);
    }
}

static inline unsigned long long next64(unsigned long long *s)
{
    unsigned long long r = rotl(s[1] * 5, 7) * 9;
    unsigned long long t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

void pb_rng_uniform(unsigned long long *s, long n, double *out, double scale)
{
    for (long i = 0; i < n; i++) out[i] = (double)(next64(s) >> 11) * 0x1.0p-53 * scale;
}

void pb_rng_normal(unsigned long long *s, long n, double *out)
{
    for (long i = 0; i < n; i += 2) {
        double u1 = (double)(next64(s) >> 11) * 0x1.0p-53;
        double u2 = (double)(next64(s) >> 11) * 0x1.0p-53;
        double r = std::sqrt(-2.0 * std::log(1.0 - u1));
        double t = 2.0 * M_PI * u2;
        out[i] = r * std::cos(t);
        if (i + 1 < n) out[i + 1] = r * std::sin(t);
    }
}

} // extern "C"
