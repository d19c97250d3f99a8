// pb_radial.h -- device evaluation of the Brinkmanlet radial functions H1, H2.
//
// Mathematics of periodicBrinkman.h:19-176 (singH, regH, diffH, vR).  The
// closed forms contain e^{-+ar} erfc(r/d -+ ad/2) products; with
// z = r/d, x = ad/2, s = ar = 2xz they are rewritten through the scaled
// complementary error function erfcx(u) = e^{u^2} erfc(u):
//   erfc(z)               = E erfcx(z)
//   e^{ s} erfc(z + x)    = E e^{-x^2} erfcx(z + x)
//   e^{-s} erfc(z - x)    = E e^{-x^2} erfcx(z - x)                 (z >= x)
//                         = 2 e^{-s} - E e^{-x^2} erfcx(x - z)      (z <  x)
// with E = e^{-z^2}, so one exp and three erfcx replace three exp and three
// erfc per pair.  Near r = 0 the regularized functions use their Taylor series
// in z^2 (coefficients from moments of the blob transform, computed on the
// host) and the singular function its series in s; the regions are the same
// as in the FP64 oracle so both evaluate the same expressions.
#pragma once

#include "pb_internal.h"

namespace pbdev {

struct sing_coeffs {
    double p1[PB_NSING], p2[PB_NSING];
};

// p1[j] = (-1)^j (j+1)^2/(j+2)!,  p2[j] = (-1)^(j+1) (j+1)(j-1)/(j+2)!
constexpr sing_coeffs make_sing_coeffs()
{
    sing_coeffs c{};
    double f = 1;
    for (int j = 0; j < PB_NSING; j++) {
        f *= (j + 2);
        double sg = (j % 2) ? -1.0 : 1.0;
        c.p1[j] = sg * (j + 1) * (j + 1) / f;
        c.p2[j] = -sg * (j + 1) * (j - 1) / f;
    }
    return c;
}

__device__ __forceinline__ void sing_H(double r, double a, double k4pa2, double &H1, double &H2)
{
    constexpr sing_coeffs C = make_sing_coeffs();
    double s = a * r;
    if (s < 0.2) {
        double q1 = 0, q2 = 0;
#pragma unroll
        for (int j = PB_NSING - 1; j >= 0; j--) {
            q1 = fma(q1, s, C.p1[j]);
            q2 = fma(q2, s, C.p2[j]);
        }
        const double inv4pi = 0.0795774715459476678844418816862571887;
        H1 = a * q1 * inv4pi / s;
        H2 = a * a * a * q2 * inv4pi / (s * s * s);
    } else {
        double es = exp(-s);
        double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
        H1 = k4pa2 * ir3 * (-1.0 + es * fma(s, s + 1.0, 1.0));
        H2 = k4pa2 * ir3 * ir2 * (3.0 - es * fma(s, s + 3.0, 3.0));
    }
}

__device__ __forceinline__ void reg_series(double z, const pb_radial &P, double &H1, double &H2)
{
    double z2 = z * z, h1 = 0, h2 = 0;
    // kept rolled: unrolling batches all coefficient loads into registers
#pragma unroll 2
    for (int m = PB_NSER - 1; m >= 0; m--) {
        h1 = fma(h1, z2, P.c1[m]);
        h2 = fma(h2, z2, P.c2[m]);
    }
    H1 = h1;
    H2 = h2;
}

// regH closed form (periodicBrinkman.h:63-72) in erfcx form.
__device__ __forceinline__ void reg_direct(double r, double z, const pb_radial &P, double &H1,
                                           double &H2)
{
    double s = P.a * r, s2 = s * s, x = P.x;
    double E = exp(-z * z);
    double Erf = 1.0 - E * erfcx(z);
    double Ap = E * P.ex2 * erfcx(z + x);
    double Am = (x >= z) ? E * P.ex2 * erfcx(x - z) : 2.0 * exp(-s) - E * P.ex2 * erfcx(z - x);
    double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
    double d2 = P.d * P.d, r2 = r * r;
    double k8 = 0.5 * P.k4pa2;
    H1 = -Erf * P.k4pa2 * ir3 + P.kg * E * (d2 + 2.0 * r2) * ir2 + k8 * ir3 * ((1.0 + s + s2) * Am - (1.0 - s + s2) * Ap);
    H2 = 3.0 * Erf * P.k4pa2 * ir3 * ir2 - P.kg * E * (3.0 * d2 + 2.0 * r2) * ir2 * ir2
         + k8 * ir3 * ir2 * ((3.0 - 3.0 * s + s2) * Ap - (3.0 + 3.0 * s + s2) * Am);
}

// The scalars of pb_radial without its series tables: kernels copy these into
// registers once instead of re-reading the constant struct in their loops.
struct radial_core {
    double a, d, inv_d, x, ex2, k4pa2, kg;
};

__device__ __forceinline__ radial_core core_of(const pb_radial &P)
{
    return radial_core{P.a, P.d, P.inv_d, P.x, P.ex2, P.k4pa2, P.kg};
}

// diffH closed form (periodicBrinkman.h:95-104) in erfcx form.  R is
// pb_radial or radial_core.
template <class R>
__device__ __forceinline__ void diff_direct(double r, double z, const R &P, double &H1, double &H2)
{
    double s = P.a * r, s2 = s * s, x = P.x;
    double E = exp(-z * z);
    double T0 = erfcx(z);
    double Tp = P.ex2 * erfcx(z + x);
    double Tm, extra = 0.0;
    if (z >= x) {
        Tm = P.ex2 * erfcx(z - x);
    } else {
        Tm = -P.ex2 * erfcx(x - z);
        extra = 2.0 * exp(-s);
    }
    double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
    double d2 = P.d * P.d, r2 = r * r;
    double a1 = 1.0 + s + s2, a2 = 1.0 - s + s2;
    double b1 = 3.0 + 3.0 * s + s2, b2 = 3.0 - 3.0 * s + s2;
    double B1 = -T0 + 0.5 * (a1 * Tm + a2 * Tp);
    double B2 = 3.0 * T0 - 0.5 * (b1 * Tm + b2 * Tp);
    double k3 = P.k4pa2 * ir3;
    H1 = E * (B1 * k3 - P.kg * (d2 + 2.0 * r2) * ir2) + 0.5 * extra * a1 * k3;
    H2 = E * (B2 * k3 * ir2 + P.kg * (3.0 * d2 + 2.0 * r2) * ir2 * ir2) - 0.5 * extra * b1 * k3 * ir2;
}

__device__ __forceinline__ void reg_H(double r, const pb_radial &P, double &H1, double &H2)
{
    double z = r * P.inv_d;
    if (z < P.zs)
        reg_series(z, P, H1, H2);
    else
        reg_direct(r, z, P, H1, H2);
}

__device__ __forceinline__ void diff_H(double r, const pb_radial &P, double &H1, double &H2)
{
    double z = r * P.inv_d;
    if (r <= 0.0) { // the singular self term is dropped (periodicBrinkman.h:105-111)
        H1 = -P.c1[0];
        H2 = -P.c2[0];
    } else if (z < P.zs) {
        double s1, s2, g1, g2;
        sing_H(r, P.a, P.k4pa2, s1, s2);
        reg_series(z, P, g1, g2);
        H1 = s1 - g1;
        H2 = s2 - g2;
    } else {
        diff_direct(r, z, P, H1, H2);
    }
}

// Radial part of K_R = diffH(d) - diffH(e) (periodicBrinkman.h:152-168).
// Regions (the same as the oracle's):
//   e > 0, r < 8e and r < 3d : K = regH(e) - regH(d)  (singular parts cancel)
//   otherwise                : K = diffH(d) [- diffH(e) when e > 0 and r < 8e]
// with regH from its series for r/w < zs and diffH = singH - regH-series there.
__device__ __forceinline__ void vR_H(double r, const pb_kparams &K, double &H1, double &H2)
{
    if (K.has_e && r < K.e8 && r < K.d3) {
        double p1, p2, q1, q2;
        reg_H(r, K.pe, p1, p2);
        reg_H(r, K.pd, q1, q2);
        H1 = p1 - q1;
        H2 = p2 - q2;
        return;
    }
    diff_H(r, K.pd, H1, H2);
    if (K.has_e && r < K.e8) {
        double t1, t2;
        diff_H(r, K.pe, t1, t2);
        H1 -= t1;
        H2 -= t2;
    }
}

// All regions, not inlined: kernels whose common path is the table call this
// for the rare near field so its library constants stay out of their loops.
// (H1, H2) come back in registers.
__device__ __noinline__ double2 vR_H_call(double r, const pb_kparams *K)
{
    double2 H;
    vR_H(r, *K, H.x, H.y);
    return H;
}

// Where the table of diffH(d) lives: z = r/d in [tz0, tz0 + tn/tinv).
// sqrt(x) for normal x > 0 from the hardware reciprocal square root and one
// Newton step on the root (r += (x - r^2) / (2 sqrt x)): a few ulp, without the
// library's scaling and special-value handling (the callers guarantee x is
// a normal positive number).
__device__ __forceinline__ double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double r = x * y;
    return fma(0.5 * y, fma(-r, r, x), r);
}

struct tabview {
    double inv_d, tz0, tinv;
    int tn;
};

__device__ __forceinline__ tabview tabview_of(const pb_kparams &K)
{
    return tabview{K.pd.inv_d, K.tz0, K.tinv, K.tn};
}

// Table evaluation of diffH(d) for r >= K.rcommon: interval i of z = r/d and
// Horner in t on the interleaved (H1, H2) coefficients held in LDS.
template <int DEG>
__device__ __forceinline__ void table_H(double r, const tabview &V, const double2 *__restrict__ tab,
                                        double &H1, double &H2)
{
    const double u = (r * V.inv_d - V.tz0) * V.tinv;
    int i = (int)u;
    i = i < V.tn ? i : V.tn - 1;
    const double t = 2.0 * (u - (double)i) - 1.0;
    const double2 *c = tab + __umul24(i, DEG + 1); // i < 2^24: a full-rate 24-bit multiply
    double h1 = c[DEG].x, h2 = c[DEG].y;
#pragma unroll
    for (int k = DEG - 1; k >= 0; k--) {
        h1 = fma(h1, t, c[k].x);
        h2 = fma(h2, t, c[k].y);
    }
    H1 = h1;
    H2 = h2;
}

struct tabview32 {
    float inv_d, tz0, tinv;
    int tn;
};

__device__ __forceinline__ tabview32 tabview32_of(const pb_kparams &K)
{
    return tabview32{(float)K.pd.inv_d, (float)K.tz0, (float)K.tinv, K.tn};
}

template <int DEG>
__device__ __forceinline__ void table_H32(float r, const tabview32 &V, const float2 *__restrict__ tab,
                                          float &H1, float &H2)
{
    const float u = (r * V.inv_d - V.tz0) * V.tinv;
    int i = (int)u;
    i = i < V.tn ? i : V.tn - 1;
    const float t = 2.0f * (u - (float)i) - 1.0f;
    const float2 *c = tab + i * (DEG + 1);
    float h1 = c[DEG].x, h2 = c[DEG].y;
#pragma unroll
    for (int k = DEG - 1; k >= 0; k--) {
        h1 = fmaf(h1, t, c[k].x);
        h2 = fmaf(h2, t, c[k].y);
    }
    H1 = h1;
    H2 = h2;
}

// FP32 far field, r >= 3d and r >= 8e: only diffH(d) contributes.
__device__ __forceinline__ void diff_far32(float r, const pb_radial32 &P, float &H1, float &H2)
{
    float z = r * P.inv_d, s = P.a * r, s2 = s * s, x = P.x;
    float E = __expf(-z * z);
    float T0 = erfcxf(z);
    float Tp = P.ex2 * erfcxf(z + x);
    float Tm, extra = 0.0f;
    if (z >= x) {
        Tm = P.ex2 * erfcxf(z - x);
    } else {
        Tm = -P.ex2 * erfcxf(x - z);
        extra = 2.0f * __expf(-s);
    }
    float ir = 1.0f / r, ir2 = ir * ir, ir3 = ir2 * ir;
    float r2 = r * r;
    float a1 = 1.0f + s + s2, a2 = 1.0f - s + s2;
    float b1 = 3.0f + 3.0f * s + s2, b2 = 3.0f - 3.0f * s + s2;
    float B1 = -T0 + 0.5f * (a1 * Tm + a2 * Tp);
    float B2 = 3.0f * T0 - 0.5f * (b1 * Tm + b2 * Tp);
    float k3 = P.k4pa2 * ir3;
    H1 = E * (B1 * k3 - P.kg * (P.d2 + 2.0f * r2) * ir2) + 0.5f * extra * a1 * k3;
    H2 = E * (B2 * k3 * ir2 + P.kg * (3.0f * P.d2 + 2.0f * r2) * ir2 * ir2) - 0.5f * extra * b1 * k3 * ir2;
}

} // namespace pbdev
