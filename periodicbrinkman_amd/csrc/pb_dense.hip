// pb_dense.hip -- dense 3nt x 3ns Ewald matrix (the output of the reference's
// periodicBrinkman, kernelReal :219-314 and kernelFourier :325-425, and of the
// host shell sums realShells :564-693 / fourierShells :703-823), plus batched
// pointwise vR / vF.
//
// One thread owns one (target, source) 3x3 block and loops over the image
// set and the k set, so no reduction is needed; consecutive threads take
// consecutive targets (coalesced column segments of A).  The k-space factor
// b(k) = phi_d(k)/(k^2 (k^2 + a^2)) / tau is tabulated once per call, by a
// kernel on the same stream (the call enqueues work and returns).
#include "pb_radial.h"

#include <cmath>

namespace {

struct box3i {
    int lo[3], hi[3];
};

__device__ __forceinline__ bool in_cube(int l0, int l1, int l2, const int *m)
{
    return abs(l0) <= m[0] && abs(l1) <= m[1] && abs(l2) <= m[2];
}

// real_mode: 0 none, 1 home box only (overwrite A and absA), 2 layer hi \ lo (add),
// 3 cube hi (add).  k_mode: 0 none, 1 layer hi \ lo excluding k = 0 (add).
__global__ __launch_bounds__(256) void k_dense(double *__restrict__ A, int lda,
                                               const double *__restrict__ y, int nt,
                                               const double *__restrict__ x, int ns,
                                               pb_kparams kp, pb_box box, box3i rb, int real_mode,
                                               const double4 *__restrict__ ktab, int nk,
                                               double *__restrict__ absA)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nt * ns) return;
    int i = (int)(p % nt), j = (int)(p / nt);
    double acc[9], aab[9];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] = aab[q] = 0.0;
    const double r0 = y[3 * (size_t)i] - x[3 * (size_t)j];
    const double r1 = y[3 * (size_t)i + 1] - x[3 * (size_t)j + 1];
    const double r2 = y[3 * (size_t)i + 2] - x[3 * (size_t)j + 2];
    if (real_mode) {
        const int *hi = real_mode == 1 ? rb.lo : rb.hi; // home: lo = hi = 0
        for (int l2 = -hi[2]; l2 <= hi[2]; l2++)
            for (int l1 = -hi[1]; l1 <= hi[1]; l1++)
                for (int l0 = -hi[0]; l0 <= hi[0]; l0++) {
                    if (real_mode == 1 && (l0 | l1 | l2)) continue;
                    if (real_mode == 2 && in_cube(l0, l1, l2, rb.lo)) continue;
                    double d0 = r0 - l0 * box.L[0], d1 = r1 - l1 * box.L[1], d2 = r2 - l2 * box.L[2];
                    double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    double H1, H2;
                    pbdev::vR_H(r, kp, H1, H2);
                    double dv[3] = {d0, d1, d2};
#pragma unroll
                    for (int b = 0; b < 3; b++)
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            double v = dv[a] * dv[b] * H2 + (a == b ? H1 : 0.0);
                            acc[b * 3 + a] += v;
                            aab[b * 3 + a] += fabs(v);
                        }
                }
    }
    for (int q = 0; q < nk; q++) {
        const double4 kk = ktab[q]; // (k0, k1, k2, b(k)/tau); 0 outside the layer and at k = 0
        if (kk.w == 0.0) continue;  // (the same entry for every thread: a uniform branch)
        double bh = kk.w * cos(kk.x * r0 + kk.y * r1 + kk.z * r2);
        double kv[3] = {kk.x, kk.y, kk.z};
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                double v;
                if (a == b) v = (kv[(a + 1) % 3] * kv[(a + 1) % 3] + kv[(a + 2) % 3] * kv[(a + 2) % 3]) * bh;
                else v = -kv[a] * kv[b] * bh;
                acc[b * 3 + a] += v;
                aab[b * 3 + a] += fabs(v);
            }
    }
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            size_t o = (size_t)(3 * j + b) * (size_t)lda + (size_t)(3 * i + a);
            if (real_mode == 1) {
                A[o] = acc[b * 3 + a];
                if (absA) absA[o] = aab[b * 3 + a];
            } else {
                A[o] += acc[b * 3 + a];
                if (absA) absA[o] += aab[b * 3 + a];
            }
        }
}

// k table of the cube |l_i| <= hi_i: entry q = (k, b(k)/tau) with
// b = (1 + kappa |k|^2) e^{-d^2 |k|^2/4} / (|k|^2 (|k|^2 + a^2)), weight 0 for
// the inner cube |l_i| <= lo_i and for k = 0 (periodicBrinkman.h:189-195, with
// the 1/tau of fourierShells :718,745).
struct kcube {
    int lo[3], hi[3];
    double k[3];                    // 2 pi / L per axis
    double kappa, d2q, a2, inv_tau; // (1 - e^{-a^2 d^2/4})/a^2, d^2/4, a^2, 1/(Lx Ly Lz)
};

__global__ void k_ktable(kcube kc, long long n, double4 *__restrict__ out)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int n0 = 2 * kc.hi[0] + 1, n1 = 2 * kc.hi[1] + 1;
    const int l0 = (int)(q % n0) - kc.hi[0], l1 = (int)((q / n0) % n1) - kc.hi[1], l2 = (int)(q / n0 / n1) - kc.hi[2];
    const double k0 = kc.k[0] * l0, k1 = kc.k[1] * l1, k2 = kc.k[2] * l2;
    const double kn = k0 * k0 + k1 * k1 + k2 * k2;
    const bool inner = abs(l0) <= kc.lo[0] && abs(l1) <= kc.lo[1] && abs(l2) <= kc.lo[2];
    double b = 0.0;
    if (!inner && kn > 0.0) b = (1.0 + kc.kappa * kn) * exp(-kc.d2q * kn) / (kn * (kn + kc.a2)) * kc.inv_tau;
    out[q] = make_double4(k0, k1, k2, b);
}

__global__ void k_vR_batch(double *__restrict__ A, const double *__restrict__ xh, int n,
                           pb_kparams kp)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double d[3] = {xh[3 * (size_t)p], xh[3 * (size_t)p + 1], xh[3 * (size_t)p + 2]};
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double H1, H2;
    pbdev::vR_H(r, kp, H1, H2);
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) A[9 * (size_t)p + b * 3 + a] = d[a] * d[b] * H2 + (a == b ? H1 : 0.0);
}

__global__ void k_vF_batch(double *__restrict__ A, const double *__restrict__ xh,
                           const double *__restrict__ kv, int n, double a, double d)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double *x = xh + 3 * (size_t)p, *k = kv + 3 * (size_t)p;
    double kx = x[0] * k[0] + x[1] * k[1] + x[2] * k[2];
    double kn = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    double X = a * d * a * d;
    double a1 = -expm1(-X / 4);
    double bhat = (1.0 + a1 * d * d * kn / X) * exp(-d * d * kn / 4) * cos(kx) / (kn * (kn + a * a));
    for (int b = 0; b < 3; b++)
        for (int c = 0; c < 3; c++) {
            double v;
            if (b == c) v = (kn - k[c] * k[c]) * bhat;
            else v = -k[c] * k[b] * bhat;
            A[9 * (size_t)p + b * 3 + c] = v;
        }
}

} // namespace

int pb_run_dense(pb_context *c, hipStream_t s, double *dA, int lda, const double *y, int nt, const double *x,
                 int ns, const pb_kparams &kp, const pb_box &box, const int *lo_real, const int *hi_real,
                 int real_mode, const int *lo_k, const int *hi_k, int k_mode, double *absA)
{
    // k table of the whole cube hi_k, built on the device into the context's
    // buffer (no host upload, no synchronization: the call stays
    // asynchronous); entries of the inner cube lo_k and k = 0 get weight 0
    int nk = 0;
    const double4 *dk = nullptr;
    if (k_mode) {
        const long long n = (2LL * hi_k[0] + 1) * (2LL * hi_k[1] + 1) * (2LL * hi_k[2] + 1);
        if (n > (1LL << 28)) return pb_fail(c, PB_ERR_UNSUPPORTED, "too many Fourier shells");
        int st;
        if ((st = pb_reserve(c, &c->dense_tmp, sizeof(double4) * (size_t)n))) return st;
        kcube kc;
        for (int q = 0; q < 3; q++) {
            kc.lo[q] = lo_k[q];
            kc.hi[q] = hi_k[q];
        }
        const double a = kp.pd.a, d = kp.pd.d, X = a * d * a * d;
        kc.kappa = -std::expm1(-X / 4) * d * d / X; // (1 - e^{-a^2 d^2/4}) / a^2
        kc.d2q = d * d / 4;
        kc.a2 = a * a;
        kc.inv_tau = 1.0 / (box.L[0] * box.L[1] * box.L[2]);
        for (int q = 0; q < 3; q++) kc.k[q] = 2 * M_PI / box.L[q];
        hipLaunchKernelGGL(k_ktable, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, kc, n,
                           (double4 *)c->dense_tmp.ptr);
        nk = (int)n;
        dk = (const double4 *)c->dense_tmp.ptr;
    }
    box3i rb;
    for (int q = 0; q < 3; q++) {
        rb.lo[q] = lo_real ? lo_real[q] : 0;
        rb.hi[q] = hi_real ? hi_real[q] : 0;
    }
    long long npair = (long long)nt * ns;
    if (npair > 0)
        hipLaunchKernelGGL(k_dense, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, dA, lda, y,
                           nt, x, ns, kp, box, rb, real_mode, dk, nk, absA);
    return pb_hip_check(c, hipGetLastError(), "dense launch");
}

int pb_run_vR_batch(hipStream_t s, double *A, const double *xh, int n, const pb_kparams &kp)
{
    if (n > 0) hipLaunchKernelGGL(k_vR_batch, dim3((n + 255) / 256), dim3(256), 0, s, A, xh, n, kp);
    return hipGetLastError() == hipSuccess ? PB_OK : PB_ERR_HIP;
}

int pb_run_vF_batch(hipStream_t s, double *A, const double *xh, const double *k, int n, double a,
                    double d)
{
    if (n > 0) hipLaunchKernelGGL(k_vF_batch, dim3((n + 255) / 256), dim3(256), 0, s, A, xh, k, n, a, d);
    return hipGetLastError() == hipSuccess ? PB_OK : PB_ERR_HIP;
}
