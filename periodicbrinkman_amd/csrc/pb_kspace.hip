// pb_kspace.hip -- k-space part of the Ewald sum (the fftBrinkman algorithm,
// fftBrinkman.h:34-287):
//   1. spread each force with the Gaussian g(x) = (pi eta)^{-3/2} e^{-|x|^2/eta}
//      onto P points per axis of an n0 x n1 x n2 grid (centred window);
//   2. forward FFT of the three fields;
//   3. multiply each mode by b(k) (k^2 I - k k^T),
//      b = phi_d(k) e^{eta k^2/2} / (k^2 (k^2 + a^2)) / N_g, and zero k = 0;
//   4. inverse FFT;
//   5. gather u(y) = h^3 sum_window g(x_g - y) G(x_g).
// T is double (plain form) or float (mixed form).  Window indices and
// wrapped positions are always computed in FP64 with the oracle's formula.
#include "pb_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int kPmax = 32;

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };

__device__ __forceinline__ int floordiv(int a, int b)
{
    int q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__device__ __forceinline__ int wrapi(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// ---- spreading ----------------------------------------------------------------
// One wavefront owns a 4 x 4 x 4 block of grid points, one point per lane, and
// gathers every source whose centred window (first index
// f0 = floor(x/h - P/2) + 1, the oracle's formula) touches the block.
// Candidates come from the real-space cell list: the cells covering the
// x-interval [(4b - P/2) h, (4b + 3 + P/2) h) per axis, unwrapped so that a
// source near the opposite face enters with its periodic image.  Lanes test 64
// candidates at a time and append the hits to an LDS ring; every 64 hits each
// lane computes the 4+4+4 block-local weights of one hit, then all lanes
// accumulate their own point.  Every grid point is written exactly once, so
// there are no atomics and the grid needs no clearing.
constexpr int kSpreadWaves = 4;

template <typename T> struct spread_lds {
    uint2 q[128];      // (source index, image code)
    T w[64][17];       // wx[4], wy[4], wz[4], scale * f[3]; odd stride: no bank conflicts
};

template <typename T>
__device__ __forceinline__ void spread_flush(spread_lds<T> &W, int qh, int cnt, int lane,
                                             const double4 *__restrict__ spos,
                                             const double4 *__restrict__ sfrc, const pb_box &box,
                                             const pb_kgrid &g, const int *g0, const int *lidx,
                                             T &a0, T &a1, T &a2)
{
    if (lane < cnt) {
        const uint2 e = W.q[(qh + lane) & 127];
        const double4 x = spos[e.x];
        const double4 f = sfrc[e.x];
        const double xc[3] = {x.x, x.y, x.z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int l = (int)((e.y >> (3 * a)) & 7) - 4;
            const int f0 = (int)floor(xc[a] * g.inv_h[a] - 0.5 * g.P[a]) + 1 + l * g.n[a];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int gq = g0[a] + q;
                const int p = gq - f0;
                const double dx = (double)(gq - l * g.n[a]) * g.h[a] - xc[a];
                T w = (T)0;
                if (p >= 0 && p < g.P[a]) {
                    if (sizeof(T) == 8)
                        w = (T)exp(-dx * dx / g.eta);
                    else
                        w = (T)__expf(-(float)(dx * dx) * (float)(1.0 / g.eta));
                }
                W.w[lane][4 * a + q] = w;
            }
        }
        W.w[lane][12] = (T)(f.x * g.scale);
        W.w[lane][13] = (T)(f.y * g.scale);
        W.w[lane][14] = (T)(f.z * g.scale);
    }
    __builtin_amdgcn_wave_barrier();
    for (int k = 0; k < cnt; k++) {
        const T t = W.w[k][lidx[0]] * W.w[k][4 + lidx[1]] * W.w[k][8 + lidx[2]];
        a0 += t * W.w[k][12];
        a1 += t * W.w[k][13];
        a2 += t * W.w[k][14];
    }
    __builtin_amdgcn_wave_barrier();
}

template <typename T>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_spread_blocks(
    const int *__restrict__ cstart, const double4 *__restrict__ spos,
    const double4 *__restrict__ sfrc, pb_cells cells, pb_box box, pb_kgrid g,
    T *__restrict__ grid)
{
    __shared__ spread_lds<T> lds[kSpreadWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nb0 = g.n[0] >> 2, nb1 = g.n[1] >> 2, nb2 = g.n[2] >> 2;
    const int bid = blockIdx.x * kSpreadWaves + wave;
    if (bid >= nb0 * nb1 * nb2) return; // whole wavefront; no block barriers below
    spread_lds<T> &W = lds[wave];
    const int b[3] = {bid / (nb1 * nb2), (bid / nb2) % nb1, bid % nb2};
    const int g0[3] = {4 * b[0], 4 * b[1], 4 * b[2]};
    const int lidx[3] = {lane >> 4, (lane >> 2) & 3, lane & 3};
    int ulo[3], uhi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double lo = (g0[a] - 0.5 * g.P[a]) * g.h[a], hi = (g0[a] + 3 + 0.5 * g.P[a]) * g.h[a];
        ulo[a] = (int)floor(lo * cells.inv_h[a] - 1e-9);
        uhi[a] = (int)floor(hi * cells.inv_h[a] + 1e-9);
    }
    T a0 = 0, a1 = 0, a2 = 0;
    int qh = 0, qn = 0;
    for (int u0 = ulo[0]; u0 <= uhi[0]; u0++) {
        const int l0 = floordiv(u0, cells.nc[0]), w0 = u0 - l0 * cells.nc[0];
        for (int u1 = ulo[1]; u1 <= uhi[1]; u1++) {
            const int l1 = floordiv(u1, cells.nc[1]), w1 = u1 - l1 * cells.nc[1];
            const int colbase = (w0 * cells.nc[1] + w1) * cells.nc[2];
            // the z range splits into runs of one image shift, each a contiguous
            // run of the z-major sorted sources
            for (int zs = ulo[2]; zs <= uhi[2];) {
                const int l2 = floordiv(zs, cells.nc[2]);
                const int zend = min(uhi[2], (l2 + 1) * cells.nc[2] - 1);
                const int js = cstart[colbase + zs - l2 * cells.nc[2]];
                const int je = cstart[colbase + zend - l2 * cells.nc[2] + 1];
                zs = zend + 1;
                const unsigned code = (unsigned)(l0 + 4) | ((unsigned)(l1 + 4) << 3) | ((unsigned)(l2 + 4) << 6);
                for (int j0 = js; j0 < je; j0 += 64) {
                    const int j = j0 + lane;
                    bool hit = false;
                    if (j < je) {
                        const double4 x = spos[j];
                        const int f0x = (int)floor(x.x * g.inv_h[0] - 0.5 * g.P[0]) + 1 + l0 * g.n[0];
                        const int f0y = (int)floor(x.y * g.inv_h[1] - 0.5 * g.P[1]) + 1 + l1 * g.n[1];
                        const int f0z = (int)floor(x.z * g.inv_h[2] - 0.5 * g.P[2]) + 1 + l2 * g.n[2];
                        hit = f0x > g0[0] - g.P[0] && f0x <= g0[0] + 3 && f0y > g0[1] - g.P[1] &&
                              f0y <= g0[1] + 3 && f0z > g0[2] - g.P[2] && f0z <= g0[2] + 3;
                    }
                    const unsigned long long m = __ballot(hit);
                    if (m == 0ull) continue;
                    if (hit) {
                        const unsigned pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                        W.q[(qh + qn + pre) & 127] = make_uint2((unsigned)j, code);
                    }
                    qn += __popcll(m);
                    if (qn >= 64) {
                        __builtin_amdgcn_wave_barrier();
                        spread_flush<T>(W, qh, 64, lane, spos, sfrc, box, g, g0, lidx, a0, a1, a2);
                        qh = (qh + 64) & 127;
                        qn -= 64;
                    }
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (qn > 0) spread_flush<T>(W, qh, qn, lane, spos, sfrc, box, g, g0, lidx, a0, a1, a2);
    const size_t idx = 2 * (((size_t)(g0[0] + lidx[0]) * g.n[1] + (g0[1] + lidx[1])) * g.n[2] + (g0[2] + lidx[2]));
    const size_t fs = 2 * (size_t)g.ng;
    grid[idx] = a0;
    grid[idx + 1] = (T)0;
    grid[fs + idx] = a1;
    grid[fs + idx + 1] = (T)0;
    grid[2 * fs + idx] = a2;
    grid[2 * fs + idx + 1] = (T)0;
}

// ---- interpolation (gather) --------------------------------------------------
// u(y) = h^3 (pi eta)^{-3/2} sum_window e^{-|x_g - y|^2/eta} G(x_g).  One
// wavefront per target: the 3P weights and wrapped indices are computed once
// into LDS, then the lanes sweep the P^3 window with consecutive lanes on
// consecutive z (contiguous memory) and the partial sums meet in a wave
// reduction.
constexpr int kInterpWaves = 4;

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T, int P>
__global__ __launch_bounds__(64 * kInterpWaves) void k_interp(int nt, const double4 *__restrict__ tpos,
                                                            const int *__restrict__ tperm, pb_kgrid g,
                                                            const T *__restrict__ grid,
                                                            double *__restrict__ u, int accumulate)
{
    __shared__ T wts[kInterpWaves][3][P];
    __shared__ int ix[kInterpWaves][3][P];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * kInterpWaves + wave;
    if (i >= nt) return; // whole wavefront; no block barriers below
    const double4 y = tpos[i];
    for (int k = lane; k < 3 * P; k += 64) {
        const int a = k / P, p = k % P;
        const double yc = a == 0 ? y.x : (a == 1 ? y.y : y.z);
        const int f0 = (int)floor(yc * g.inv_h[a] - 0.5 * P) + 1;
        const double dx = (double)(f0 + p) * g.h[a] - yc;
        if (sizeof(T) == 8)
            wts[wave][a][p] = (T)exp(-dx * dx / g.eta);
        else
            wts[wave][a][p] = (T)__expf(-(float)(dx * dx) * (float)(1.0 / g.eta));
        int j = f0 + p;
        j = j < 0 ? j + g.n[a] : (j >= g.n[a] ? j - g.n[a] : j);
        ix[wave][a][p] = j;
    }
    __builtin_amdgcn_wave_barrier();
    const size_t fs = 2 * (size_t)g.ng;
    T a0 = 0, a1 = 0, a2 = 0;
    constexpr int kIter = (P * P * P + 63) / 64;
#pragma unroll 9
    for (int it = 0; it < kIter; it++) {
        const int k = it * 64 + lane;
        if (P * P * P % 64 == 0 || k < P * P * P) {
            const int pz = k % P, py = (k / P) % P, px = k / (P * P);
            const size_t off = 2 * (((size_t)ix[wave][0][px] * g.n[1] + ix[wave][1][py]) * g.n[2] + ix[wave][2][pz]);
            const T wt = wts[wave][0][px] * wts[wave][1][py] * wts[wave][2][pz];
            a0 += wt * grid[off];
            a1 += wt * grid[fs + off];
            a2 += wt * grid[2 * fs + off];
        }
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane == 0) {
        const double sc = g.scale * g.vol;
        const size_t o = 3 * (size_t)tperm[i];
        if (accumulate) {
            u[o] += sc * (double)a0;
            u[o + 1] += sc * (double)a1;
            u[o + 2] += sc * (double)a2;
        } else {
            u[o] = sc * (double)a0;
            u[o + 1] = sc * (double)a1;
            u[o + 2] = sc * (double)a2;
        }
    }
}

// Per-axis window sizes that differ (fftBrinkman's P[3]) take this path.
template <typename T>
__global__ __launch_bounds__(256) void k_interp_any(int nt, const double4 *__restrict__ tpos,
                                                    const int *__restrict__ tperm, pb_kgrid g,
                                                    const T *__restrict__ grid,
                                                    double *__restrict__ u, int accumulate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const double4 y = tpos[i];
    const double yc[3] = {y.x, y.y, y.z};
    int f0[3];
    for (int a = 0; a < 3; a++) f0[a] = (int)floor(yc[a] * g.inv_h[a] - 0.5 * g.P[a]) + 1;
    const size_t fs = 2 * (size_t)g.ng;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int px = 0; px < g.P[0]; px++) {
        const double dxx = (double)(f0[0] + px) * g.h[0] - yc[0];
        const double wx = exp(-dxx * dxx / g.eta);
        const int jx = ((f0[0] + px) % g.n[0] + g.n[0]) % g.n[0];
        for (int py = 0; py < g.P[1]; py++) {
            const double dxy = (double)(f0[1] + py) * g.h[1] - yc[1];
            const double wxy = wx * exp(-dxy * dxy / g.eta);
            const int jy = ((f0[1] + py) % g.n[1] + g.n[1]) % g.n[1];
            for (int pz = 0; pz < g.P[2]; pz++) {
                const double dxz = (double)(f0[2] + pz) * g.h[2] - yc[2];
                const double wt = wxy * exp(-dxz * dxz / g.eta);
                const int jz = ((f0[2] + pz) % g.n[2] + g.n[2]) % g.n[2];
                const size_t idx = 2 * (((size_t)jx * g.n[1] + jy) * g.n[2] + jz);
                a0 += wt * (double)grid[idx];
                a1 += wt * (double)grid[fs + idx];
                a2 += wt * (double)grid[2 * fs + idx];
            }
        }
    }
    const double sc = g.scale * g.vol;
    const size_t o = 3 * (size_t)tperm[i];
    if (accumulate) {
        u[o] += sc * a0;
        u[o + 1] += sc * a1;
        u[o + 2] += sc * a2;
    } else {
        u[o] = sc * a0;
        u[o + 1] = sc * a1;
        u[o + 2] = sc * a2;
    }
}

template <typename T>
void launch_interp(hipStream_t s, int nt, const pb_kgrid &g, const void *tpos, const void *tperm,
                   const T *grid, double *u, int acc)
{
    dim3 gd((nt + kInterpWaves - 1) / kInterpWaves), bd(64 * kInterpWaves);
    dim3 gd1((nt + 255) / 256), bd1(256);
    const double4 *tp = (const double4 *)tpos;
    const int *pm = (const int *)tperm;
    bool same = g.P[0] == g.P[1] && g.P[1] == g.P[2];
    switch (same ? g.P[0] : -1) {
    case 4: hipLaunchKernelGGL((k_interp<T, 4>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 6: hipLaunchKernelGGL((k_interp<T, 6>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 8: hipLaunchKernelGGL((k_interp<T, 8>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 10: hipLaunchKernelGGL((k_interp<T, 10>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 12: hipLaunchKernelGGL((k_interp<T, 12>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 16: hipLaunchKernelGGL((k_interp<T, 16>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 20: hipLaunchKernelGGL((k_interp<T, 20>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    case 24: hipLaunchKernelGGL((k_interp<T, 24>), gd, bd, 0, s, nt, tp, pm, g, grid, u, acc); break;
    default: hipLaunchKernelGGL(k_interp_any<T>, gd1, bd1, 0, s, nt, tp, pm, g, grid, u, acc); break;
    }
}

// Batched radix-2 Stockham FFT of length N along one axis, B pencils per
// workgroup staged in LDS (ping-pong).  Forward uses e^{-2 pi i k/N}.
template <typename T, int N>
__global__ __launch_bounds__(256) void k_fft_axis(typename cplx<T>::type *__restrict__ data,
                                                  int axis, int n0, int n1, int n2, int B,
                                                  const typename cplx<T>::type *__restrict__ tw,
                                                  int inverse, long long field_stride)
{
    using C = typename cplx<T>::type;
    extern __shared__ unsigned char smem_raw[];
    C *X = reinterpret_cast<C *>(smem_raw);
    C *Y = X + (size_t)N * B;
    C *base = data + (size_t)blockIdx.y * field_stride;
    const long long pencil0 = (long long)blockIdx.x * B;
    const int tid = threadIdx.x;
    // addressing of element (p, b): pencil q = pencil0 + b
    auto addr = [&](int p, int b) -> size_t {
        long long q = pencil0 + b;
        if (axis == 2) return (size_t)q * N + p;
        if (axis == 1) {
            long long i0 = q / n2, i2 = q % n2;
            return (size_t)(i0 * (long long)n1 * n2 + (long long)p * n2 + i2);
        }
        return (size_t)((long long)p * n1 * n2 + q);
    };
    const int tot = N * B;
    if (axis == 2) {
        for (int idx = tid; idx < tot; idx += blockDim.x) {
            int b = idx / N, p = idx % N;
            X[b * N + p] = base[addr(p, b)];
        }
    } else {
        for (int idx = tid; idx < tot; idx += blockDim.x) {
            int p = idx / B, b = idx % B;
            X[b * N + p] = base[addr(p, b)];
        }
    }
    __syncthreads();
    const int half = N / 2;
    for (int Ns = 1; Ns < N; Ns <<= 1) {
        for (int idx = tid; idx < B * half; idx += blockDim.x) {
            int b = idx / half, k = idx % half;
            C a = X[b * N + k], c = X[b * N + k + half];
            int km = k & (Ns - 1);
            C w = tw[km * (N / (2 * Ns))];
            if (inverse) w.y = -w.y;
            C cw;
            cw.x = c.x * w.x - c.y * w.y;
            cw.y = c.x * w.y + c.y * w.x;
            int o = ((k - km) << 1) + km;
            C s, d;
            s.x = a.x + cw.x;
            s.y = a.y + cw.y;
            d.x = a.x - cw.x;
            d.y = a.y - cw.y;
            Y[b * N + o] = s;
            Y[b * N + o + Ns] = d;
        }
        __syncthreads();
        C *t = X;
        X = Y;
        Y = t;
    }
    if (axis == 2) {
        for (int idx = tid; idx < tot; idx += blockDim.x) {
            int b = idx / N, p = idx % N;
            base[addr(p, b)] = X[b * N + p];
        }
    } else {
        for (int idx = tid; idx < tot; idx += blockDim.x) {
            int p = idx / B, b = idx % B;
            base[addr(p, b)] = X[b * N + p];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_multiply(typename cplx<T>::type *__restrict__ grid,
                                                  pb_kgrid g, double k0, double k1, double k2,
                                                  double c, double a2, const double *__restrict__ fac)
{
    using C = typename cplx<T>::type;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.ng) return;
    int q2 = (int)(q % g.n[2]);
    int q1 = (int)((q / g.n[2]) % g.n[1]);
    int q0 = (int)(q / ((long long)g.n[2] * g.n[1]));
    // signed frequency, Nyquist taken as -n/2 (the reference's centred shift)
    int f0 = (2 * q0 >= g.n[0]) ? q0 - g.n[0] : q0;
    int f1 = (2 * q1 >= g.n[1]) ? q1 - g.n[1] : q1;
    int f2 = (2 * q2 >= g.n[2]) ? q2 - g.n[2] : q2;
    double kx = k0 * f0, ky = k1 * f1, kz = k2 * f2;
    double kn = kx * kx + ky * ky + kz * kz;
    C *F0 = grid, *F1 = grid + g.ng, *F2 = grid + 2 * g.ng;
    if (kn == 0.0) {
        C z;
        z.x = 0;
        z.y = 0;
        F0[q] = z;
        F1[q] = z;
        F2[q] = z;
        return;
    }
    // fac holds e^{-(d^2/4 - eta/2) k_a^2} per axis and frequency
    double b = (1.0 + c * kn) * fac[q0] * fac[g.n[0] + q1] * fac[g.n[0] + g.n[1] + q2] /
               (kn * (kn + a2) * (double)g.ng);
    C v0 = F0[q], v1 = F1[q], v2 = F2[q];
    T bk = (T)b, tkn = (T)kn, tx = (T)kx, ty = (T)ky, tz = (T)kz;
    T kr = tx * v0.x + ty * v1.x + tz * v2.x;
    T ki = tx * v0.y + ty * v1.y + tz * v2.y;
    C o0, o1, o2;
    o0.x = bk * (tkn * v0.x - tx * kr);
    o0.y = bk * (tkn * v0.y - tx * ki);
    o1.x = bk * (tkn * v1.x - ty * kr);
    o1.y = bk * (tkn * v1.y - ty * ki);
    o2.x = bk * (tkn * v2.x - tz * kr);
    o2.y = bk * (tkn * v2.y - tz * ki);
    F0[q] = o0;
    F1[q] = o1;
    F2[q] = o2;
}

template <typename T, int N>
void launch_fft(hipStream_t s, typename cplx<T>::type *grid, const pb_kgrid &g, int axis,
                const typename cplx<T>::type *tw, int inverse)
{
    using C = typename cplx<T>::type;
    int B;
    if (axis == 2) {
        B = 8;
        while (B > 1 && (size_t)2 * N * B * sizeof(C) > 65536) B >>= 1;
    } else {
        B = 16;
        while (B > 1 && ((size_t)2 * N * B * sizeof(C) > 65536 || g.n[2] % B)) B >>= 1;
    }
    long long npencil = g.ng / N;
    dim3 grid_dim((unsigned)(npencil / B), 3);
    size_t lds = (size_t)2 * N * B * sizeof(C);
    hipLaunchKernelGGL((k_fft_axis<T, N>), grid_dim, dim3(256), lds, s, grid, axis, g.n[0],
                       g.n[1], g.n[2], B, tw, inverse, (long long)g.ng);
}

template <typename T>
int fft_dispatch(hipStream_t s, typename cplx<T>::type *grid, const pb_kgrid &g, int axis,
                 const typename cplx<T>::type *tw, int inverse)
{
    switch (g.n[axis]) {
    case 2: launch_fft<T, 2>(s, grid, g, axis, tw, inverse); break;
    case 4: launch_fft<T, 4>(s, grid, g, axis, tw, inverse); break;
    case 8: launch_fft<T, 8>(s, grid, g, axis, tw, inverse); break;
    case 16: launch_fft<T, 16>(s, grid, g, axis, tw, inverse); break;
    case 32: launch_fft<T, 32>(s, grid, g, axis, tw, inverse); break;
    case 64: launch_fft<T, 64>(s, grid, g, axis, tw, inverse); break;
    case 128: launch_fft<T, 128>(s, grid, g, axis, tw, inverse); break;
    case 256: launch_fft<T, 256>(s, grid, g, axis, tw, inverse); break;
    case 512: launch_fft<T, 512>(s, grid, g, axis, tw, inverse); break;
    default: return PB_ERR_UNSUPPORTED;
    }
    return PB_OK;
}

// Twiddles e^{-2 pi i k/n}, k < n/2, for the three axes, stored consecutively
// at offsets 0, n0/2, n0/2 + n1/2 (computed in long double).
template <typename T>
int upload_twiddles(pb_context *c, hipStream_t s, const pb_kgrid &g)
{
    using C = typename cplx<T>::type;
    int prec = sizeof(T) == 8 ? 0 : 1;
    int key = (g.n[0] * 1024 + g.n[1]) * 1024 + g.n[2];
    if (c->twiddle_n == key && c->twiddle_prec == prec) return PB_OK;
    size_t tot = (size_t)(g.n[0] + g.n[1] + g.n[2]) / 2 + 3;
    std::vector<C> h(tot);
    size_t off = 0;
    for (int a = 0; a < 3; a++) {
        int n = g.n[a];
        for (int k = 0; k < n / 2; k++) {
            long double ang = -2.0L * 3.141592653589793238462643383279502884L * k / n;
            h[off + k].x = (T)cosl(ang);
            h[off + k].y = (T)sinl(ang);
        }
        off += n / 2 > 0 ? n / 2 : 1;
    }
    int st;
    if ((st = pb_reserve(c, &c->twiddle, tot * sizeof(C)))) return st;
    if ((st = pb_hip_check(c, hipMemcpyAsync(c->twiddle.ptr, h.data(), tot * sizeof(C),
                                             hipMemcpyHostToDevice, s), "twiddle upload")))
        return st;
    if ((st = pb_hip_check(c, hipStreamSynchronize(s), "twiddle sync"))) return st;
    c->twiddle_n = key;
    c->twiddle_prec = prec;
    return PB_OK;
}

template <typename T>
int run_kspace_t(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells,
                 const pb_box &box, double a, double d, int ns, int nt, double *u, bool accumulate)
{
    using C = typename cplx<T>::type;
    int st;
    for (int ax = 0; ax < 3; ax++) {
        int n = g.n[ax];
        if (n < 4 || (n & (n - 1)) || n > 512)
            return pb_fail(c, PB_ERR_UNSUPPORTED, "FFT grid sizes must be powers of two in [4, 512]");
        if (g.P[ax] < 1 || g.P[ax] > kPmax || g.P[ax] > n)
            return pb_fail(c, PB_ERR_INVALID, "window points P must be in [1, min(32, grid)]");
    }
    if ((st = pb_reserve(c, &c->grid, 3 * (size_t)g.ng * sizeof(C)))) return st;
    if ((st = upload_twiddles<T>(c, s, g))) return st;
    // per-axis factors e^{-(d^2/4 - eta/2) k^2}
    std::vector<double> fac((size_t)g.n[0] + g.n[1] + g.n[2]);
    double beta = d * d / 4 - g.eta / 2;
    size_t off = 0;
    for (int ax = 0; ax < 3; ax++) {
        for (int q = 0; q < g.n[ax]; q++) {
            int f = (2 * q >= g.n[ax]) ? q - g.n[ax] : q;
            double k = 2 * M_PI * f / box.L[ax];
            fac[off + q] = std::exp(-beta * k * k);
        }
        off += g.n[ax];
    }
    if ((st = pb_reserve(c, &c->axis_fac, fac.size() * sizeof(double)))) return st;
    if ((st = pb_hip_check(c, hipMemcpyAsync(c->axis_fac.ptr, fac.data(), fac.size() * sizeof(double),
                                             hipMemcpyHostToDevice, s), "factor upload")))
        return st;
    C *grid = (C *)c->grid.ptr;
    pb_mark(c, s, 1);
    {
        long long nb = (long long)(g.n[0] / 4) * (g.n[1] / 4) * (g.n[2] / 4);
        hipLaunchKernelGGL(k_spread_blocks<T>, dim3((unsigned)((nb + kSpreadWaves - 1) / kSpreadWaves)),
                           dim3(64 * kSpreadWaves), 0, s, (const int *)c->cell_start.ptr,
                           (const double4 *)c->spos.ptr, (const double4 *)c->sfrc.ptr, cells, box, g,
                           (T *)grid);
    }
    const C *tw = (const C *)c->twiddle.ptr;
    const C *tws[3] = {tw, tw + g.n[0] / 2, tw + g.n[0] / 2 + g.n[1] / 2};
    pb_mark(c, s, 2);
    for (int ax = 2; ax >= 0; ax--)
        if ((st = fft_dispatch<T>(s, grid, g, ax, tws[ax], 0))) return pb_fail(c, st, "FFT size");
    pb_mark(c, s, 3);
    double om = -std::expm1(-a * a * d * d / 4);
    hipLaunchKernelGGL(k_multiply<T>, dim3((unsigned)((g.ng + 255) / 256)), dim3(256), 0, s, grid, g,
                       2 * M_PI / box.L[0], 2 * M_PI / box.L[1], 2 * M_PI / box.L[2], om / (a * a),
                       a * a, (const double *)c->axis_fac.ptr);
    pb_mark(c, s, 4);
    for (int ax = 0; ax < 3; ax++)
        if ((st = fft_dispatch<T>(s, grid, g, ax, tws[ax], 1))) return pb_fail(c, st, "FFT size");
    pb_mark(c, s, 5);
    if (nt > 0) launch_interp<T>(s, nt, g, c->tpos.ptr, c->tperm.ptr, (const T *)grid, u, accumulate ? 1 : 0);
    pb_mark(c, s, 6);
    return pb_hip_check(c, hipGetLastError(), "k-space launch");
}

} // namespace

int pb_run_kspace(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells,
                  const pb_box &box, double a, double d, int ns, int nt, double *u, bool accumulate,
                  int precision)
{
    if (precision == PB_PREC_MIXED)
        return run_kspace_t<float>(c, s, g, cells, box, a, d, ns, nt, u, accumulate);
    return run_kspace_t<double>(c, s, g, cells, box, a, d, ns, nt, u, accumulate);
}
