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

__device__ __forceinline__ int wrapi(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ void window(double x, double h, double inv_h, int P, double eta,
                                       int &first, T *w)
{
    double sft = x * inv_h;
    first = (int)floor(sft - 0.5 * P) + 1;
    for (int p = 0; p < P; p++) {
        double dx = (first + p) * h - x;
        if (sizeof(T) == 8)
            w[p] = (T)exp(-dx * dx / eta);
        else
            w[p] = (T)__expf(-(float)(dx * dx) * (float)(1.0 / eta));
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_spread(int ns, const double4 *__restrict__ spos,
                                                const double4 *__restrict__ sfrc, pb_kgrid g,
                                                T *__restrict__ grid)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    double4 x = spos[j], f = sfrc[j];
    int f0[3];
    T w0[kPmax], w1[kPmax], w2[kPmax];
    window<T>(x.x, g.h[0], g.inv_h[0], g.P[0], g.eta, f0[0], w0);
    window<T>(x.y, g.h[1], g.inv_h[1], g.P[1], g.eta, f0[1], w1);
    window<T>(x.z, g.h[2], g.inv_h[2], g.P[2], g.eta, f0[2], w2);
    const size_t fs = 2 * (size_t)g.ng;
    const T fx = (T)f.x * (T)g.scale, fy = (T)f.y * (T)g.scale, fz = (T)f.z * (T)g.scale;
    for (int px = 0; px < g.P[0]; px++) {
        const int jx = wrapi(f0[0] + px, g.n[0]);
        for (int py = 0; py < g.P[1]; py++) {
            const int jy = wrapi(f0[1] + py, g.n[1]);
            const T wxy = w0[px] * w1[py];
            const size_t row = ((size_t)jx * g.n[1] + jy) * g.n[2];
            for (int pz = 0; pz < g.P[2]; pz++) {
                const int jz = wrapi(f0[2] + pz, g.n[2]);
                const T wt = wxy * w2[pz];
                const size_t idx = 2 * (row + jz);
                atomicAdd(&grid[idx], fx * wt);
                atomicAdd(&grid[fs + idx], fy * wt);
                atomicAdd(&grid[2 * fs + idx], fz * wt);
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(int nt, const double4 *__restrict__ tpos,
                                                const int *__restrict__ tperm, pb_kgrid g,
                                                const T *__restrict__ grid,
                                                double *__restrict__ u, int accumulate)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double4 y = tpos[i];
    int f0[3];
    T w0[kPmax], w1[kPmax], w2[kPmax];
    window<T>(y.x, g.h[0], g.inv_h[0], g.P[0], g.eta, f0[0], w0);
    window<T>(y.y, g.h[1], g.inv_h[1], g.P[1], g.eta, f0[1], w1);
    window<T>(y.z, g.h[2], g.inv_h[2], g.P[2], g.eta, f0[2], w2);
    const size_t fs = 2 * (size_t)g.ng;
    T a0 = 0, a1 = 0, a2 = 0;
    for (int px = 0; px < g.P[0]; px++) {
        const int jx = wrapi(f0[0] + px, g.n[0]);
        for (int py = 0; py < g.P[1]; py++) {
            const int jy = wrapi(f0[1] + py, g.n[1]);
            const size_t row = ((size_t)jx * g.n[1] + jy) * g.n[2];
            T b0 = 0, b1 = 0, b2 = 0;
            for (int pz = 0; pz < g.P[2]; pz++) {
                const int jz = wrapi(f0[2] + pz, g.n[2]);
                const size_t idx = 2 * (row + jz);
                const T wz = w2[pz];
                b0 += wz * grid[idx];
                b1 += wz * grid[fs + idx];
                b2 += wz * grid[2 * fs + idx];
            }
            const T wxy = w0[px] * w1[py];
            a0 += wxy * b0;
            a1 += wxy * b1;
            a2 += wxy * b2;
        }
    }
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

template <typename T>
__global__ void k_zero(T *__restrict__ p, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = (T)0;
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
int run_kspace_t(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_box &box, double a,
                 double d, int ns, int nt, double *u, bool accumulate)
{
    using C = typename cplx<T>::type;
    int st;
    for (int ax = 0; ax < 3; ax++) {
        int n = g.n[ax];
        if (n < 2 || (n & (n - 1)) || n > 512)
            return pb_fail(c, PB_ERR_UNSUPPORTED, "FFT grid sizes must be powers of two in [2, 512]");
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
    long long nreal = 3 * 2 * g.ng;
    hipLaunchKernelGGL(k_zero<T>, dim3(2048), dim3(256), 0, s, (T *)grid, nreal);
    pb_mark(c, s, 1);
    if (ns > 0)
        hipLaunchKernelGGL(k_spread<T>, dim3((ns + 255) / 256), dim3(256), 0, s, ns,
                           (const double4 *)c->spos.ptr, (const double4 *)c->sfrc.ptr, g, (T *)grid);
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
    if (nt > 0)
        hipLaunchKernelGGL(k_gather<T>, dim3((nt + 255) / 256), dim3(256), 0, s, nt,
                           (const double4 *)c->tpos.ptr, (const int *)c->tperm.ptr, g, (const T *)grid,
                           u, accumulate ? 1 : 0);
    pb_mark(c, s, 6);
    return pb_hip_check(c, hipGetLastError(), "k-space launch");
}

} // namespace

int pb_run_kspace(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_box &box, double a,
                  double d, int ns, int nt, double *u, bool accumulate, int precision)
{
    if (precision == PB_PREC_MIXED)
        return run_kspace_t<float>(c, s, g, box, a, d, ns, nt, u, accumulate);
    return run_kspace_t<double>(c, s, g, box, a, d, ns, nt, u, accumulate);
}
