// pb_mrhs.hip -- real-space part for many force sets on one particle set.
//
// For R right-hand sides the pair interaction is a contraction: for a group
// of 32 targets and 16 sources,
//   U_a[t][r] += sum_{s, b} K_ab(t, s) F[s][b][r],   K_ab = H1 delta_ab + H2 d_a d_b
// (K_R of periodicBrinkman.h:152-176, zero outside rc), i.e. for each
// component a a dense [32 x 48] . [48 x R] product, done with
// v_mfma_f32_32x32x2f32 (FP32 inputs, FP32 accumulation).  The contraction
// index is ordered k = 16 b + s, so the lane holding A[t][k] (lane t + 32 (k & 1))
// and the lane holding B[k][j] (lane j + 32 (k & 1)) both need only the
// sources s of their own parity: every lane evaluates the radial functions of
// its target with 8 of the 16 sources (FP64, the table or the rare-region
// call) and feeds both operands from registers.
//
// Candidate sources come from the same column walk as the single-force pair
// kernel (tiles of <= 64 targets split into halves of 32, one wavefront each);
// the kept sources collect in an LDS ring and are contracted 16 at a time.
#include "pb_radial.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMWaves = 4;
#ifndef MRHS_Q2
#define MRHS_Q2 4 // sources per lane parity and contraction step, 64 sets per launch
#endif
#ifndef MRHS_Q1
#define MRHS_Q1 8 // the same, 32 sets per launch
#endif

struct mrhs_lds {
    double4 ring[128]; // kept candidate sources: image position, index (as bits)
};

__device__ __forceinline__ unsigned lane_prefix(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ double uniform(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return uniform(v);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return uniform(v);
}

__device__ __forceinline__ int floordiv(int a, int b)
{
    int q = a / b;
    return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// Contract the 2Q ring entries from head (cnt <= 2Q of them valid) into acc.
template <int Q, int NB, int DEG>
__device__ __forceinline__ void contract(const mrhs_lds &W, int head, int cnt, int lane, double y0, double y1,
                                           double y2, bool active, double rc2, double rcommon,
                                           const pbdev::tabview &tv, const double2 *__restrict__ tab,
                                           const pb_kparams *kpp, const float *__restrict__ Fs, int RP, int c0,
                                           f32x16 (&acc)[3][NB])
{
    const int t = lane & 31, kh = lane >> 5;
    (void)t;
    float H1[Q], H2[Q], dd[Q][3];
    int src[Q];
    double4 scs[Q];
    // the ring entries, then the B operands (forces of this lane's sources),
    // issued before the radial functions so that their latency overlaps them
    // (measured at 2e5 x 64 sets: call 36.5 -> 32.7 ms, although 12 registers
    // spill at four sources per parity)
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int s = 2 * q + kh;
        const bool valid = s < cnt;
        scs[q] = W.ring[(head + (valid ? s : 0)) & 127];
        src[q] = valid ? (int)__double_as_longlong(scs[q].w) : -1;
    }
    float bvs[3 * Q][NB];
#pragma unroll
    for (int kk = 0; kk < 3 * Q; kk++) {
        const int b = kk / Q, q = kk % Q;
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
            bvs[kk][nb] = src[q] >= 0 ? Fs[((size_t)src[q] * 3 + b) * RP + c0 + 32 * nb + (lane & 31)] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int s = 2 * q + kh;
        const bool valid = s < cnt;
        const double4 sc = scs[q];
        const double d0 = y0 - sc.x, d1 = y1 - sc.y, d2 = y2 - sc.z;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        const bool in = valid && active && r2 < rc2;
        double h1 = 0, h2 = 0;
        if (in) {
            const double r = sqrt(r2);
            if (r >= rcommon) {
                pbdev::table_H<DEG>(r, tv, tab, h1, h2);
            } else {
                const double2 H = pbdev::vR_H_call(r, kpp);
                h1 = H.x;
                h2 = H.y;
            }
        }
        H1[q] = (float)h1;
        H2[q] = (float)h2;
        dd[q][0] = (float)d0;
        dd[q][1] = (float)d1;
        dd[q][2] = (float)d2;
    }
#pragma unroll
    for (int kk = 0; kk < 3 * Q; kk++) {
        const int b = kk / Q, q = kk % Q;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float av = (a == b ? H1[q] : 0.0f) + H2[q] * dd[q][a] * dd[q][b];
#pragma unroll
            for (int nb = 0; nb < NB; nb++)
                acc[a][nb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bvs[kk][nb], acc[a][nb], 0, 0, 0);
        }
    }
}

template <int Q, int NB, int DEG>
__global__ __launch_bounds__(64 * kMWaves, 2) void k_pair_mrhs(
    const int2 *__restrict__ tiles, const int *__restrict__ ntiles_ptr, const int *__restrict__ cstart,
    const double4 *__restrict__ spos, const float *__restrict__ Fs, int RP, int c0, int R,
    const double4 *__restrict__ tpos, const int *__restrict__ tperm, int nt, pb_cells cells, pb_box box,
    const pb_kparams *__restrict__ kpp, double *__restrict__ u)
{
    const pb_kparams &kp = *kpp;
    const pbdev::tabview tv = pbdev::tabview_of(kp);
    const double rcommon = kp.rcommon;
    __shared__ mrhs_lds lds[kMWaves];
    extern __shared__ double dyn_lds[]; // the radial table
    const int ntab = kp.tn * (DEG + 1) * 2;
    for (int k = threadIdx.x; k < ntab; k += blockDim.x) dyn_lds[k] = kp.tab64[k];
    __syncthreads();
    const double2 *tab = reinterpret_cast<const double2 *>(dyn_lds);

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int item = blockIdx.x * kMWaves + wave;
    const int tile = item >> 1, half = item & 1;
    if (tile >= *ntiles_ptr) return; // whole wavefront; no block barriers below
    const int2 T = tiles[tile];
    const int tb = T.x + 32 * half, cnt = min(32, T.y - 32 * half);
    if (cnt <= 0) return;
    mrhs_lds &W = lds[wave];
    const int t = lane & 31;
    const bool active = t < cnt;
    const double4 p = tpos[tb + (active ? t : 0)];
    const double bmin0 = wave_min(p.x), bmin1 = wave_min(p.y), bmin2 = wave_min(p.z);
    const double bmax0 = wave_max(p.x), bmax1 = wave_max(p.y), bmax2 = wave_max(p.z);
    const double y0 = p.x, y1 = p.y, y2 = p.z;

    const int nc0 = cells.nc[0], nc1 = cells.nc[1], nc2 = cells.nc[2];
    const double rcx = cells.rc * (1.0 + 1e-12);
    const double rcx2 = rcx * rcx, rc2 = cells.rc2;
    const int ulo0 = (int)floor((bmin0 - rcx) * cells.inv_h[0]), uhi0 = (int)floor((bmax0 + rcx) * cells.inv_h[0]);
    const int ulo1 = (int)floor((bmin1 - rcx) * cells.inv_h[1]), uhi1 = (int)floor((bmax1 + rcx) * cells.inv_h[1]);
    const int ny = uhi1 - ulo1 + 1, ncols = (uhi0 - ulo0 + 1) * ny;

    f32x16 acc[3][NB];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[a][nb][v] = 0.0f;
    int head = 0, pend = 0;

    for (int cb = 0; cb < ncols; cb += 64) {
        const int colv = cb + lane;
        int ze = -1, l0 = 0, l1 = 0, colbase = 0, l2 = 0, zend = -1, ja = 0, jea = 0;
        bool ok = colv < ncols;
        if (ok) {
            const int u0 = ulo0 + colv / ny, u1 = ulo1 + colv % ny;
            const double g0 = fmax(0.0, fmax(u0 * cells.h[0] - bmax0, bmin0 - (u0 + 1) * cells.h[0]));
            const double g1 = fmax(0.0, fmax(u1 * cells.h[1] - bmax1, bmin1 - (u1 + 1) * cells.h[1]));
            const double g2 = g0 * g0 + g1 * g1;
            ok = g2 < rcx2;
            if (ok) {
                const double rem = sqrt(rcx2 - g2);
                const int zs = (int)floor((bmin2 - rem) * cells.inv_h[2]);
                ze = (int)floor((bmax2 + rem) * cells.inv_h[2]);
                l0 = floordiv(u0, nc0);
                l1 = floordiv(u1, nc1);
                colbase = ((u0 - l0 * nc0) * nc1 + (u1 - l1 * nc1)) * nc2;
                l2 = floordiv(zs, nc2);
                zend = min(ze, (l2 + 1) * nc2 - 1);
                ja = cstart[colbase + zs - l2 * nc2];
                jea = cstart[colbase + zend - l2 * nc2 + 1];
            }
        }
        unsigned long long cm = __ballot(ok);
        while (cm) {
            const int i = __builtin_ctzll(cm);
            cm &= cm - 1;
            const int cl0 = __builtin_amdgcn_readlane(l0, i), cl1 = __builtin_amdgcn_readlane(l1, i);
            const int cze = __builtin_amdgcn_readlane(ze, i), cbase = __builtin_amdgcn_readlane(colbase, i);
            int cl2 = __builtin_amdgcn_readlane(l2, i), czend = __builtin_amdgcn_readlane(zend, i);
            int j = __builtin_amdgcn_readlane(ja, i), je = __builtin_amdgcn_readlane(jea, i);
            const double sh0 = cl0 * box.L[0], sh1 = cl1 * box.L[1];
            for (;;) {
                const double sh2 = cl2 * box.L[2];
                for (; j < je; j += 64) {
                    const int nk = min(64, je - j);
                    bool keep = false;
                    double sx = 0, sy = 0, sz = 0;
                    if (lane < nk) {
                        const double4 xs = spos[j + lane];
                        sx = xs.x + sh0;
                        sy = xs.y + sh1;
                        sz = xs.z + sh2;
                        const double g0 = fmax(0.0, fmax(bmin0 - sx, sx - bmax0));
                        const double g1 = fmax(0.0, fmax(bmin1 - sy, sy - bmax1));
                        const double g2 = fmax(0.0, fmax(bmin2 - sz, sz - bmax2));
                        keep = g0 * g0 + g1 * g1 + g2 * g2 < rcx2;
                    }
                    const unsigned long long km = __ballot(keep);
                    __builtin_amdgcn_wave_barrier();
                    if (keep)
                        W.ring[(head + pend + lane_prefix(km)) & 127] =
                            make_double4(sx, sy, sz, __longlong_as_double((long long)(j + lane)));
                    pend += __popcll(km);
                    __builtin_amdgcn_wave_barrier();
                    while (pend >= 2 * Q) {
                        contract<Q, NB, DEG>(W, head, 2 * Q, lane, y0, y1, y2, active, rc2, rcommon, tv, tab, kpp, Fs,
                                             RP, c0, acc);
                        head = (head + 2 * Q) & 127;
                        pend -= 2 * Q;
                    }
                }
                if (czend >= cze) break;
                const int zsn = czend + 1;
                cl2 = floordiv(zsn, nc2);
                czend = min(cze, (cl2 + 1) * nc2 - 1);
                j = cstart[cbase + zsn - cl2 * nc2];
                je = cstart[cbase + czend - cl2 * nc2 + 1];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (pend > 0)
        contract<Q, NB, DEG>(W, head, pend, lane, y0, y1, y2, active, rc2, rcommon, tv, tab, kpp, Fs, RP, c0, acc);
    // C[i][j]: column j = lane & 31, row i = 8 (v / 4) + 4 (lane / 32) + v % 4
    const int kh = lane >> 5;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = 8 * (v >> 2) + 4 * kh + (v & 3);
        if (i >= cnt) continue;
        const size_t o = 3 * (size_t)tperm[tb + i];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int r = c0 + 32 * nb + (lane & 31);
            if (r >= R) continue;
            double *ur = u + (size_t)r * 3 * nt + o;
#pragma unroll
            for (int a = 0; a < 3; a++) ur[a] = (double)acc[a][nb][v]; // each (set, target) once per call
        }
    }
}

// u[i] += ur[i], i < n
__global__ void k_add(long long n, double *__restrict__ u, const double *__restrict__ ur)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += ur[i];
}

// Fs[j][b][rp] = f[rp][perm[j]][b] (zero for rp >= R): the forces of the
// sorted sources, RHS index fastest.
__global__ void k_sort_forces(int ns, int R, int RP, const int *__restrict__ perm, const float *__restrict__ f,
                              float *__restrict__ Fs)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)ns * 3 * RP) return;
    const int rp = (int)(q % RP);
    const long long jb = q / RP;
    const int b = (int)(jb % 3), j = (int)(jb / 3);
    Fs[q] = rp < R ? f[(size_t)rp * 3 * ns + 3 * (size_t)perm[j] + b] : 0.0f;
}

// Forces of sets r0 .. r0 + nr - 1 from the sorted FP32 array Fs[j][3][RP]
// to the per-set FP64 arrays frc[r][j] the spreading reads: tiles of 64
// sources x 16 sets through LDS, so that the reads (runs of sets) and the
// writes (runs of sources) are both contiguous.
__global__ __launch_bounds__(256) void k_force_sets(int ns, int r0, int nr, int RP, const float *__restrict__ Fs,
                                                    double4 *__restrict__ frc)
{
    __shared__ float T[64][3][17];
    const int j0 = (int)blockIdx.x * 64, t0 = (int)blockIdx.y * 16;
    for (int idx = threadIdx.x; idx < 64 * 3 * 16; idx += 256) {
        const int r = idx % 16, jb = idx / 16, b = jb % 3, j = jb / 3;
        const int jj = j0 + j, rr = t0 + r;
        T[j][b][r] = (jj < ns && rr < nr) ? Fs[(size_t)jj * 3 * RP + (size_t)b * RP + r0 + rr] : 0.0f;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * 16; idx += 256) {
        const int j = idx % 64, r = idx / 64;
        const int jj = j0 + j, rr = t0 + r;
        if (jj < ns && rr < nr) frc[(size_t)rr * ns + jj] = make_double4(T[j][0][r], T[j][1][r], T[j][2][r], 0.0);
    }
}

} // namespace

int pb_force_sets(pb_context *c, hipStream_t s, int ns, int r0, int nr, int RP, double4 *frc)
{
    if (ns > 0 && nr > 0)
        hipLaunchKernelGGL(k_force_sets, dim3((unsigned)((ns + 63) / 64), (unsigned)((nr + 15) / 16)), dim3(256), 0,
                           s, ns, r0, nr, RP, (const float *)c->mrhs_f.ptr, frc);
    return pb_hip_check(c, hipGetLastError(), "force sets launch");
}

void pb_add(hipStream_t s, long long n, double *u, const double *ur)
{
    if (n > 0) hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, u, ur);
}

int pb_sort_forces(pb_context *c, hipStream_t s, int ns, int R, int RP, const float *f)
{
    int st;
    if ((st = pb_reserve(c, &c->mrhs_f, sizeof(float) * (size_t)(ns > 0 ? ns : 1) * 3 * RP))) return st;
    const long long n = (long long)ns * 3 * RP;
    if (n > 0)
        hipLaunchKernelGGL(k_sort_forces, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ns, R, RP,
                           (const int *)c->sperm.ptr, f, (float *)c->mrhs_f.ptr);
    return pb_hip_check(c, hipGetLastError(), "force sort launch");
}

int pb_run_mrhs_real(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_kparams *kp,
                     size_t table_doubles, int table_deg, const pb_box &box, int ns, int nt, int R, int RP,
                     double *u)
{
    (void)ns;
    if (nt <= 0 || R <= 0) return PB_OK;
    if (!table_doubles) return pb_fail(c, PB_ERR_UNSUPPORTED, "the multi-RHS path needs the radial table");
    const int *ntiles = (const int *)c->qoff.ptr + cells.nkey_t / 16;
    const unsigned blocks = (unsigned)((2LL * c->max_tiles + kMWaves - 1) / kMWaves);
    const size_t lds = table_doubles * sizeof(double);
    // 64 right-hand sides per launch (two 32 x 32 accumulator blocks per
    // component, 8 sources per contraction step), a last 32 with one block
    if (table_deg != PB_TAB_DEG && table_deg != PB_TAB_DEG_MIXED && table_deg != PB_TAB_DEG_MID)
        return pb_fail(c, PB_ERR_INVALID, "radial table of an unsupported degree");
    auto k64 = table_deg == PB_TAB_DEG   ? k_pair_mrhs<MRHS_Q2, 2, PB_TAB_DEG>
               : table_deg == PB_TAB_DEG_MID ? k_pair_mrhs<MRHS_Q2, 2, PB_TAB_DEG_MID>
                                             : k_pair_mrhs<MRHS_Q2, 2, PB_TAB_DEG_MIXED>;
    auto k32 = table_deg == PB_TAB_DEG   ? k_pair_mrhs<MRHS_Q1, 1, PB_TAB_DEG>
               : table_deg == PB_TAB_DEG_MID ? k_pair_mrhs<MRHS_Q1, 1, PB_TAB_DEG_MID>
                                             : k_pair_mrhs<MRHS_Q1, 1, PB_TAB_DEG_MIXED>;
    for (int c0 = 0; c0 < RP; c0 += 64)
        hipLaunchKernelGGL(RP - c0 >= 64 ? k64 : k32, dim3(blocks), dim3(64 * kMWaves), lds, s,
                           (const int2 *)c->tiles.ptr, ntiles, (const int *)c->cell_start.ptr,
                           (const double4 *)c->spos.ptr, (const float *)c->mrhs_f.ptr, RP, c0, R,
                           (const double4 *)c->tpos.ptr, (const int *)c->tperm.ptr, nt, cells, box, kp, u);
    return pb_hip_check(c, hipGetLastError(), "multi-RHS pair launch");
}
