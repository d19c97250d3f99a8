// pb_real.hip -- real-space image sum over cell-listed neighbours.
//
// u(y_t) += sum over source images with |y_t - x_j - l L| < rc of
//           K_R(y_t - x_j - l L) f_j,   K_R = S from diffH(d) - diffH(e)
// (periodicBrinkman.h:152-176; the reference sums all pairs over image cubes,
// realShells :564-693, kernelReal :219-314).
//
// One wavefront owns one target tile (<= 64 targets of one cell, one per
// lane).  It walks the (2st+1)^3 neighbour cells; each source is loaded once
// per wave (uniform address -> scalar loads) and tested against all 64
// targets.  In-range (target, source, image) triples are compacted with a
// ballot into a 128-entry LDS ring; whenever 64 are pending every lane
// evaluates one of them, so the expensive exp/erfcx work runs with full
// lanes regardless of how sparse the neighbourhood test is.  Results are
// added into per-target FP64 accumulators in LDS (ds_add_f64).  In mixed
// precision, pairs beyond rfar (>= 3d and >= 8e) go to a second ring and are
// evaluated in FP32; their results are still accumulated in FP64.
#include "pb_radial.h"

namespace {

constexpr int kWaves = 4;  // wavefronts per workgroup, each independent
constexpr int kQ = 128;    // ring capacity (power of two)

struct wave_lds {
    double y[64][3];
    double acc[64][3];
    uint2 q64[kQ];
    uint2 q32[kQ];
};

// Broadcast lane k's double to the whole wavefront (k uniform): v_readlane.
__device__ __forceinline__ double lane_bcast(double v, int k)
{
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, k);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ unsigned lane_prefix(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ void decode(const uint2 e, const pb_box &box, int &t, double &s0,
                                       double &s1, double &s2)
{
    t = e.y & 63;
    s0 = (double)((int)((e.y >> 6) & 7) - 4) * box.L[0];
    s1 = (double)((int)((e.y >> 9) & 7) - 4) * box.L[1];
    s2 = (double)((int)((e.y >> 12) & 7) - 4) * box.L[2];
}

template <bool TABLE>
__device__ __forceinline__ void eval64(wave_lds &W, const uint2 e, const double4 *__restrict__ spos,
                                       const double4 *__restrict__ sfrc, const pb_box &box,
                                       const pb_kparams &kp, const pb_kparams *kpp,
                                       const double2 *__restrict__ tab)
{
    int t;
    double s0, s1, s2;
    decode(e, box, t, s0, s1, s2);
    double4 xs = spos[e.x];
    double4 fs = sfrc[e.x];
    double d0 = W.y[t][0] - (xs.x + s0);
    double d1 = W.y[t][1] - (xs.y + s1);
    double d2 = W.y[t][2] - (xs.z + s2);
    double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double H1, H2;
    if (TABLE) {
        if (r >= kp.rcommon) {
            pbdev::table_H(r, kp, tab, H1, H2);
        } else {
            double H[2];
            pbdev::vR_H_call(r, kpp, H);
            H1 = H[0];
            H2 = H[1];
        }
    } else {
        pbdev::vR_H_split(r, kp, H1, H2);
    }
    double g = H2 * (d0 * fs.x + d1 * fs.y + d2 * fs.z);
    atomicAdd(&W.acc[t][0], H1 * fs.x + g * d0);
    atomicAdd(&W.acc[t][1], H1 * fs.y + g * d1);
    atomicAdd(&W.acc[t][2], H1 * fs.z + g * d2);
}

template <bool TABLE>
__device__ __forceinline__ void eval32(wave_lds &W, const uint2 e, const double4 *__restrict__ spos,
                                       const double4 *__restrict__ sfrc, const pb_box &box,
                                       const pb_kparams &kp, const float2 *__restrict__ tab32)
{
    int t;
    double s0, s1, s2;
    decode(e, box, t, s0, s1, s2);
    double4 xs = spos[e.x];
    double4 fs = sfrc[e.x];
    float d0 = (float)(W.y[t][0] - (xs.x + s0));
    float d1 = (float)(W.y[t][1] - (xs.y + s1));
    float d2 = (float)(W.y[t][2] - (xs.z + s2));
    float f0 = (float)fs.x, f1 = (float)fs.y, f2 = (float)fs.z;
    float r = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    float H1, H2;
    if (TABLE)
        pbdev::table_H32(r, kp, tab32, H1, H2);
    else
        pbdev::diff_far32(r, kp.pf, H1, H2);
    float g = H2 * (d0 * f0 + d1 * f1 + d2 * f2);
    atomicAdd(&W.acc[t][0], (double)(H1 * f0 + g * d0));
    atomicAdd(&W.acc[t][1], (double)(H1 * f1 + g * d1));
    atomicAdd(&W.acc[t][2], (double)(H1 * f2 + g * d2));
}

template <bool MIXED, bool TABLE>
__global__ __launch_bounds__(64 * kWaves) void k_pair(
    const pb_tile *__restrict__ tiles, const int *__restrict__ ntiles_ptr,
    const int *__restrict__ cstart, const double4 *__restrict__ spos,
    const double4 *__restrict__ sfrc, const double4 *__restrict__ tpos,
    const int *__restrict__ tperm, pb_cells cells, pb_box box,
    const pb_kparams *__restrict__ kpp, double *__restrict__ u,
    unsigned long long *__restrict__ counters)
{
    const pb_kparams &kp = *kpp;
    __shared__ wave_lds lds[kWaves];
    extern __shared__ double dyn_lds[]; // the radial table: FP64 then (mixed) FP32 copy
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntab = TABLE ? kp.tn * (PB_TAB_DEG + 1) * 2 : 0;
    if (TABLE) { // every wavefront takes part before any may leave
        for (int k = threadIdx.x; k < ntab; k += blockDim.x) dyn_lds[k] = kp.tab64[k];
        if (MIXED) {
            float *t32 = reinterpret_cast<float *>(dyn_lds + ntab);
            for (int k = threadIdx.x; k < ntab; k += blockDim.x) t32[k] = kp.tab32[k];
        }
        __syncthreads();
    }
    const double2 *tab = reinterpret_cast<const double2 *>(dyn_lds);
    const float2 *tab32 = reinterpret_cast<const float2 *>(dyn_lds + ntab);
    // work item = (tile, column of the stencil): the (2st0+1)(2st1+1) columns
    // of one tile run on different wavefronts and meet in u through atomics
    const int ncol = (2 * cells.st[0] + 1) * (2 * cells.st[1] + 1);
    const long long item = (long long)blockIdx.x * kWaves + wave;
    const int tile_id = (int)(item / ncol), col = (int)(item % ncol);
    if (tile_id >= *ntiles_ptr) return; // whole wavefront leaves; no block barriers below
    wave_lds &W = lds[wave];
    const pb_tile T = tiles[tile_id];
    const bool active = lane < T.count;
    double y0 = 0, y1 = 0, y2 = 0;
    if (active) {
        double4 p = tpos[T.first + lane];
        y0 = p.x;
        y1 = p.y;
        y2 = p.z;
    }
    W.y[lane][0] = y0;
    W.y[lane][1] = y1;
    W.y[lane][2] = y2;
    W.acc[lane][0] = W.acc[lane][1] = W.acc[lane][2] = 0.0;
    __builtin_amdgcn_wave_barrier();

    const int nc0 = cells.nc[0], nc1 = cells.nc[1], nc2 = cells.nc[2];
    const int c2 = T.cell % nc2, c1 = (T.cell / nc2) % nc1, c0 = T.cell / (nc1 * nc2);
    const int o0 = col / (2 * cells.st[1] + 1) - cells.st[0];
    const int o1 = col % (2 * cells.st[1] + 1) - cells.st[1];
    const int u0 = c0 + o0, w0 = ((u0 % nc0) + nc0) % nc0, l0 = (u0 - w0) / nc0;
    const int u1 = c1 + o1, w1 = ((u1 % nc1) + nc1) % nc1, l1 = (u1 - w1) / nc1;
    const double yy0 = y0 - l0 * box.L[0], yy1 = y1 - l1 * box.L[1];
    const unsigned code01 = (unsigned)lane | ((unsigned)(l0 + 4) << 6) | ((unsigned)(l1 + 4) << 9);
    const int colbase = (w0 * nc1 + w1) * nc2;
    const double rc2 = cells.rc2;

    // iterator over the cells of this column (o2) and their sources: sources
    // are fetched 64 at a time, one per lane (coalesced), and broadcast to the
    // wavefront one by one with readlane for the 64-target test
    int o2 = -cells.st[2] - 1, j = 0, je = 0, k = 0, nk = 0;
    double yy2 = 0, cx = 0, cy = 0, cz = 0;
    unsigned code = 0;
    bool done = false;
    int qh = 0, qn = 0, fh = 0, fn = 0;
    unsigned long long n64 = 0, n32 = 0;
    for (;;) {
        // fill: test candidates until a ring holds a full batch or the column ends
        while (!done && qn < 64 && (!MIXED || fn < 64)) {
            if (k >= nk) {
                if (j >= je) {
                    if (++o2 > cells.st[2]) {
                        done = true;
                        break;
                    }
                    const int uu = c2 + o2, w2 = ((uu % nc2) + nc2) % nc2, l2 = (uu - w2) / nc2;
                    j = cstart[colbase + w2];
                    je = cstart[colbase + w2 + 1];
                    yy2 = y2 - l2 * box.L[2];
                    code = code01 | ((unsigned)(l2 + 4) << 12);
                    continue;
                }
                nk = min(64, je - j);
                if (lane < nk) {
                    const double4 xs = spos[j + lane];
                    cx = xs.x;
                    cy = xs.y;
                    cz = xs.z;
                }
                k = 0;
            }
            const double sx = lane_bcast(cx, k), sy = lane_bcast(cy, k), sz = lane_bcast(cz, k);
            const double d0 = yy0 - sx, d1 = yy1 - sy, d2 = yy2 - sz;
            const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
            const bool in = active && r2 < rc2;
            const unsigned jk = (unsigned)(j + k);
            if (!MIXED) {
                const unsigned long long m = __ballot(in);
                if (in) W.q64[(qh + qn + lane_prefix(m)) & (kQ - 1)] = make_uint2(jk, code);
                qn += __popcll(m);
            } else {
                const bool far = in && r2 >= kp.rfar2;
                const bool near = in && !far;
                const unsigned long long mn = __ballot(near), mf = __ballot(far);
                if (near) W.q64[(qh + qn + lane_prefix(mn)) & (kQ - 1)] = make_uint2(jk, code);
                if (far) W.q32[(fh + fn + lane_prefix(mf)) & (kQ - 1)] = make_uint2(jk, code);
                qn += __popcll(mn);
                fn += __popcll(mf);
            }
            if (++k >= nk) j += nk;
        }
        // evaluate: one batch per ring that is full (or, at the end, non-empty)
        const int cnt = qn >= 64 ? 64 : (done ? qn : 0);
        if (cnt > 0) {
            __builtin_amdgcn_wave_barrier();
            if (lane < cnt) eval64<TABLE>(W, W.q64[(qh + lane) & (kQ - 1)], spos, sfrc, box, kp, kpp, tab);
            qh = (qh + cnt) & (kQ - 1);
            qn -= cnt;
            n64 += cnt;
        }
        if (MIXED) {
            const int fc = fn >= 64 ? 64 : (done ? fn : 0);
            if (fc > 0) {
                __builtin_amdgcn_wave_barrier();
                if (lane < fc) eval32<TABLE>(W, W.q32[(fh + lane) & (kQ - 1)], spos, sfrc, box, kp, tab32);
                fh = (fh + fc) & (kQ - 1);
                fn -= fc;
                n32 += fc;
            }
        }
        if (done && qn == 0 && (!MIXED || fn == 0)) break;
    }
    __builtin_amdgcn_wave_barrier();
    if (active) {
        const size_t o = 3 * (size_t)tperm[T.first + lane];
        atomicAdd(&u[o + 0], W.acc[lane][0]);
        atomicAdd(&u[o + 1], W.acc[lane][1]);
        atomicAdd(&u[o + 2], W.acc[lane][2]);
    }
    if (lane == 0 && counters) {
        atomicAdd(&counters[0], n64);
        atomicAdd(&counters[1], n32);
    }
}

} // namespace

int pb_run_real(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_kparams *kp,
                size_t table_doubles, const pb_box &box, int ns, int nt, double *u, bool accumulate,
                int precision)
{
    (void)ns;
    int st;
    if ((st = pb_reserve(c, &c->counters, 4 * sizeof(unsigned long long)))) return st;
    if ((st = pb_hip_check(c, hipMemsetAsync(c->counters.ptr, 0, 4 * sizeof(unsigned long long), s),
                           "memset counters")))
        return st;
    if (nt <= 0) return PB_OK;
    if (!accumulate &&
        (st = pb_hip_check(c, hipMemsetAsync(u, 0, 3 * (size_t)nt * sizeof(double), s), "memset u")))
        return st;
    size_t max_tiles = (size_t)cells.ncell + (size_t)nt / 64 + 1;
    size_t ncol = (size_t)(2 * cells.st[0] + 1) * (2 * cells.st[1] + 1);
    unsigned blocks = (unsigned)((max_tiles * ncol + kWaves - 1) / kWaves);
    const int *ntiles = (const int *)c->tile_off.ptr + cells.ncell;
    pb_mark(c, s, PB_NPHASE - 1); // the pair-kernel interval covers only this launch
    const bool mixed = precision == PB_PREC_MIXED;
    const size_t ntab = table_doubles; // 0: closed forms per pair
    const size_t lds = ntab ? ntab * sizeof(double) + (mixed ? ntab * sizeof(float) : 0) : 0;
#define PB_PAIR_ARGS                                                                                  \
    (const pb_tile *)c->tiles.ptr, ntiles, (const int *)c->cell_start.ptr, (const double4 *)c->spos.ptr, \
        (const double4 *)c->sfrc.ptr, (const double4 *)c->tpos.ptr, (const int *)c->tperm.ptr, cells, box, \
        kp, u, (unsigned long long *)c->counters.ptr
    dim3 gd(blocks), bd(64 * kWaves);
    if (mixed && ntab) hipLaunchKernelGGL((k_pair<true, true>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (mixed) hipLaunchKernelGGL((k_pair<true, false>), gd, bd, 0, s, PB_PAIR_ARGS);
    else if (ntab) hipLaunchKernelGGL((k_pair<false, true>), gd, bd, lds, s, PB_PAIR_ARGS);
    else hipLaunchKernelGGL((k_pair<false, false>), gd, bd, 0, s, PB_PAIR_ARGS);
#undef PB_PAIR_ARGS
    return pb_hip_check(c, hipGetLastError(), "pair kernel launch");
}
