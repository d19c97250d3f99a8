// pb_real.hip -- real-space image sum over cell-listed neighbours.
//
// u(y_t) += sum over source images with |y_t - x_j - l L| < rc of
//           K_R(y_t - x_j - l L) f_j,   K_R = S from diffH(d) - diffH(e)
// (periodicBrinkman.h:152-176; the reference sums all pairs over image cubes,
// realShells :564-693, kernelReal :219-314).
//
// Work item = (tile, split): a tile is <= 64 targets of one quarter brick
// (2x2x4 cells; half a brick or a brick where targets are sparse), one
// target per lane; its bounding box B decides the
// candidate sources: the (x, y) cell columns within rc of B, each trimmed in
// z to sqrt(rc^2 - gap^2), which are contiguous runs of the z-major sorted
// sources.  The columns are dealt round-robin to `nsplit` wavefronts, which
// meet in u through FP64 atomics.  The lanes of a wavefront set up 64 of its
// columns in parallel (culling, image shifts, first run of sources); the
// candidates of the lanes' current runs, numbered in lane order, are loaded
// 64 at a time (one per lane, a round drawing on as many columns as it takes
// to fill), kept when within rc of the tile box, and staged in LDS.
//
// Each staged candidate is broadcast to the 64 lane targets and tested in
// FP32 (positions relative to the tile's lower corner, rc enlarged by 1e-4
// relative, so every pair in range passes); the passing (target, source, image) triples are
// compacted with a ballot into an LDS ring, one LDS write per tested source;
// each time 64 are pending every lane evaluates one, so the evaluation runs
// with full lanes however sparse the test is.  A batch applies the exact
// cutoff r < rc in FP64 and moves its rare pairs with r < rcommon (series and
// blob regions) to a third ring that is evaluated 64 at a time out of line.
// Results go to per-target FP64 accumulators in LDS (ds_add_f64).  With an
// FP32 far field (the MIXED template), pairs beyond rfar (>= 3d and >= 8e) by
// the FP32 distance go to the second half of the ring and are evaluated in
// FP32, and still accumulated in FP64.
#include "pb_radial.h"

namespace {

constexpr int kWaves = 4; // wavefronts per workgroup, each independent
static_assert(kWaves == pb_pair_waves && kWaves == 4, "pair workgroup size");
constexpr int kQ = 128;   // ring capacity (power of two)



struct wave_lds {
    double y[64][3];
    double acc[64][3];
    float4 stage[64];  // staged candidates: c = position - tile origin (FP32), |c|^2
    uint2 stage_jc[64]; // their source indices and image codes ((l0+4) | (l1+4) << 3 | (l2+4) << 6) << 6
    uint2 q[2 * kQ];   // [0, kQ): FP64 ring; [kQ, 2 kQ): FP32 far ring (without one: all FP64)
    uint2 qr[kQ];      // r < rcommon, moved out of evaluated FP64 batches
    union {
        int2 col_ej[64];  // per column lane: first candidate of its run in the round, first source of the run
        uint2 trash[64];  // (while testing) where a lane whose test fails writes: the ring write takes no branch
    };
    unsigned col_code[64]; // per column lane: image code of its run
};

// Element j of a source array through a 32-bit byte offset on the array's
// (uniform) base: a scalar-base load, no 64-bit address arithmetic per lane
// (source counts stay far below 2^27).
__device__ __forceinline__ double4 src_at(const double4 *__restrict__ a, unsigned j)
{
    return *reinterpret_cast<const double4 *>(reinterpret_cast<const char *>(a) + (j << 5));
}

// Inclusive prefix sum over the 64 lanes by DPP (shifts within rows of 16,
// then the row totals broadcast to the rows above): no LDS round trips.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); // row_bcast:31 into rows 2, 3
    return v;
}

__device__ __forceinline__ unsigned lane_prefix(unsigned long long m, unsigned base = 0u)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, base));
}

// Uniform copy of a value every lane holds (kept in SGPRs).
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

// ring entry: x = source index, y = lane | (l0+4) << 6 | (l1+4) << 9 | (l2+4) << 12
__device__ __forceinline__ void decode(const uint2 e, const pb_box &box, int &t, double &s0,
                                       double &s1, double &s2)
{
    t = e.y & 63;
    s0 = (double)((int)((e.y >> 6) & 7) - 4) * box.L[0];
    s1 = (double)((int)((e.y >> 9) & 7) - 4) * box.L[1];
    s2 = (double)((int)((e.y >> 12) & 7) - 4) * box.L[2];
}

__device__ __forceinline__ void accumulate(wave_lds &W, int t, double H1, double H2, double d0,
                                           double d1, double d2, const double4 &fs)
{
    const double g = H2 * (d0 * fs.x + d1 * fs.y + d2 * fs.z);
    atomicAdd(&W.acc[t][0], H1 * fs.x + g * d0);
    atomicAdd(&W.acc[t][1], H1 * fs.y + g * d1);
    atomicAdd(&W.acc[t][2], H1 * fs.z + g * d2);
}

// FP64 batch of cnt entries from the ring at head.  Pairs with r < rcommon
// go to the rare ring; the others take the common far region of K_R (only
// diffH(d)), from the table or the closed form.
// Kernel constants the batches use, copied into registers once.
struct batch_consts {
    pbdev::tabview tv;
    pbdev::tabview32 tv32;
    pbdev::radial_core core;
    pb_radial32 pf;
    double rcommon, rcommon2;
};

template <bool TABLE, int DEG, int RM>
__device__ __forceinline__ void batch64(wave_lds &W, const int lane, const int cnt, const int head,
                                        const int rh, int &rn, unsigned &nin, const double4 *__restrict__ spos,
                                        const double4 *__restrict__ sfrc, const pb_box &box, const double rc2,
                                        const batch_consts &K, const double2 *__restrict__ tab)
{
    __builtin_amdgcn_wave_barrier();
    bool on = lane < cnt;
    uint2 e = make_uint2(0u, 0u);
    int t = 0;
    double d0 = 0, d1 = 0, d2 = 0, r2 = 0;
    double4 fs = make_double4(0, 0, 0, 0);
    if (on) {
        e = W.q[(head + lane) & RM];
        double s0, s1, s2;
        decode(e, box, t, s0, s1, s2);
        const double4 xs = src_at(spos, e.x);
        fs = src_at(sfrc, e.x);
        d0 = W.y[t][0] - (xs.x + s0);
        d1 = W.y[t][1] - (xs.y + s1);
        d2 = W.y[t][2] - (xs.z + s2);
        r2 = d0 * d0 + d1 * d1 + d2 * d2;
        on = r2 < rc2; // the exact cutoff (the FP32 candidate test lets a few more through)
    }
    nin += __popcll(__ballot(on));
    const bool rare = on && r2 < K.rcommon2; // (r2 = 0, a target on a source, is rare)
    const unsigned long long mr = __ballot(rare);
    if (mr) {
        if (rare) W.qr[(rh + rn + lane_prefix(mr)) & (kQ - 1)] = e;
        rn += __popcll(mr);
    }
    if (on && !rare) {
        const double r = pbdev::sqrt_pos(r2); // r2 >= rcommon^2 > 0
        double H1, H2;
        if (TABLE)
            pbdev::table_H<DEG>(r, K.tv, tab, H1, H2);
        else
            pbdev::diff_direct(r, r * K.core.inv_d, K.core, H1, H2);
        accumulate(W, t, H1, H2, d0, d1, d2, fs);
    }
}

// Rare batch: every region of K_R, through an out-of-line call so its library
// constants stay out of the main loop.
__device__ __forceinline__ void batch_rare(wave_lds &W, const int lane, const int cnt, const int head,
                                           const double4 *__restrict__ spos,
                                           const double4 *__restrict__ sfrc, const pb_box &box,
                                           const pb_kparams *kpp)
{
    __builtin_amdgcn_wave_barrier();
    if (lane < cnt) {
        const uint2 e = W.qr[(head + lane) & (kQ - 1)];
        int t;
        double s0, s1, s2;
        decode(e, box, t, s0, s1, s2);
        const double4 xs = src_at(spos, e.x);
        const double4 fs = src_at(sfrc, e.x);
        const double d0 = W.y[t][0] - (xs.x + s0);
        const double d1 = W.y[t][1] - (xs.y + s1);
        const double d2 = W.y[t][2] - (xs.z + s2);
        const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const double2 H = pbdev::vR_H_call(r, kpp);
        accumulate(W, t, H.x, H.y, d0, d1, d2, fs);
    }
}

// FP32 batch of the far ring (r >= rfar: only diffH(d), in FP32).
template <bool TABLE, int DEG>
__device__ __forceinline__ void batch32(wave_lds &W, const int lane, const int cnt, const int head, unsigned &nin,
                                        const double4 *__restrict__ spos,
                                        const double4 *__restrict__ sfrc, const pb_box &box, const double rc2,
                                        const batch_consts &K, const float2 *__restrict__ tab32)
{
    __builtin_amdgcn_wave_barrier();
    bool on = lane < cnt;
    uint2 e = make_uint2(0u, 0u);
    int t = 0;
    double dd0 = 0, dd1 = 0, dd2 = 0;
    if (on) {
        e = W.q[kQ + ((head + lane) & (kQ - 1))];
        double s0, s1, s2;
        decode(e, box, t, s0, s1, s2);
        const double4 xs = src_at(spos, e.x);
        dd0 = W.y[t][0] - (xs.x + s0);
        dd1 = W.y[t][1] - (xs.y + s1);
        dd2 = W.y[t][2] - (xs.z + s2);
        on = dd0 * dd0 + dd1 * dd1 + dd2 * dd2 < rc2; // the exact cutoff
    }
    nin += __popcll(__ballot(on));
    if (on) {
        const double4 fs = src_at(sfrc, e.x);
        const float d0 = (float)dd0, d1 = (float)dd1, d2 = (float)dd2;
        const float f0 = (float)fs.x, f1 = (float)fs.y, f2 = (float)fs.z;
        const float r = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
        float H1, H2;
        if (TABLE)
            pbdev::table_H32<DEG>(r, K.tv32, tab32, H1, H2);
        else
            pbdev::diff_far32(r, K.pf, H1, H2);
        const float g = H2 * (d0 * f0 + d1 * f1 + d2 * f2);
        atomicAdd(&W.acc[t][0], (double)(H1 * f0 + g * d0));
        atomicAdd(&W.acc[t][1], (double)(H1 * f1 + g * d1));
        atomicAdd(&W.acc[t][2], (double)(H1 * f2 + g * d2));
    }
}

template <bool MIXED, bool TABLE, int DEG>
__device__ __forceinline__ void pair_body(
    const int blk, const int2 *__restrict__ tiles, const int *__restrict__ ntiles_ptr,
    const int *__restrict__ cstart, const double4 *__restrict__ spos,
    const double4 *__restrict__ sfrc, const double4 *__restrict__ tpos,
    const int *__restrict__ tperm, int nsplit, pb_cells cells, pb_box box,
    const pb_kparams *__restrict__ kpp, double *__restrict__ u, uint4 *__restrict__ stats)
{
    const pb_kparams &kp = *kpp;
    const batch_consts K{pbdev::tabview_of(kp), pbdev::tabview32_of(kp), pbdev::core_of(kp.pd), kp.pf,
                         kp.rcommon, kp.rcommon * kp.rcommon};
    const double rfar2 = kp.rfar2;
    __shared__ wave_lds lds[kWaves];
    extern __shared__ double dyn_lds[]; // the radial table: FP64 then (FP32 far field) FP32 copy
    const int ntab = TABLE ? kp.tn * (DEG + 1) * 2 : 0;
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

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long item = (long long)blk * kWaves + wave;
    const int tile = (int)(item / nsplit), split = (int)(item % nsplit);
    if (tile >= *ntiles_ptr) { // whole wavefront leaves (with nsplit = kWaves the whole workgroup: the only
                               // block barrier, at the end, is taken only then)
        if (lane == 0) stats[item] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const int2 T = tiles[tile];
    const int t0 = T.x;
    wave_lds &W = lds[wave];
    const bool active = lane < T.y;
    const double4 p = tpos[active ? t0 + lane : t0];
    W.y[lane][0] = p.x;
    W.y[lane][1] = p.y;
    W.y[lane][2] = p.z;
    W.acc[lane][0] = W.acc[lane][1] = W.acc[lane][2] = 0.0;
    // bounding box of the tile (inactive lanes repeat the first target)
    const double bmin0 = wave_min(p.x), bmin1 = wave_min(p.y), bmin2 = wave_min(p.z);
    const double bmax0 = wave_max(p.x), bmax1 = wave_max(p.y), bmax2 = wave_max(p.z);
    // the candidate test runs in FP32 on positions relative to the tile's
    // lower corner, against rc enlarged by 1e-4 (relative) so that every pair
    // in range passes; the batches apply the exact cutoff in FP64.  With c a
    // candidate and y the lane's target, |y - c|^2 < R^2 is tested as
    // |c|^2 - 2 c.y < R^2 - |y|^2: three FMAs per candidate (|c|^2 staged with
    // it), the right side per lane.  Coordinates are below ~2.5 rc, so the
    // rounding of the expanded form stays near 1e-6 rc^2, far inside the
    // 1e-4 enlargement.  Inactive lanes get a bound no candidate meets.
    const float y0 = (float)(p.x - bmin0), y1 = (float)(p.y - bmin1), y2 = (float)(p.z - bmin2);
    const float ty0 = -2.0f * y0, ty1 = -2.0f * y1, ty2 = -2.0f * y2, yy = y0 * y0 + y1 * y1 + y2 * y2;
    __builtin_amdgcn_wave_barrier();

    const int nc0 = cells.nc[0], nc1 = cells.nc[1], nc2 = cells.nc[2];
    const double rcx = cells.rc * (1.0 + 1e-12); // culling is conservative; the test is exact
    const double rcx2 = rcx * rcx;
    const int ulo0 = (int)floor((bmin0 - rcx) * cells.inv_h[0]), uhi0 = (int)floor((bmax0 + rcx) * cells.inv_h[0]);
    const int ulo1 = (int)floor((bmin1 - rcx) * cells.inv_h[1]), uhi1 = (int)floor((bmax1 + rcx) * cells.inv_h[1]);
    const int ny = uhi1 - ulo1 + 1;
    const int ncols = (uhi0 - ulo0 + 1) * ny;
    const double rc2 = cells.rc2;
    const float rc2f = (float)(rc2 * (1.0 + 1e-4)), rfar2f = (float)rfar2;
    const float qin = active ? rc2f - yy : -1e30f, qfar = rfar2f - yy; // test bounds on |c|^2 - 2 c.y

    // FP64 ring index mask: kQ entries beside the far ring, else 2 kQ (the
    // ring is then checked every two candidates: at most 63 + 128 pending)
    constexpr int RM = MIXED ? kQ - 1 : 2 * kQ - 1;
    int qh = 0, qn = 0, fh = 0, fn = 0, rh = 0, rn = 0;
    unsigned n64 = 0, n32 = 0, ntest = 0, nread = 0;
    // evaluate every full ring (or, at the end, every non-empty one)
    auto flush = [&](const bool last) {
        while (qn >= 64 || (last && qn > 0)) {
            const int c = qn < 64 ? qn : 64;
            batch64<TABLE, DEG, RM>(W, lane, c, qh, rh, rn, n64, spos, sfrc, box, rc2, K, tab);
            qh = (qh + c) & RM;
            qn -= c;
            if (rn >= 64) { // a batch adds at most 64: the rare ring (kQ) never overflows
                batch_rare(W, lane, 64, rh, spos, sfrc, box, kpp);
                rh = (rh + 64) & (kQ - 1);
                rn -= 64;
            }
        }
        while (rn >= 64 || (last && rn > 0)) {
            const int c = rn < 64 ? rn : 64;
            batch_rare(W, lane, c, rh, spos, sfrc, box, kpp);
            rh = (rh + c) & (kQ - 1);
            rn -= c; // (counted when classified in batch64)
        }
        if (MIXED) {
            while (fn >= 64 || (last && fn > 0)) {
                const int c = fn < 64 ? fn : 64;
                batch32<TABLE, DEG>(W, lane, c, fh, n32, spos, sfrc, box, rc2, K, tab32);
                fh = (fh + c) & (kQ - 1);
                fn -= c;
            }
        }
    };

    // append the lanes of one tested candidate that pass the FP32 test to the
    // rings (with the FP32 far field: r >= rfar by the FP32 distance to the
    // far ring; the FP32 decision near rfar only picks the precision)
    auto append = [&](const unsigned jk, const float q, const unsigned code) {
        const bool in = q < qin;
        const unsigned long long m = __ballot(in);
        if (!MIXED) { // (the caller flushes)
            const unsigned slot = lane_prefix(m, (unsigned)(qh + qn));
            uint2 *dst = in ? &W.q[slot & RM] : &W.trash[lane];
            *dst = make_uint2(jk, code);
            qn += __popcll(m);
        } else {
            const bool far = in && q >= qfar;
            const unsigned long long mf = __ballot(far);
            const unsigned pm = lane_prefix(m), pf = lane_prefix(mf);
            if (in) {
                const unsigned v = far ? (((unsigned)(fh + fn) + pf) & (kQ - 1)) | kQ
                                       : ((unsigned)(qh + qn) + pm - pf) & (kQ - 1);
                W.q[v] = make_uint2(jk, code);
            }
            const int nf = __popcll(mf);
            qn += __popcll(m) - nf;
            fn += nf;
            if (qn >= 64 || fn >= 64) flush(false);
        }
    };

    // this wavefront's columns are split, split + nsplit, ...; the lanes set
    // up 64 of them at a time, one column per lane.  A column's z range runs
    // through one or more periodic images, each a run of consecutive sources.
    // The candidates of the lanes' current runs are numbered in lane order and
    // staged 64 at a time, so a staging round draws on many columns (the runs
    // are short: a few sources each at the benchmark density).
    for (int cb = split; cb < ncols; cb += 64 * nsplit) {
        const int colv = cb + nsplit * lane;
        int ze = -1, l0 = 0, l1 = 0, colbase = 0, l2 = 0, zend = -1, j = 0, je = 0;
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
                j = cstart[colbase + zs - l2 * nc2];
                je = cstart[colbase + zend - l2 * nc2 + 1];
            }
        }
        const unsigned code01 = ((unsigned)(l0 + 4) | ((unsigned)(l1 + 4) << 3)) << 6;
        for (;;) {
            // lanes whose run is used up move to the next image's run of their column
            bool adv = ok && j >= je && zend < ze;
            while (__ballot(adv)) {
                if (adv) {
                    const int zsn = zend + 1;
                    l2 = floordiv(zsn, nc2);
                    zend = min(ze, (l2 + 1) * nc2 - 1);
                    j = cstart[colbase + zsn - l2 * nc2];
                    je = cstart[colbase + zend - l2 * nc2 + 1];
                }
                adv = ok && j >= je && zend < ze;
            }
            const int len = ok && je > j ? je - j : 0;
            const unsigned long long nz = __ballot(len > 0);
            if (!nz) break;
            const unsigned rcode = code01 | ((unsigned)(l2 + 4) << 12);
            // the first column with candidates (uniform): when it alone holds
            // the round's candidates (long runs: dense sources) they are its
            // next nk sources; else the candidates are numbered over the
            // lanes' runs (prefix sum) and each lane finds its column
            const int c0 = __builtin_ctzll(nz);
            const int len0 = __builtin_amdgcn_readlane(len, c0);
            int nk = 64, excl = 0;
            bool one = len0 >= 64;
            if (!one) {
                const int incl = wave_incl_scan(len);
                nk = min(64, __builtin_amdgcn_readlane(incl, 63));
                excl = incl - len;
                one = len0 >= nk;
                if (!one) {
                    __builtin_amdgcn_wave_barrier();
                    W.col_ej[lane] = make_int2(excl, j);
                    W.col_code[lane] = rcode;
                    __builtin_amdgcn_wave_barrier();
                }
            }
            bool keep = false;
            double sx = 0, sy = 0, sz = 0;
            int src = 0;
            unsigned ccode = 0;
            if (lane < nk) {
                if (one) {
                    src = __builtin_amdgcn_readlane(j, c0) + lane;
                    ccode = (unsigned)__builtin_amdgcn_readlane((int)rcode, c0);
                } else {
                    // candidate `lane`: the last column whose first candidate is
                    // at or before it (columns without candidates share the
                    // next one's start and are passed over)
                    int c = 0; // (c + step stays at or below 63)
#pragma unroll
                    for (int step = 32; step >= 1; step >>= 1)
                        if (W.col_ej[c + step].x <= lane) c += step;
                    const int2 ej = W.col_ej[c];
                    src = ej.y + (lane - ej.x);
                    ccode = W.col_code[c];
                }
                const double4 xs = spos[src];
                sx = xs.x + (double)((int)((ccode >> 6) & 7) - 4) * box.L[0];
                sy = xs.y + (double)((int)((ccode >> 9) & 7) - 4) * box.L[1];
                sz = xs.z + (double)((int)((ccode >> 12) & 7) - 4) * box.L[2];
                const double g0 = fmax(0.0, fmax(bmin0 - sx, sx - bmax0));
                const double g1 = fmax(0.0, fmax(bmin1 - sy, sy - bmax1));
                const double g2 = fmax(0.0, fmax(bmin2 - sz, sz - bmax2));
                keep = g0 * g0 + g1 * g1 + g2 * g2 < rcx2;
            }
            // the lanes' runs move past the candidates this round takes
            j += one ? (lane == c0 ? nk : 0) : min(len, max(0, 64 - excl));
            const unsigned long long km = __ballot(keep);
            const int nkeep = __popcll(km);
            nread += nk;
            ntest += nkeep;
            __builtin_amdgcn_wave_barrier();
            if (keep) {
                const float c0f = (float)(sx - bmin0), c1f = (float)(sy - bmin1), c2f = (float)(sz - bmin2);
                const unsigned slot = lane_prefix(km);
                W.stage[slot] = make_float4(c0f, c1f, c2f, c0f * c0f + c1f * c1f + c2f * c2f);
                W.stage_jc[slot] = make_uint2((unsigned)src, ccode);
            }
            __builtin_amdgcn_wave_barrier();
            // two candidates per step: independent dependency chains
            int k = 0;
            for (; k + 1 < nkeep; k += 2) {
                const float4 sa = W.stage[k], sb = W.stage[k + 1]; // broadcast reads
                const uint4 jc = *reinterpret_cast<const uint4 *>(&W.stage_jc[k]);
                const float qa = fmaf(ty0, sa.x, fmaf(ty1, sa.y, fmaf(ty2, sa.z, sa.w)));
                const float qb = fmaf(ty0, sb.x, fmaf(ty1, sb.y, fmaf(ty2, sb.z, sb.w)));
                append(jc.x, qa, jc.y | (unsigned)lane);
                append(jc.z, qb, jc.w | (unsigned)lane);
                if (!MIXED && qn >= 64) flush(false);
            }
            if (k < nkeep) {
                const float4 sa = W.stage[k];
                const uint2 jc = W.stage_jc[k];
                append(jc.x, fmaf(ty0, sa.x, fmaf(ty1, sa.y, fmaf(ty2, sa.z, sa.w))), jc.y | (unsigned)lane);
                if (!MIXED && qn >= 64) flush(false);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    flush(true);
    __builtin_amdgcn_wave_barrier();
    if (nsplit == kWaves) {
        // the workgroup's wavefronts share the tile (and left together above
        // when there was none): one of them adds the four partial sums to u
        __syncthreads();
        if (wave == 0 && active) {
            const size_t o = 3 * (size_t)tperm[t0 + lane];
#pragma unroll
            for (int c = 0; c < 3; c++)
                atomicAdd(&u[o + c], (lds[0].acc[lane][c] + lds[1].acc[lane][c]) +
                                         (lds[2].acc[lane][c] + lds[3].acc[lane][c]));
        }
    } else if (active && (n64 | n32)) {
        const size_t o = 3 * (size_t)tperm[t0 + lane];
        atomicAdd(&u[o + 0], W.acc[lane][0]);
        atomicAdd(&u[o + 1], W.acc[lane][1]);
        atomicAdd(&u[o + 2], W.acc[lane][2]);
    }
    // per-item statistics: one plain store each (same-address atomics from
    // every wavefront would serialize in L2)
    if (lane == 0) stats[item] = make_uint4(n64, n32, ntest, nread);
}

__global__ void k_zero(long long n, double *__restrict__ a)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

#define PB_PAIR_PARAMS                                                                                 \
    const int2 *__restrict__ tiles, const int *__restrict__ ntiles_ptr, const int *__restrict__ cstart,  \
        const double4 *__restrict__ spos, const double4 *__restrict__ sfrc,                              \
        const double4 *__restrict__ tpos, const int *__restrict__ tperm, int nsplit, pb_cells cells,      \
        pb_box box, const pb_kparams *__restrict__ kpp, double *__restrict__ u, uint4 *__restrict__ stats
#define PB_PAIR_FWD (int)blockIdx.x, tiles, ntiles_ptr, cstart, spos, sfrc, tpos, tperm, nsplit, cells, box, kpp, u, stats

// The table path fits 4 wavefronts per SIMD (build-time PB_PAIR_OCC); the
// closed forms need more registers and run with the compiler's own choice.
#ifndef PB_PAIR_OCC
#define PB_PAIR_OCC 4
#endif
template <bool MIXED, int DEG>
__global__ __launch_bounds__(64 * kWaves, PB_PAIR_OCC) void k_pair_table(PB_PAIR_PARAMS)
{
    pair_body<MIXED, true, DEG>(PB_PAIR_FWD);
}

template <bool MIXED>
__global__ __launch_bounds__(64 * kWaves) void k_pair_direct(PB_PAIR_PARAMS)
{
    pair_body<MIXED, false, PB_TAB_DEG>(PB_PAIR_FWD);
}

// Grouped: every system of a class, block b on system map[b].x.
#define PB_GPAIR_FWD                                                                                    \
    m.y, S.tiles, S.ntiles, S.sstart, S.spos, S.sfrc, S.tpos, S.tperm, nsplit, S.cells, S.box, S.kp, S.u, S.stats

template <bool MIXED, int DEG>
__global__ __launch_bounds__(64 * kWaves, PB_PAIR_OCC) void k_gpair_table(const pb_sys *__restrict__ sys_,
                                                                          const int2 *__restrict__ map_, int nsplit)
{
    const int2 m = pb_map_at(map_, blockIdx.x);
    const auto &S = ((pb_const_sys)sys_)[m.x];
    pair_body<MIXED, true, DEG>(PB_GPAIR_FWD);
}

template <bool MIXED>
__global__ __launch_bounds__(64 * kWaves) void k_gpair_direct(const pb_sys *__restrict__ sys_,
                                                              const int2 *__restrict__ map_, int nsplit)
{
    const int2 m = pb_map_at(map_, blockIdx.x);
    const auto &S = ((pb_const_sys)sys_)[m.x];
    pair_body<MIXED, false, PB_TAB_DEG>(PB_GPAIR_FWD);
}
#undef PB_GPAIR_FWD

} // namespace

int pb_group_pair(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int nsplit, bool far32, int tdeg,
                  size_t lds)
{
    if (nblk <= 0) return PB_OK;
    dim3 gd(nblk), bd(64 * kWaves);
#define PB_GP(M, D) hipLaunchKernelGGL((k_gpair_table<M, D>), gd, bd, lds, s, sys, map, nsplit)
    if (tdeg == PB_TAB_DEG_MIXED && far32) PB_GP(true, PB_TAB_DEG_MIXED);
    else if (tdeg == PB_TAB_DEG_MID && far32) PB_GP(true, PB_TAB_DEG_MID);
    else if (tdeg == PB_TAB_DEG && far32) PB_GP(true, PB_TAB_DEG);
    else if (tdeg == PB_TAB_DEG_MIXED) PB_GP(false, PB_TAB_DEG_MIXED);
    else if (tdeg == PB_TAB_DEG_MID) PB_GP(false, PB_TAB_DEG_MID);
    else if (tdeg == PB_TAB_DEG) PB_GP(false, PB_TAB_DEG);
#undef PB_GP
    else if (tdeg == 0 && far32) hipLaunchKernelGGL((k_gpair_direct<true>), gd, bd, 0, s, sys, map, nsplit);
    else if (tdeg == 0) hipLaunchKernelGGL((k_gpair_direct<false>), gd, bd, 0, s, sys, map, nsplit);
    else return PB_ERR_INVALID;
    return hipGetLastError() == hipSuccess ? PB_OK : PB_ERR_HIP;
}

void pb_zero_u(hipStream_t s, int n, double *u)
{
    if (n <= 0) return;
    const long long n3 = 3LL * n;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, n3, u);
}

int pb_run_real(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_kparams *kp,
                size_t table_doubles, int table_deg, const pb_box &box, int ns, int nt, double *u, bool accumulate,
                bool far32, double rounds)
{
    (void)ns;
    int st;
    c->stat_items = 0;
    c->stat_group = false;
    if (nt <= 0) return PB_OK;
    if (!accumulate) pb_zero_u(s, nt, u);
    const int max_tiles = c->max_tiles; // tiles built with the target binning
    const int *ntiles = (const int *)c->qoff.ptr + cells.nkey_t / 16;
    const int nsplit = pb_pair_split(c->cu_count, nt, rounds);
    const unsigned blocks = (unsigned)(((long long)max_tiles * nsplit + kWaves - 1) / kWaves);
    if ((st = pb_reserve(c, &c->counters, sizeof(uint4) * (size_t)blocks * kWaves))) return st;
    c->stat_items = (long long)blocks * kWaves;
    const bool mixed = far32;
    const size_t ntab = table_doubles; // 0: closed forms per pair
    const size_t lds = ntab ? ntab * sizeof(double) + (mixed ? ntab * sizeof(float) : 0) : 0;
    pb_mark(c, s, PB_NPHASE - 1); // the pair-kernel interval covers only this launch
#define PB_PAIR_ARGS                                                                                   \
    (const int2 *)c->tiles.ptr, ntiles, (const int *)c->cell_start.ptr, (const double4 *)c->spos.ptr,   \
        (const double4 *)c->sfrc.ptr, (const double4 *)c->tpos.ptr, (const int *)c->tperm.ptr, nsplit,    \
        cells, box, kp, u, (uint4 *)c->counters.ptr
    dim3 gd(blocks), bd(64 * kWaves);
    if (ntab && table_deg != PB_TAB_DEG && table_deg != PB_TAB_DEG_MIXED && table_deg != PB_TAB_DEG_MID)
        return pb_fail(c, PB_ERR_INVALID, "radial table of an unsupported degree");
    const bool lo = table_deg == PB_TAB_DEG_MIXED, mid = table_deg == PB_TAB_DEG_MID;
    if (mixed && ntab && lo) hipLaunchKernelGGL((k_pair_table<true, PB_TAB_DEG_MIXED>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (mixed && ntab && mid) hipLaunchKernelGGL((k_pair_table<true, PB_TAB_DEG_MID>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (mixed && ntab) hipLaunchKernelGGL((k_pair_table<true, PB_TAB_DEG>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (mixed) hipLaunchKernelGGL((k_pair_direct<true>), gd, bd, 0, s, PB_PAIR_ARGS);
    else if (ntab && lo) hipLaunchKernelGGL((k_pair_table<false, PB_TAB_DEG_MIXED>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (ntab && mid) hipLaunchKernelGGL((k_pair_table<false, PB_TAB_DEG_MID>), gd, bd, lds, s, PB_PAIR_ARGS);
    else if (ntab) hipLaunchKernelGGL((k_pair_table<false, PB_TAB_DEG>), gd, bd, lds, s, PB_PAIR_ARGS);
    else hipLaunchKernelGGL((k_pair_direct<false>), gd, bd, 0, s, PB_PAIR_ARGS);
    pb_mark(c, s, PB_NPHASE); // the pair interval ends with the launch, on its own stream
#undef PB_PAIR_ARGS
    return pb_hip_check(c, hipGetLastError(), "pair kernel launch");
}
