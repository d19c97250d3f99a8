// pb_bin.hip -- cell binning of sources and targets (counting sort).
//
// Both use the same grid of fine cells (side >= rc/3).  Sources are ordered
// z-major, so the cells of one (x, y) column are one contiguous run of the
// sorted array.  Targets are ordered by a brick-Morton key (bricks of 4x4x4
// cells in z-major order, Morton order inside a brick), so the targets of a
// quarter brick (2x2x4 cells, 16 consecutive keys) are contiguous; the pair
// kernel's tiles are runs of <= 64 of them, or a whole half brick or brick
// where that holds <= 64 targets (quarter_tiles).
//
// Positions are wrapped into [0, L) with the same operation sequence as the
// oracle (v - floor(v/L)*L, no contraction) so both place a point on a cell
// or window boundary identically.
#include "pb_internal.h"

namespace {

__device__ __forceinline__ double wrap_coord(double v, double L)
{
    double w = __dsub_rn(v, __dmul_rn(floor(v / L), L));
    if (w >= L) w = __dsub_rn(w, L);
    if (w < 0.0) w = 0.0;
    return w;
}


__global__ void k_bin_count(int n, const double *__restrict__ pos, pb_box box, pb_cells cells,
                            int brick_order, int *__restrict__ count, int *__restrict__ key_of,
                            int *__restrict__ rank)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ci[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double w = wrap_coord(pos[3 * (size_t)i + c], box.L[c]);
        int k = (int)(w * cells.inv_h[c]);
        ci[c] = k >= cells.nc[c] ? cells.nc[c] - 1 : k;
    }
    const int key = brick_order ? pb_target_key(ci[0], ci[1], ci[2], cells)
                                : (ci[0] * cells.nc[1] + ci[1]) * cells.nc[2] + ci[2];
    key_of[i] = key;
    rank[i] = atomicAdd(&count[key], 1);
}

__global__ void k_bin_scatter(int n, const double *__restrict__ pos, const double *__restrict__ frc,
                              pb_box box, const int *__restrict__ start,
                              const int *__restrict__ key_of, const int *__restrict__ rank,
                              double4 *__restrict__ opos, double4 *__restrict__ ofrc,
                              int *__restrict__ operm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dst = start[key_of[i]] + rank[i];
    double w[3];
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = wrap_coord(pos[3 * (size_t)i + c], box.L[c]);
    opos[dst] = make_double4(w[0], w[1], w[2], 0.0);
    if (ofrc) ofrc[dst] = make_double4(frc[3 * (size_t)i], frc[3 * (size_t)i + 1], frc[3 * (size_t)i + 2], 0.0);
    if (operm) operm[dst] = i;
}

// Exclusive scan of n ints, out[n] = total, in three passes: per-block sums of
// 1024 elements, a one-workgroup scan of those, then per-block scans with the
// block offsets added.
constexpr int kScanBlock = 1024;

__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int w = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += w;
        __syncthreads();
    }
    total = sh[255];
    int ex = sh[t] - v;
    __syncthreads();
    return ex;
}

__global__ __launch_bounds__(256) void k_scan_sums(int n, const int *__restrict__ in, int *__restrict__ sums)
{
    __shared__ int sh[256];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    int v = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) v += base + q < n ? in[base + q] : 0;
    int tot;
    block_exclusive_scan(v, sh, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_top(int nb, int *__restrict__ sums)
{
    // exclusive scan of nb <= 256 * 16 block sums in place
    __shared__ int sh[256];
    const int per = (nb + 255) / 256, lo = threadIdx.x * per, hi = min(nb, lo + per);
    int v = 0;
    for (int i = lo; i < hi; i++) v += sums[i];
    int tot;
    int run = block_exclusive_scan(v, sh, tot);
    for (int i = lo; i < hi; i++) {
        int w = sums[i];
        sums[i] = run;
        run += w;
    }
    if (threadIdx.x == 0) sums[nb] = tot;
}

__global__ __launch_bounds__(256) void k_scan_apply(int n, const int *__restrict__ in,
                                                    const int *__restrict__ sums, int nb,
                                                    int *__restrict__ out)
{
    __shared__ int sh[256];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    int v[4], s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        v[q] = base + q < n ? in[base + q] : 0;
        s += v[q];
    }
    int tot;
    int run = block_exclusive_scan(s, sh, tot) + sums[blockIdx.x];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = sums[nb];
}

// Plain fill (used instead of hipMemsetAsync so the binning can be captured
// into a graph).
__global__ void k_fill_int(int n, int *__restrict__ a, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Target tiles from the counts of a brick's quarters (qi = quarter within the
// brick, c its count, h the count of its half brick, B of the brick): a brick
// with <= 64 targets is one tile (4x4x4 cells), else a half brick with <= 64
// (2x4x4), else each quarter (2x2x4) in runs of <= 64.  Sparse targets thus
// share one candidate walk per tile instead of a few lanes each.  Returns the
// tiles that quarter qi writes and their span (targets from the quarter's
// first one).
__device__ __forceinline__ int quarter_tiles(int qi, int c, int h, int B, int &span)
{
    if (B <= 64) {
        span = B;
        return qi == 0 && B > 0;
    }
    if (h <= 64) {
        span = h;
        return (qi & 1) == 0 && h > 0;
    }
    span = c;
    return (c + 63) >> 6;
}

__device__ __forceinline__ int quarter_tiles_at(const int *__restrict__ tstart, int q, int &first, int &span)
{
    const int b = q >> 2, hq = q >> 1;
    first = tstart[16 * q];
    const int c = tstart[16 * q + 16] - first;
    const int h = tstart[32 * hq + 32] - tstart[32 * hq];
    const int B = tstart[64 * b + 64] - tstart[64 * b];
    return quarter_tiles(q & 3, c, h, B, span);
}

__global__ void k_quarter_tiles(int nq, const int *__restrict__ tstart, int *__restrict__ qcount)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int first, span;
    qcount[q] = quarter_tiles_at(tstart, q, first, span);
}

__global__ void k_write_tiles(int nq, const int *__restrict__ tstart, const int *__restrict__ qoff,
                              int2 *__restrict__ tiles)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int first, span;
    const int nt = quarter_tiles_at(tstart, q, first, span);
    const int o = qoff[q];
    for (int i = 0; i < nt; i++) tiles[o + i] = make_int2(first + 64 * i, min(64, span - 64 * i));
}

// ---- fused binning of sources and targets (small key spaces) -----------------
// Five launches instead of seventeen: one fill (which can also zero the
// output), one count and one scatter over both point sets, and the two
// launches of the chunked scans (the target one also writes the tiles).  Used
// when ncell + nkey_t <= kFusedKeys.
constexpr int kFusedKeys = 1 << 18;

// Zeroes both count arrays and, with u, the nu doubles of the output (one
// launch instead of two).
__global__ void k_fill_int2(int n0, int *__restrict__ a0, int n1, int *__restrict__ a1, long long nu,
                            double *__restrict__ u)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n0)
        a0[i] = 0;
    else if (i - n0 < n1)
        a1[i - n0] = 0;
    else if (i - n0 - n1 < nu)
        u[i - n0 - n1] = 0.0;
}

__device__ __forceinline__ int cell_key(const double *__restrict__ pos, size_t i, const pb_box &box,
                                        const pb_cells &cells, bool brick_order)
{
    int ci[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double w = wrap_coord(pos[3 * i + c], box.L[c]);
        const int k = (int)(w * cells.inv_h[c]);
        ci[c] = k >= cells.nc[c] ? cells.nc[c] - 1 : k;
    }
    if (brick_order) return pb_target_key(ci[0], ci[1], ci[2], cells);
    return (ci[0] * cells.nc[1] + ci[1]) * cells.nc[2] + ci[2];
}

// threads [0, ns): sources (z-major cell keys); [ns, ns + nt): targets
// (brick-Morton keys)
__device__ __forceinline__ void bin_count2(int i, int ns, const double *__restrict__ xs, int nt,
                                           const double *__restrict__ yt, const pb_box &box, const pb_cells &cells,
                                           int *__restrict__ scount, int *__restrict__ skey, int *__restrict__ srank,
                                           int *__restrict__ tcount, int *__restrict__ tkey, int *__restrict__ trank)
{
    if (i < ns) {
        const int key = cell_key(xs, (size_t)i, box, cells, false);
        skey[i] = key;
        srank[i] = atomicAdd(&scount[key], 1);
    } else if (i - ns < nt) {
        const int j = i - ns;
        const int key = cell_key(yt, (size_t)j, box, cells, true);
        tkey[j] = key;
        trank[j] = atomicAdd(&tcount[key], 1);
    }
}

__global__ void k_bin_count2(int ns, const double *__restrict__ xs, int nt, const double *__restrict__ yt,
                             pb_box box, pb_cells cells, int *__restrict__ scount, int *__restrict__ skey,
                             int *__restrict__ srank, int *__restrict__ tcount, int *__restrict__ tkey,
                             int *__restrict__ trank)
{
    bin_count2(blockIdx.x * blockDim.x + threadIdx.x, ns, xs, nt, yt, box, cells, scount, skey, srank, tcount, tkey,
               trank);
}

__device__ __forceinline__ void bin_scatter2(int i, int ns, const double *__restrict__ xs,
                                             const double *__restrict__ fs, int nt, const double *__restrict__ yt,
                                             const pb_box &box, const int *__restrict__ sstart,
                                             const int *__restrict__ skey, const int *__restrict__ srank,
                                             double4 *__restrict__ spos, double4 *__restrict__ sfrc,
                                             int *__restrict__ sperm, const int *__restrict__ tstart,
                                             const int *__restrict__ tkey, const int *__restrict__ trank,
                                             double4 *__restrict__ tpos, int *__restrict__ tperm)
{
    const bool src = i < ns;
    const int j = src ? i : i - ns;
    if (!src && j >= nt) return;
    const double *pos = src ? xs : yt;
    const int dst = (src ? sstart[skey[j]] + srank[j] : tstart[tkey[j]] + trank[j]);
    double w[3];
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = wrap_coord(pos[3 * (size_t)j + c], box.L[c]);
    if (src) {
        spos[dst] = make_double4(w[0], w[1], w[2], 0.0);
        if (fs) sfrc[dst] = make_double4(fs[3 * (size_t)j], fs[3 * (size_t)j + 1], fs[3 * (size_t)j + 2], 0.0);
        sperm[dst] = j;
    } else {
        tpos[dst] = make_double4(w[0], w[1], w[2], 0.0);
        tperm[dst] = j;
    }
}

__global__ void k_bin_scatter2(int ns, const double *__restrict__ xs, const double *__restrict__ fs, int nt,
                               const double *__restrict__ yt, pb_box box, const int *__restrict__ sstart,
                               const int *__restrict__ skey, const int *__restrict__ srank,
                               double4 *__restrict__ spos, double4 *__restrict__ sfrc, int *__restrict__ sperm,
                               const int *__restrict__ tstart, const int *__restrict__ tkey,
                               const int *__restrict__ trank, double4 *__restrict__ tpos, int *__restrict__ tperm)
{
    bin_scatter2(blockIdx.x * blockDim.x + threadIdx.x, ns, xs, fs, nt, yt, box, sstart, skey, srank, spos, sfrc,
                 sperm, tstart, tkey, trank, tpos, tperm);
}

// Exclusive scan across the 1024 threads of the workgroup; total to every thread.
__device__ __forceinline__ int scan1024(int v, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int sk = wsum[k];
        pre += k < w ? sk : 0;
        tot += sk;
    }
    __syncthreads(); // wsum is rewritten by the next call
    total = tot;
    return pre + x - v;
}

// Exclusive scan of n counts into start[0..n] (start[n] = total), 16
// consecutive counts per thread and 16384 per round.  With tiles != null the
// counts are target keys (n a multiple of 64): each thread's 16 keys are one
// quarter brick, whose targets become tiles by quarter_tiles.
__device__ void scan_counts(const int *__restrict__ count, int n, int *__restrict__ start, int *wsum,
                            int2 *__restrict__ tiles, int *__restrict__ ntiles)
{
    int carry = 0, tcarry = 0;
    auto load16 = [&](int i0, int (&v)[16]) {
        if (i0 + 16 <= n) {
            const int4 *c4 = reinterpret_cast<const int4 *>(count + i0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int4 t = c4[q];
                v[4 * q] = t.x;
                v[4 * q + 1] = t.y;
                v[4 * q + 2] = t.z;
                v[4 * q + 3] = t.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) v[q] = i0 + q < n ? count[i0 + q] : 0;
        }
    };
    for (int base = 0; base < n; base += 16 * 1024) {
        const int i0 = base + 16 * (int)threadIdx.x;
        int v[16];
        load16(i0, v);
        int sum = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) sum += v[q];
        int tot;
        const int first = scan1024(sum, wsum, tot) + carry;
        int run = first;
        if (i0 + 16 <= n) {
            int4 *o4 = reinterpret_cast<int4 *>(start + i0);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int4 t;
                t.x = run;
                run += v[4 * q];
                t.y = run;
                run += v[4 * q + 1];
                t.z = run;
                run += v[4 * q + 2];
                t.w = run;
                run += v[4 * q + 3];
                o4[q] = t;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                if (i0 + q < n) start[i0 + q] = run;
                run += v[q];
            }
        }
        carry += tot;
        if (tiles) { // the 4 quarters of a brick are 4 consecutive lanes
            const int h = sum + __shfl_xor(sum, 1, 64);
            const int B = h + __shfl_xor(h, 2, 64);
            int span;
            const int nq = quarter_tiles((int)threadIdx.x & 3, sum, h, B, span);
            int ttot;
            int o = scan1024(nq, wsum, ttot) + tcarry;
            for (int k = 0; k < nq; k++) tiles[o + k] = make_int2(first + 64 * k, min(64, span - 64 * k));
            tcarry += ttot;
        }
    }
    if (threadIdx.x == 0) {
        start[n] = carry;
        if (ntiles) *ntiles = tcarry;
    }
}



// ---- chunked scans of the fused binning ------------------------------------
// The two count arrays (sources: cell counts; targets: brick-Morton key
// counts, whose quarter bricks also become tiles) are scanned in two launches
// of one workgroup per chunk of kChunk keys: the chunk totals of counts and
// tiles, then each chunk's scan offset by the totals of the chunks before it
// in the same array.  Many workgroups in flight instead of one per array
// working through the array in rounds.
constexpr int kChunk = 4096; // 256 threads x 16 keys; a multiple of a brick's 64 keys

__device__ __forceinline__ void load16(const int *__restrict__ count, int n, int i0, int (&v)[16])
{
    if (i0 + 16 <= n) {
        const int4 *c4 = reinterpret_cast<const int4 *>(count + i0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int4 t = c4[q];
            v[4 * q] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = i0 + q < n ? count[i0 + q] : 0;
    }
}

// Exclusive scan across the 256 threads of the workgroup; total to every thread.
__device__ __forceinline__ int scan256(int v, int *wsum, int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sk = wsum[k];
        pre += k < w ? sk : 0;
        tot += sk;
    }
    __syncthreads(); // wsum is rewritten by the next call
    total = tot;
    return pre + x - v;
}

// Tiles of the thread's quarter brick (16 target keys summing to c; the 4
// quarters of a brick are 4 consecutive lanes).
__device__ __forceinline__ int thread_tiles(int c, int &span)
{
    const int h = c + __shfl_xor(c, 1, 64);
    const int B = h + __shfl_xor(h, 2, 64);
    return quarter_tiles((int)threadIdx.x & 3, c, h, B, span);
}

// blocks [0, nchs): source chunks; [nchs, nchs + ncht): target chunks.
__global__ __launch_bounds__(256) void k_bin_sums2(const int *__restrict__ scount, int ncell,
                                                   const int *__restrict__ tcount, int nkey_t, int nchs,
                                                   int2 *__restrict__ csum)
{
    __shared__ int wsum[4];
    const bool tgt = (int)blockIdx.x >= nchs; // uniform per workgroup
    const int ch = tgt ? (int)blockIdx.x - nchs : (int)blockIdx.x;
    int v[16];
    load16(tgt ? tcount : scount, tgt ? nkey_t : ncell, ch * kChunk + 16 * (int)threadIdx.x, v);
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) sum += v[q];
    int span, nq = 0;
    if (tgt) nq = thread_tiles(sum, span);
    int tot, ttot;
    scan256(sum, wsum, tot);
    scan256(nq, wsum, ttot);
    if (threadIdx.x == 0) csum[blockIdx.x] = make_int2(tot, ttot);
}

__global__ __launch_bounds__(256) void k_bin_apply2(const int *__restrict__ scount, int ncell, int *__restrict__ sstart,
                                                    const int *__restrict__ tcount, int nkey_t, int *__restrict__ tstart,
                                                    int nchs, int ncht, const int2 *__restrict__ csum,
                                                    int2 *__restrict__ tiles, int *__restrict__ ntiles)
{
    __shared__ int wsum[4];
    const bool tgt = (int)blockIdx.x >= nchs; // uniform per workgroup
    const int ch = tgt ? (int)blockIdx.x - nchs : (int)blockIdx.x;
    const int n = tgt ? nkey_t : ncell, nch = tgt ? ncht : nchs, cbase = tgt ? nchs : 0;
    int *__restrict__ start = tgt ? tstart : sstart;
    // totals of the chunks before this one
    int pc = 0, pt = 0;
    for (int k = (int)threadIdx.x; k < ch; k += 256) {
        const int2 t = csum[cbase + k];
        pc += t.x;
        pt += t.y;
    }
    int carry, tcarry;
    scan256(pc, wsum, carry);
    scan256(pt, wsum, tcarry);
    const int i0 = ch * kChunk + 16 * (int)threadIdx.x;
    int v[16];
    load16(tgt ? tcount : scount, n, i0, v);
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) sum += v[q];
    int tot;
    const int first = scan256(sum, wsum, tot) + carry;
    int run = first;
    if (i0 + 16 <= n) {
        int4 *o4 = reinterpret_cast<int4 *>(start + i0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            int4 t;
            t.x = run;
            run += v[4 * q];
            t.y = run;
            run += v[4 * q + 1];
            t.z = run;
            run += v[4 * q + 2];
            t.w = run;
            run += v[4 * q + 3];
            o4[q] = t;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (i0 + q < n) start[i0 + q] = run;
            run += v[q];
        }
    }
    const bool last = ch == nch - 1 && threadIdx.x == 0;
    if (last) start[n] = carry + tot;
    if (tgt) {
        int span;
        const int nq = thread_tiles(sum, span);
        int ttot;
        const int o = scan256(nq, wsum, ttot) + tcarry;
        for (int k = 0; k < nq; k++) tiles[o + k] = make_int2(first + 64 * k, min(64, span - 64 * k));
        if (last) *ntiles = tcarry + ttot;
    }
}

// ---- grouped binning: every system of a batched call in one launch each ----
__global__ __launch_bounds__(256) void k_gbin_count(const pb_sys *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = pb_map_at(map, blockIdx.x);
    const auto &S = ((pb_const_sys)sys)[m.x];
    const pb_box box = S.box;
    const pb_cells cells = S.cells;
    bin_count2(m.y * 256 + (int)threadIdx.x, S.ns, S.xn, S.nt, S.y, box, cells, S.scount, S.skey, S.srank,
               S.tcount, S.tkey, S.trank);
}

__global__ __launch_bounds__(256) void k_gbin_scatter(const pb_sys *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = pb_map_at(map, blockIdx.x);
    const auto &S = ((pb_const_sys)sys)[m.x];
    const pb_box box = S.box;
    bin_scatter2(m.y * 256 + (int)threadIdx.x, S.ns, S.xn, S.f, S.nt, S.y, box, S.sstart, S.skey, S.srank, S.spos,
                 S.f ? S.sfrc : nullptr, S.sperm, S.tstart, S.tkey, S.trank, S.tpos, S.tperm);
}

__global__ __launch_bounds__(1024) void k_gbin_scan(const pb_sys *__restrict__ sys, const int2 *__restrict__ map)
{
    __shared__ int wsum[16];
    const auto &S = ((pb_const_sys)sys)[pb_map_at(map, blockIdx.x).x];
    scan_counts(S.scount, S.cells.ncell, S.sstart, wsum, nullptr, nullptr);
    scan_counts(S.tcount, S.cells.nkey_t, S.tstart, wsum, S.tiles, S.ntiles);
}

__global__ void k_fill_map(int2 *__restrict__ map, long long n, int per, long long valid)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[i] = i < valid ? make_int2((int)(i / per), (int)(i % per)) : make_int2(-1, 0);
}

// Gather map for systems of per blocks each, read after the XCD renumbering
// of pb_xcd_block (n = 8 nsys per8 entries, per8 = ceil(per / 8)): at each
// step of the dispatch the 8 XCDs take the 8 eighths of one system, so the
// systems pass through the device one after another (one grid in the shared
// cache at a time) while each XCD's L2 sees a contiguous run of targets.
__global__ void k_fill_map_xcd(int2 *__restrict__ map, int nsys, int per)
{
    const int per8 = (per + 7) / 8;
    const long long n = 8LL * nsys * per8, L = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= n) return;
    const long long steps = (long long)nsys * per8;
    const int x = (int)(L / steps);
    const long long k = L % steps;
    const int sys = (int)(k / per8), b = x * per8 + (int)(k % per8);
    map[L] = b < per ? make_int2(sys, b) : make_int2(-1, 0);
}

__global__ void k_fill_zero(long long n, int *__restrict__ a)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}
} // namespace

// Exclusive scan of n ints from in to out (out[n] = total) on stream s.
static int run_scan(pb_context *c, hipStream_t s, int n, const int *in, int *out)
{
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb > 256 * 16) return pb_fail(c, PB_ERR_UNSUPPORTED, "too many cells for the scan");
    int st;
    if ((st = pb_reserve(c, &c->scan_tmp, sizeof(int) * (nb + 1)))) return st;
    int *sums = (int *)c->scan_tmp.ptr;
    hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(256), 0, s, n, in, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, nb, sums);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, n, in, (const int *)sums, nb, out);
    return PB_OK;
}

int pb_build_tiles(pb_context *c, hipStream_t s, const pb_cells &cells, int nt, int *max_tiles)
{
    // a quarter of a brick is 2x2x4 cells: 16 consecutive target keys
    const int nq = cells.nkey_t / 16;
    const size_t mt = (size_t)nq + (size_t)nt / 64 + 1;
    int st;
    if ((st = pb_reserve(c, &c->qcount, sizeof(int) * (nq + 1)))) return st;
    if ((st = pb_reserve(c, &c->qoff, sizeof(int) * (nq + 1)))) return st;
    if ((st = pb_reserve(c, &c->tiles, sizeof(int2) * mt))) return st;
    const int blocks = (nq + 255) / 256;
    hipLaunchKernelGGL(k_quarter_tiles, dim3(blocks), dim3(256), 0, s, nq, (const int *)c->tcell_start.ptr,
                       (int *)c->qcount.ptr);
    if ((st = run_scan(c, s, nq, (const int *)c->qcount.ptr, (int *)c->qoff.ptr))) return st;
    hipLaunchKernelGGL(k_write_tiles, dim3(blocks), dim3(256), 0, s, nq, (const int *)c->tcell_start.ptr,
                       (const int *)c->qoff.ptr, (int2 *)c->tiles.ptr);
    *max_tiles = (int)mt;
    return pb_hip_check(c, hipGetLastError(), "tile launch");
}

int pb_bin_points(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_box &box, int n,
                  const double *pos, const double *frc, bool is_source)
{
    const int nkey = is_source ? cells.ncell : cells.nkey_t;
    pb_buffer &count = is_source ? c->cell_count : c->tcell_count;
    pb_buffer &start = is_source ? c->cell_start : c->tcell_start;
    pb_buffer &key_of = is_source ? c->cell_of_s : c->cell_of_t;
    pb_buffer &rank = is_source ? c->rank_s : c->rank_t;
    pb_buffer &opos = is_source ? c->spos : c->tpos;
    int st;
    size_t nn = (size_t)(n > 0 ? n : 1);
    if ((st = pb_reserve(c, &count, sizeof(int) * (nkey + 1)))) return st;
    if ((st = pb_reserve(c, &start, sizeof(int) * (nkey + 1)))) return st;
    if ((st = pb_reserve(c, &key_of, sizeof(int) * nn))) return st;
    if ((st = pb_reserve(c, &rank, sizeof(int) * nn))) return st;
    if ((st = pb_reserve(c, &opos, sizeof(double4) * nn))) return st;
    if (is_source) {
        if ((st = pb_reserve(c, &c->sfrc, sizeof(double4) * nn))) return st;
        if ((st = pb_reserve(c, &c->sperm, sizeof(int) * nn))) return st;
    } else {
        if ((st = pb_reserve(c, &c->tperm, sizeof(int) * nn))) return st;
    }
    hipLaunchKernelGGL(k_fill_int, dim3((nkey + 1 + 255) / 256), dim3(256), 0, s, nkey + 1, (int *)count.ptr, 0);
    if (n > 0) {
        int blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_bin_count, dim3(blocks), dim3(256), 0, s, n, pos, box, cells,
                           is_source ? 0 : 1, (int *)count.ptr, (int *)key_of.ptr, (int *)rank.ptr);
    }
    if ((st = run_scan(c, s, nkey, (const int *)count.ptr, (int *)start.ptr))) return st;
    if (n > 0) {
        int blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_bin_scatter, dim3(blocks), dim3(256), 0, s, n, pos, frc, box,
                           (const int *)start.ptr, (const int *)key_of.ptr,
                           (const int *)rank.ptr, (double4 *)opos.ptr,
                           (is_source && frc) ? (double4 *)c->sfrc.ptr : nullptr,
                           is_source ? (int *)c->sperm.ptr : (int *)c->tperm.ptr);
    }
    if (!is_source && (st = pb_build_tiles(c, s, cells, n, &c->max_tiles))) return st;
    return pb_hip_check(c, hipGetLastError(), "binning launch");
}

int pb_bin_all(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_box &box, int ns, const double *xn,
               const double *f, int nt, const double *y, double *zero_u)
{
    if (!pb_bin_fused_ok(cells)) {
        int st;
        if (zero_u) pb_zero_u(s, nt, zero_u);
        if ((st = pb_bin_points(c, s, cells, box, ns, xn, f, true))) return st;
        return pb_bin_points(c, s, cells, box, nt, y, nullptr, false);
    }
    int st;
    const size_t nsn = (size_t)(ns > 0 ? ns : 1), ntn = (size_t)(nt > 0 ? nt : 1);
    const int ncell = cells.ncell, nkt = cells.nkey_t, nq = nkt / 16;
    const size_t mt = (size_t)std::min(nq, nt) + (size_t)nt / 64 + 1; // non-empty quarters <= nt
    if ((st = pb_reserve(c, &c->cell_count, sizeof(int) * (ncell + 1)))) return st;
    if ((st = pb_reserve(c, &c->cell_start, sizeof(int) * (ncell + 1)))) return st;
    if ((st = pb_reserve(c, &c->cell_of_s, sizeof(int) * nsn))) return st;
    if ((st = pb_reserve(c, &c->rank_s, sizeof(int) * nsn))) return st;
    if ((st = pb_reserve(c, &c->spos, sizeof(double4) * nsn))) return st;
    if ((st = pb_reserve(c, &c->sfrc, sizeof(double4) * nsn))) return st;
    if ((st = pb_reserve(c, &c->sperm, sizeof(int) * nsn))) return st;
    if ((st = pb_reserve(c, &c->tcell_count, sizeof(int) * (nkt + 1)))) return st;
    if ((st = pb_reserve(c, &c->tcell_start, sizeof(int) * (nkt + 1)))) return st;
    if ((st = pb_reserve(c, &c->cell_of_t, sizeof(int) * ntn))) return st;
    if ((st = pb_reserve(c, &c->rank_t, sizeof(int) * ntn))) return st;
    if ((st = pb_reserve(c, &c->tpos, sizeof(double4) * ntn))) return st;
    if ((st = pb_reserve(c, &c->tperm, sizeof(int) * ntn))) return st;
    if ((st = pb_reserve(c, &c->qoff, sizeof(int) * (nq + 1)))) return st;
    if ((st = pb_reserve(c, &c->tiles, sizeof(int2) * mt))) return st;
    const int nchs = (ncell + kChunk - 1) / kChunk, ncht = (nkt + kChunk - 1) / kChunk;
    if ((st = pb_reserve(c, &c->scan_tmp, sizeof(int2) * (size_t)(nchs + ncht)))) return st;
    int *scount = (int *)c->cell_count.ptr, *tcount = (int *)c->tcell_count.ptr;
    const long long nu = zero_u ? 3LL * std::max(nt, 0) : 0;
    const long long nfill = (long long)ncell + nkt + nu;
    hipLaunchKernelGGL(k_fill_int2, dim3((unsigned)((nfill + 255) / 256)), dim3(256), 0, s, ncell, scount, nkt, tcount,
                       nu, zero_u);
    const long long npts = (long long)ns + nt;
    if (npts > 0)
        hipLaunchKernelGGL(k_bin_count2, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, ns, xn, nt, y, box,
                           cells, scount, (int *)c->cell_of_s.ptr, (int *)c->rank_s.ptr, tcount,
                           (int *)c->cell_of_t.ptr, (int *)c->rank_t.ptr);
    int2 *csum = (int2 *)c->scan_tmp.ptr;
    hipLaunchKernelGGL(k_bin_sums2, dim3((unsigned)(nchs + ncht)), dim3(256), 0, s, (const int *)scount, ncell,
                       (const int *)tcount, nkt, nchs, csum);
    hipLaunchKernelGGL(k_bin_apply2, dim3((unsigned)(nchs + ncht)), dim3(256), 0, s, (const int *)scount, ncell,
                       (int *)c->cell_start.ptr, (const int *)tcount, nkt, (int *)c->tcell_start.ptr, nchs, ncht,
                       (const int2 *)csum, (int2 *)c->tiles.ptr, (int *)c->qoff.ptr + nq);
    if (npts > 0)
        hipLaunchKernelGGL(k_bin_scatter2, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, ns, xn, f, nt, y,
                           box, (const int *)c->cell_start.ptr, (const int *)c->cell_of_s.ptr,
                           (const int *)c->rank_s.ptr, (double4 *)c->spos.ptr, f ? (double4 *)c->sfrc.ptr : nullptr,
                           (int *)c->sperm.ptr, (const int *)c->tcell_start.ptr, (const int *)c->cell_of_t.ptr,
                           (const int *)c->rank_t.ptr, (double4 *)c->tpos.ptr, (int *)c->tperm.ptr);
    c->max_tiles = (int)mt;
    return pb_hip_check(c, hipGetLastError(), "binning launch");
}

void pb_group_bin_count(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk)
{
    if (nblk > 0) hipLaunchKernelGGL(k_gbin_count, dim3(nblk), dim3(256), 0, s, sys, map);
}

void pb_group_bin_scan(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk)
{
    if (nblk > 0) hipLaunchKernelGGL(k_gbin_scan, dim3(nblk), dim3(1024), 0, s, sys, map);
}

void pb_group_bin_scatter(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk)
{
    if (nblk > 0) hipLaunchKernelGGL(k_gbin_scatter, dim3(nblk), dim3(256), 0, s, sys, map);
}

void pb_fill_zero(hipStream_t s, long long n, int *a)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a);
}

bool pb_bin_fused_ok(const pb_cells &cells)
{
    return (long long)cells.ncell + cells.nkey_t <= kFusedKeys && cells.nkey_t % 64 == 0;
}

void pb_fill_map(hipStream_t s, int2 *map, long long n, int per, long long valid)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, map, n, per, valid);
}

long long pb_fill_map_xcd(hipStream_t s, int2 *map, int nsys, int per)
{
    const long long n = 8LL * nsys * ((per + 7) / 8);
    if (n > 0) hipLaunchKernelGGL(k_fill_map_xcd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, map, nsys, per);
    return n;
}
