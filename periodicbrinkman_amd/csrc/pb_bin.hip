// pb_bin.hip -- cell binning of sources and targets (counting sort) and the
// target tiles of the real-space pass.
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
                            int *__restrict__ count, int *__restrict__ cell_of,
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
    int cell = (ci[0] * cells.nc[1] + ci[1]) * cells.nc[2] + ci[2];
    cell_of[i] = cell;
    rank[i] = atomicAdd(&count[cell], 1);
}

__global__ void k_bin_scatter(int n, const double *__restrict__ pos, const double *__restrict__ frc,
                              pb_box box, const int *__restrict__ start,
                              const int *__restrict__ cell_of, const int *__restrict__ rank,
                              double4 *__restrict__ opos, double4 *__restrict__ ofrc,
                              int *__restrict__ operm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dst = start[cell_of[i]] + rank[i];
    double w[3];
#pragma unroll
    for (int c = 0; c < 3; c++) w[c] = wrap_coord(pos[3 * (size_t)i + c], box.L[c]);
    opos[dst] = make_double4(w[0], w[1], w[2], 0.0);
    if (ofrc) ofrc[dst] = make_double4(frc[3 * (size_t)i], frc[3 * (size_t)i + 1], frc[3 * (size_t)i + 2], 0.0);
    if (operm) operm[dst] = i;
}

// Exclusive scan of n ints in one workgroup of 1024 threads; out[n] = total.
__global__ __launch_bounds__(1024) void k_scan(int n, const int *__restrict__ in,
                                                int *__restrict__ out)
{
    __shared__ int part[1024];
    int t = threadIdx.x;
    int chunk = (n + 1023) / 1024;
    int lo = t * chunk, hi = min(n, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = lo; i < hi; i++) {
        int v = in[i];
        out[i] = run;
        run += v;
    }
    if (t == 1023) out[n] = part[1023];
}

__global__ void k_tile_count(int ncell, const int *__restrict__ count, int *__restrict__ tcount)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) tcount[c] = (count[c] + 63) >> 6;
}

__global__ void k_make_tiles(int ncell, const int *__restrict__ count,
                             const int *__restrict__ start, const int *__restrict__ toff,
                             pb_tile *__restrict__ tiles)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int n = count[c], first = start[c], o = toff[c];
    for (int k = 0; k * 64 < n; k++) {
        pb_tile t;
        t.cell = c;
        t.first = first + 64 * k;
        t.count = min(64, n - 64 * k);
        t.pad = 0;
        tiles[o + k] = t;
    }
}

} // namespace

int pb_bin_points(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_box &box, int n,
                  const double *pos, const double *frc, bool is_source)
{
    int nc = cells.ncell;
    pb_buffer &count = is_source ? c->cell_count : c->tcell_count;
    pb_buffer &start = is_source ? c->cell_start : c->tcell_start;
    pb_buffer &cell_of = is_source ? c->cell_of_s : c->cell_of_t;
    pb_buffer &rank = is_source ? c->rank_s : c->rank_t;
    pb_buffer &opos = is_source ? c->spos : c->tpos;
    int st;
    size_t nn = (size_t)(n > 0 ? n : 1);
    if ((st = pb_reserve(c, &count, sizeof(int) * (nc + 1)))) return st;
    if ((st = pb_reserve(c, &start, sizeof(int) * (nc + 1)))) return st;
    if ((st = pb_reserve(c, &cell_of, sizeof(int) * nn))) return st;
    if ((st = pb_reserve(c, &rank, sizeof(int) * nn))) return st;
    if ((st = pb_reserve(c, &opos, sizeof(double4) * nn))) return st;
    if (is_source) {
        if ((st = pb_reserve(c, &c->sfrc, sizeof(double4) * nn))) return st;
    } else {
        if ((st = pb_reserve(c, &c->tperm, sizeof(int) * nn))) return st;
    }
    if ((st = pb_hip_check(c, hipMemsetAsync(count.ptr, 0, sizeof(int) * (nc + 1), s), "memset")))
        return st;
    if (n > 0) {
        int blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_bin_count, dim3(blocks), dim3(256), 0, s, n, pos, box, cells,
                           (int *)count.ptr, (int *)cell_of.ptr, (int *)rank.ptr);
    }
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, nc, (const int *)count.ptr,
                       (int *)start.ptr);
    if (n > 0) {
        int blocks = (n + 255) / 256;
        hipLaunchKernelGGL(k_bin_scatter, dim3(blocks), dim3(256), 0, s, n, pos, frc, box,
                           (const int *)start.ptr, (const int *)cell_of.ptr,
                           (const int *)rank.ptr, (double4 *)opos.ptr,
                           is_source ? (double4 *)c->sfrc.ptr : nullptr,
                           is_source ? nullptr : (int *)c->tperm.ptr);
    }
    if (!is_source) {
        // tiles of <= 64 targets per cell
        if ((st = pb_reserve(c, &c->tile_off, sizeof(int) * (nc + 1)))) return st;
        size_t max_tiles = (size_t)nc + nn / 64 + 1;
        if ((st = pb_reserve(c, &c->tiles, sizeof(pb_tile) * max_tiles))) return st;
        if ((st = pb_reserve(c, &c->ntiles, sizeof(int) * (nc + 1)))) return st;
        hipLaunchKernelGGL(k_tile_count, dim3((nc + 255) / 256), dim3(256), 0, s, nc,
                           (const int *)count.ptr, (int *)c->ntiles.ptr);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, nc, (const int *)c->ntiles.ptr,
                           (int *)c->tile_off.ptr);
        hipLaunchKernelGGL(k_make_tiles, dim3((nc + 255) / 256), dim3(256), 0, s, nc,
                           (const int *)count.ptr, (const int *)start.ptr,
                           (const int *)c->tile_off.ptr, (pb_tile *)c->tiles.ptr);
    }
    return pb_hip_check(c, hipGetLastError(), "binning launch");
}
