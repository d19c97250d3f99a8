// pb_internal.h -- declarations shared by the HIP translation units of libpbamd.
//
// Device data layout (HBM):
//   sorted sources   spos[j] = (x, y, z, 0) wrapped into [0, L), double4
//                    sfrc[j] = (fx, fy, fz, 0), double4
//                    cell_start[c] .. cell_start[c+1]: sources of cell
//                    c = (c0 * nc1 + c1) * nc2 + c2 (z-major)
//   sorted targets   tpos[t] = (x, y, z, 0), tperm[t] = original index, in
//                    brick-Morton order; tiles[k] = (first, count <= 64)
//                    within one quarter brick (2x2x4 cells), or a whole half
//                    brick or brick where that holds <= 64 targets
//   k-space grids    real grid: (fx, fy, fz, 0) then (ux, uy, uz, 0) per point,
//                    point (i0, i1, i2) at (i0*n1 + i1)*n2 + i2 (T4);
//                    half spectrum: 3 fields of n0 x n1 x (n2/2 + 1) complex,
//                    mode (k0, k1, k2) at (k0*n1 + k1)*(n2/2 + 1) + k2
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/periodicbrinkman.h"

#define PB_NSER 36  // terms of the (r/d)^2 series of the regularized functions
#define PB_NSING 14 // terms of the (a r) series of the singular function

// Timed phases of pb_velocity: binning, spread, forward FFT, multiply,
// inverse FFT, gather, pair kernel.
#define PB_NPHASE 7
#define PB_TSLOTS 64

// Radial-function constants for one split width (the Ewald split d or the blob e).
struct pb_radial {
    double a;      // Brinkman parameter
    double d;      // split width
    double inv_d;  // 1/d
    double x;      // a d / 2
    double ex2;    // e^{-x^2}
    double k4pa2;  // 1/(4 pi a^2)
    double kg;     // (1 - e^{-x^2}) / (2 a^2 d^3 pi^{3/2})
    double zs;     // r/d below which the series is used
    double c1[PB_NSER], c2[PB_NSER]; // regH = sum c1[m] z^{2m}, sum c2[m] z^{2m}, z = r/d
};

// FP32 constants for the far field (r >= 3d), split d only.
struct pb_radial32 {
    float a, inv_d, x, ex2, k4pa2, kg, d2;
};

// Everything the real-space kernels need about the kernel K_R.
struct pb_kparams {
    pb_radial pd;   // split d
    pb_radial pe;   // blob e (valid when has_e)
    pb_radial32 pf; // FP32 far field
    int has_e;
    double e8;      // 8 e: below it the blob term is evaluated
    double d3;      // 3 d
    double rfar2;   // (mixed precision) squared radius beyond which FP32 is used
    double rcommon; // r >= rcommon: only the closed form of diffH(d) is needed
    // piecewise-polynomial table of (H1, H2) of diffH(d) on z = r/d in
    // [tz0, tz0 + tn/tinv): interval i, t = 2((z - tz0) tinv - i) - 1 in [-1, 1],
    // H = sum_k c[i][k][0|1] t^k; coefficients at tab64 / tab32 (device)
    double tz0, tinv;
    int tn, tdeg; // intervals, polynomial degree (tn = 0: no table)
    const double *tab64;
    const float *tab32;
};

#define PB_TAB_MAXDEG 15
#define PB_TAB_DEG 12      // degree of the table polynomials (FP64 form, error <= 4e-15)
#define PB_TAB_DEG_MIXED 6 // degree for the mixed form (error <= PB_TAB_TOL_MIXED): fewer bytes and FMAs per pair
// table error of the mixed form, relative to each interval's largest |H|: two
// orders below the form's 1e-6 target (1e-10 before: 150 of the 512 batched
// systems then needed degree 8; the clustered step 46.9 -> 45.4 ms)
#define PB_TAB_TOL_MIXED 1e-8
#define PB_TAB_DEG_MID 8   // the mixed form's degree where 6 does not fit the LDS budget (large alpha d)
#define PB_TAB_LDS_DOUBLES 2048 // LDS budget of the FP64 table (16 KB)

struct pb_box {
    double L[3];
    double inv_L[3];
};

// Cell grid shared by the real-space pass and spreading: fine cells of side
// >= rc/3.  Sources are sorted z-major (key = (c0*nc1 + c1)*nc2 + c2), targets
// by brick-Morton key (brick of 4x4x4 cells, Morton order inside).
struct pb_cells {
    int nc[3];      // cells per axis
    int ncell;      // nc0 * nc1 * nc2 (source keys)
    int nb[3];      // bricks per axis, ceil(nc/4)
    int nkey_t;     // nb0 * nb1 * nb2 * 64 (target keys)
    int st[3];      // cells per rc, rounded up (stencil half-width)
    double h[3];    // cell size
    double inv_h[3];
    double rc, rc2; // cutoff
};

// k-space geometry.
struct pb_kgrid {
    int n[3];       // grid
    int P[3];       // window points per axis
    long long ng;   // n0 n1 n2
    double h[3], inv_h[3];
    double eta;     // Gaussian e^{-x^2/eta}
    double scale;   // (pi eta)^{-3/2}
    double vol;     // h0 h1 h2 (gather quadrature weight)
};

// Constants of the Fourier multiply of one system.
struct pb_kmult {
    double k[3];     // 2 pi / L per axis
    double c;        // (1 - e^{-a^2 d^2/4}) / a^2
    double a2;       // a^2
    double inv_ng;   // 1 / (n0 n1 n2)
    double beta;     // d^2/4 - eta/2: e^{-beta k^2} removes the split and the two window convolutions
    int n[3];
};
pb_kmult pb_make_kmult(const pb_kgrid &g, const pb_box &box, double a, double d);

// One system of a grouped batched call (device copy): its share of the
// caller's arrays, its geometry and constants, and its share of the group's
// scratch.  A grouped launch covers every system of a class (same kernel
// template) in one grid; block b of the launch works on system map[b].x,
// block map[b].y of that system's own partition.
struct pb_sys {
    const double *xn, *f, *y;
    double *u;
    int ns, nt;
    pb_box box;
    pb_cells cells;
    pb_kgrid g;
    pb_kmult km;
    const pb_kparams *kp;
    int *scount, *sstart, *skey, *srank, *sperm; // source binning (sstart = cell_start)
    double4 *spos, *sfrc;
    int *tcount, *tstart, *tkey, *trank, *tperm; // target binning
    double4 *tpos;
    int2 *tiles;
    int *ntiles;
    void *grid4, *spec; // real grid (T4 / packed trip) and half spectrum
    uint4 *stats;       // pair-kernel statistics, one per wavefront item
    int acc_mode;       // the gather: 0 overwrites u, 2 adds with atomics
};

// ---- host helpers (pb_params.cpp) ----------------------------------------
// Build the table of diffH(d) on z in [z0, z1] with polynomials of degree deg
// whose verified error (relative to each interval's largest |H|) is <= tol:
// fills tz0, tinv, tn, tdeg and the coefficients (layout [tn][deg+1][2]);
// returns that error (tn = 0: no table within the LDS budget).
double pb_build_table(pb_kparams *k, double z0, double z1, std::vector<double> &coef, int deg = PB_TAB_DEG,
                      double tol = 4e-15);
// Window half-width (in standard deviations) of P points when none is given.
double pb_window_m(int P);
void pb_fill_radial(pb_radial *r, double a, double d);
void pb_fill_kparams(pb_kparams *k, double a, double d, double e, double rfar);
void pb_fill_box(pb_box *b, const double *L);

// ---- context ----------------------------------------------------------------
struct pb_buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// The allocation-free entry point's device scratch (pb_noalloc.hip): one
// array of PB_NOALLOC_BYTES per device.  Its first PB_NOALLOC_FIXED bytes hold
// the constants (radial table, kernel constants, FFT twiddles) at fixed
// offsets; each call carves its other buffers from the rest.
#define PB_NOALLOC_BYTES PB_NOALLOC_MAX_BYTES
#define PB_NOALLOC_FIXED (64u << 10)
// Look up the array's device address for a device (with that device current;
// once per device, at context creation) / the address found (null: none).
int pb_noalloc_scratch_init(int device);
void *pb_noalloc_scratch(int device);

struct pb_context {
    int device = 0;
    int cu_count = 256; // compute units of the device
    std::string err;
    // arena mode (the allocation-free context of a device): pb_reserve hands
    // out slices of [arena, arena + arena_end) from arena_used on and never
    // allocates or frees
    char *arena = nullptr;
    size_t arena_end = 0, arena_used = 0;
    // real space
    pb_buffer cell_count, cell_start, cell_of_s, rank_s, cell_of_t, rank_t, tcell_count,
        tcell_start;
    pb_buffer spos, sfrc, tpos, tperm, counters, scan_tmp;
    pb_buffer sperm;  // sorted source -> caller's index
    pb_buffer mrhs_f; // forces of the sorted sources for many right-hand sides
    pb_buffer mrhs_u; // the real-space part of their velocities (R x 3 nt doubles)
    pb_buffer qcount, qoff, tiles; // target tiles: int2 (first, count); count at qoff[nkey_t/16]
    int max_tiles = 0;             // bound on the tile count (launch size)
    // k space
    pb_buffer grid, spec; // real grid (T4 per point), half spectrum
    // balanced spreading: candidate counts, split-region list, slot bases and
    // totals (ints); the work items (int2); the partial sums of split regions
    pb_buffer sp_int, sp_items, sp_part;
    pb_buffer twid[2];              // FFT twiddles for n = 4 .. 512: [0] FP64, [1] FP32
    int twid_ok[2] = {0, 0};
    // batched systems: kernel constants and radial tables of every system,
    // with the parameters they were built for
    pb_buffer batch_kp, batch_tab;
    std::vector<double> batch_key;
    std::vector<size_t> batch_ntab;
    std::vector<int> batch_tdeg;
    // batched systems run on PB_BATCH_LANES streams of their own, each with
    // its own scratch (a sub-context), so small systems overlap on the device
    pb_context *sub[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t sub_stream[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t sub_event[7] = {};
    hipStream_t cap_stream = nullptr; // origin of the batched call's fan-out and graph
    // a single system's pair kernel runs on this stream, concurrently with its
    // k-space pipeline on the caller's stream (fork and join by events)
    hipStream_t pair_stream = nullptr;
    hipEvent_t pair_event[2] = {};
    // a grouped call's second pair-kernel class runs on a second stream
    hipStream_t pair_stream2 = nullptr;
    hipEvent_t pair_event2 = {};
    // grouped batched systems: scratch of every system (ints, double4 points,
    // tiles, grids, spectra), device descriptors and block maps, and the
    // launch list built for the group layout in group_key
    pb_buffer gb_int, gb_d4, gb_tiles, gb_grid, gb_spec, gb_sys, gb_map, gb_stats;
    std::vector<double> group_key;
    struct group_launch {
        int kind;      // 0 count, 1 scan, 2 scatter, 3 pair, 4 spread, 5 FFT, 6 gather
        int a, b, c;   // pair: far32, table degree, nsplit; spread: precision; FFT: precision, pass, n;
                       // gather: precision, P
        size_t lds;    // pair: dynamic LDS bytes
        size_t map;    // offset of the launch's map entries
        int nblk;
    };
    std::vector<group_launch> group_launches;
    size_t group_counts = 0;    // ints to zero at the start of gb_int (the count arrays)
    // many right-hand sides: k space of up to 64 force sets per grouped launch
    pb_buffer mr_sys, mr_map, mr_frc, mr_grid, mr_spec;
    std::vector<pb_sys> mr_sys_host; // the descriptors last uploaded to mr_sys
    long long group_items = 0;  // pair statistics items
    // the batched call's captured graph and the calls it replays
    hipGraphExec_t graph_exec = nullptr;
    std::vector<double> graph_key, last_batch_key;
    // dense / shells
    pb_buffer dense_tmp;
    // kernel constants in device memory (kernels take a pointer: a by-value
    // copy of the series tables in kernel arguments costs registers)
    pb_buffer kparams, table;
    double kp_key[7] = {-1, -1, -1, -1, -1, -1, -1};
    size_t table_doubles = 0;
    int table_deg = PB_TAB_DEG; // degree of the table built with kparams
    // per-work-item statistics of the last real-space pass (uint4 each in
    // counters): FP64 pairs, FP32 pairs, candidates tested, candidates read
    long long stat_items = 0;
    bool stat_group = false; // the last pass was a grouped batched call (statistics in gb_stats)
    // phase timing (pb_set_timing): a ring of event sets, one per call, so a
    // timed loop of up to PB_TSLOTS calls needs no host synchronization
    int timing = 0;
    int ev_ready = 0;
    int cur = 0;
    long long ncalls = 0, nsummed = 0;
    hipEvent_t ev[PB_TSLOTS][PB_NPHASE + 1] = {};
    int slot_used[PB_TSLOTS] = {};
    unsigned recorded[PB_TSLOTS] = {}; // boundaries recorded in the slot's call (bit i)
    double tsum[PB_NPHASE] = {};
};

// Record the boundary event of phase i (0 = start) when timing is on.
void pb_mark(pb_context *c, hipStream_t s, int i);
// Record the end boundary (PB_NPHASE) unless the pair kernel already did.
void pb_mark_end(pb_context *c, hipStream_t s);
// Zero 3 n doubles of u on stream s.
void pb_zero_u(hipStream_t s, int n, double *u);

int pb_reserve(pb_context *c, pb_buffer *b, size_t bytes);
int pb_fail(pb_context *c, int status, const std::string &msg);
int pb_hip_check(pb_context *c, hipError_t e, const char *what);

// ---- launchers (pb_bin.hip, pb_real.hip, pb_kspace.hip, pb_dense.hip) -------
int pb_bin_points(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_box &box, int n,
                  const double *pos, const double *frc, bool is_source);
// Many right-hand sides (pb_mrhs.hip): Fs[j][b][rp] from the caller's R force
// sets (float, set r at r * 3 ns) padded to RP columns; the FP64 source forces
// of a chunk of sets for the k-space pass; the real-space contraction added
// into u (double, set r at r * 3 nt).
int pb_sort_forces(pb_context *c, hipStream_t s, int ns, int R, int RP, const float *f);
// frc[q * ns + j] = force set r0 + q of sorted source j, q < nr
int pb_force_sets(pb_context *c, hipStream_t s, int ns, int r0, int nr, int RP, double4 *frc);
// The real-space contraction OVERWRITES u (set r at r * 3 nt) for every
// target and set r < R; pb_add adds n doubles of ur to u on stream s.
void pb_add(hipStream_t s, long long n, double *u, const double *ur);
int pb_run_mrhs_real(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_kparams *kp,
                     size_t table_doubles, int table_deg, const pb_box &box, int ns, int nt, int R, int RP,
                     double *u);

// Target tiles: runs of <= 64 sorted targets inside one quarter brick (or a
// sparse half brick or brick); the
// tile count is left on the device, *max_tiles bounds it.  Built by
// pb_bin_points for targets.
int pb_build_tiles(pb_context *c, hipStream_t s, const pb_cells &cells, int nt, int *max_tiles);
// Sources and targets together (with the tiles), in five launches when the
// key space is small, else the two pb_bin_points calls; with zero_u the 3 nt
// doubles of the output are zeroed too (by the first launch).
int pb_bin_all(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_box &box, int ns, const double *xn,
               const double *f, int nt, const double *y, double *zero_u = nullptr);

// Workgroups are dispatched to the 8 XCDs of the device round-robin; this
// renumbering gives each XCD one contiguous range of logical ids, so
// neighbouring work shares an L2 (nblk8: launched workgroups, a multiple of 8).
// It pays for the gather, whose grid reads have no other reuse; the pair and
// spread kernels measured faster without it (their spatially adjacent
// workgroups then contend for the same L2 channels).
// Sort key of a target cell (cx, cy, cz): bricks of 4 x 4 x 4 cells in z-major
// order, Morton order (x1 y1 z1 x0 y0 z0) of the low two bits inside a brick.
__device__ __forceinline__ int pb_target_key(int cx, int cy, int cz, const pb_cells &cells)
{
    const int brick = ((cx >> 2) * cells.nb[1] + (cy >> 2)) * cells.nb[2] + (cz >> 2);
    const int m = ((cx & 2) << 4) | ((cy & 2) << 3) | ((cz & 2) << 2) | ((cx & 1) << 2) | ((cy & 1) << 1) | (cz & 1);
    return brick * 64 + m;
}
__device__ __forceinline__ int pb_xcd_block(int b, int nblk8)
{
    return (b & 7) * (nblk8 >> 3) + (b >> 3);
}

// Grouped kernels read their descriptors and maps through the constant
// address space: the data pointers loaded from there are known to be global
// (scalar-base global loads and stores instead of flat ones), and the
// compiler may reload a descriptor field instead of keeping it live.
#if defined(__HIP_DEVICE_COMPILE__)
#define PB_CONSTANT_AS __attribute__((address_space(4)))
#else
#define PB_CONSTANT_AS // (the host pass only parses the kernel bodies)
#endif
typedef const PB_CONSTANT_AS pb_sys *pb_const_sys;
typedef const PB_CONSTANT_AS int *pb_const_map; // (system, block) pairs
__device__ __forceinline__ int2 pb_map_at(const int2 *map, long long i)
{
    return make_int2(((pb_const_map)map)[2 * i], ((pb_const_map)map)[2 * i + 1]);
}
// ---- grouped launches (one per phase and class, for many systems) ----------
// sys: device descriptors; map: (system, block) per launched block; nblk:
// launched blocks.  Binning: count and scatter take one block of 256 points
// (sources then targets) per map entry, the scan one workgroup per system.
void pb_group_bin_count(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk);
void pb_group_bin_scan(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk);
void pb_group_bin_scatter(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk);
void pb_fill_zero(hipStream_t s, long long n, int *a);
// map[i] = (i / per, i % per) for i < valid, (-1, 0) for valid <= i < n
void pb_fill_map(hipStream_t s, int2 *map, long long n, int per, long long valid);
// gather map (read after the XCD renumbering) that passes nsys systems of
// per blocks through all XCDs one system at a time; returns the entries
long long pb_fill_map_xcd(hipStream_t s, int2 *map, int nsys, int per);
// Pair kernel of one class (far32, table degree; tdeg 0: closed forms) with
// nsplit wavefronts per tile and lds bytes of dynamic LDS (the largest table).
int pb_group_pair(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int nsplit, bool far32,
                  int tdeg, size_t lds);
// k space of one precision (mixed: FP32 grids): spread (one block per 8^3
// region), the FFT passes (pass 0..4 as in the single-system pipeline, for
// systems whose transformed axis has length n), the gather (window P on
// every axis, or P = 0 for per-axis windows; map over blocks of
// 4 targets, XCD-renumbered: nblk a multiple of 8, entries (-1, 0) idle).
void pb_group_spread(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision);
// FP32 spreading of pb_spread_sets force sets per workgroup that share the
// particles (consecutive descriptors): map entries (set group, region)
constexpr int pb_spread_sets = 8;
void pb_group_spread_sets(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk);
int pb_group_fft(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision, int pass, int n,
                 const void *twiddles);
int pb_group_interp(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision, int P);
// blocks of one system in each grouped launch
int pb_group_fft_blocks(int precision, int pass, const pb_kgrid &g);
// The FFT passes of grid g in order (pb_group_fft's pass numbers: 0 z forward,
// 1 y forward, 2 x with the multiply, 3 y inverse, 4 z inverse, 5 and 6 the
// fused (y, z) plane passes forward and inverse); returns their number (3 or
// 5).  pb_fft_axis: the axis whose length a pass's launch is classed by.
int pb_fft_passes(int precision, const pb_kgrid &g, int *passes);
int pb_fft_axis(int pass);
constexpr int pb_interp_waves = 4; // wavefronts per gather block
// targets per gather wavefront for window P (P^2 window columns per target:
// 4 targets fill whole rounds of 64 lanes where P^2 is not a multiple of 64;
// only up to P = 12, beyond which the unrolled rounds spill registers)
constexpr __host__ __device__ int pb_interp_tw(int P) { return P >= 8 && P <= 12 && (P * P) % 64 != 0 ? 4 : 1; }
constexpr int pb_pair_waves = 4;   // wavefronts per pair-kernel workgroup
// the fused binning (one scan workgroup) applies to this key space
bool pb_bin_fused_ok(const pb_cells &cells);
// FFT twiddles of a precision (uploaded once per context)
int pb_twiddles(pb_context *c, hipStream_t s, int precision, const void **tw);

// Wavefronts per pair-kernel tile: about `rounds` rounds of resident
// wavefronts (16 per compute unit) in total over the tiles (~48 targets each);
// each extra wavefront repeats the tile's set-up.
inline int pb_pair_split(int cu_count, int nt, double rounds)
{
    const double tiles_est = std::max(1.0, nt / 48.0);
    return (int)std::min(16.0, std::max(1.0, std::round(rounds * 16 * cu_count / tiles_est)));
}
// The rounds for a form, measured beside the k space at 1e5 -> 1e5, 128^3,
// P 12 (wavefronts per tile 2, 3, 4, 6): the mixed form with the table 1.1
// (its pair kernel the shorter path; fewer, longer wavefronts leave the
// compute units to the k space: 0.461-0.464 against 0.469-0.470 ms per step
// at 2.25), the FP64 table form 2.25 (1.065 ms; 1.071-1.075 at 1.1 or 3),
// the closed forms 3 (the longer path: 1.075-1.080 against 1.111 ms).
inline double pb_pair_rounds(int precision, size_t table_doubles)
{
    if (table_doubles == 0) return 3.0;
    return precision == PB_PREC_MIXED ? 1.1 : 2.25;
}
// far32: pairs beyond rfar in FP32 (table copy when table_doubles > 0, else
// the FP32 closed form); otherwise every pair in FP64.  rounds: as in
// pb_pair_split.
int pb_run_real(pb_context *c, hipStream_t s, const pb_cells &cells, const pb_kparams *kp_dev,
                size_t table_doubles, int table_deg, const pb_box &box, int ns, int nt, double *u, bool accumulate,
                bool far32, double rounds);
// accumulate: 0 overwrite u, 1 add to u, 2 add with atomics (another kernel
// may be adding to u at the same time).  fork, when given, is called at point
// fork_at (0 before the spread, 1 after it, 2 after the forward transform, 3
// before the gather) with the stream the k space runs on, which it may
// replace for the rest of the k space.
int pb_run_kspace(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells,
                  const pb_box &box, double a, double d, int ns, int nt, double *u, int accumulate,
                  int precision, const std::function<int(hipStream_t &)> *fork = nullptr, int fork_at = 0);
// The spread of this work takes the balanced form (count, chunked items,
// reduction: large inputs), else one workgroup per region.
bool pb_spread_split(const pb_context *c, const pb_kgrid &g, int ns);
// The k table goes to c->dense_tmp (built on the device on stream s).
int pb_run_dense(pb_context *c, hipStream_t s, double *dA, int lda, const double *y, int nt, const double *x,
                 int ns, const pb_kparams &kp, const pb_box &box, const int *lo_real, const int *hi_real,
                 int real_mode, const int *lo_k, const int *hi_k, int k_mode, double *absA);
int pb_run_vR_batch(hipStream_t s, double *A, const double *xh, int n, const pb_kparams &kp);
int pb_run_vF_batch(hipStream_t s, double *A, const double *xh, const double *k, int n,
                    double a, double d);
