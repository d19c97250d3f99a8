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

#include <functional>

#include <climits>
#include <cmath>
#include <vector>

namespace {

constexpr int kPmax = 32;

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Three components packed without padding (the inverse transform's output:
// the gather reads 12 or 24 bytes per window point instead of 16 or 32).
template <typename T> struct trip {
    T x, y, z;
};

template <typename T> struct cplx;
template <> struct cplx<double> {
    using type = double2;
    using type4 = double4;
};
template <> struct cplx<float> {
    using type = float2;
    using type4 = float4;
};

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
// The grid is cut into regions of 8 x 8 x 8 points.  A workgroup (four
// wavefronts, each two 4 x 4 x 4 sub-blocks, one point per lane) spreads onto
// one region the sources whose centred window (first index
// f0 = floor(x/h - P/2) + 1, the oracle's formula) touches it.  Candidates
// come from the real-space cell list: the cells covering
// [(gA - P/2) h, (gB + P/2) h) per axis, unwrapped so that a source near the
// opposite face enters with its periodic image; these cell runs, in a fixed
// order, number the region's candidates 0 .. cand - 1.  The workgroup sets up
// the runs (one per thread), scans their lengths, tests candidates 256 at a
// time and stages the hits in LDS in a fixed order; each wavefront then takes
// the staged sources that touch its sub-block in batches.
//
// FP32, one force set: at staging, the thread that found a source evaluates
// its window weights at the region's 8 points per axis (24 values, zero
// outside the window) and records which 4-point halves of the region the
// window touches; each wavefront then takes the staged sources that touch its
// 4 x 4 x 8 column (both z halves) 32 at a time as matrix products on the
// MFMA units (A = wx F over (x, component), B = wy wz over (y, z), the
// sources the inner dimension; spread_mfma_col).  FP32 with several force
// sets (16 sources per batch, weights evaluated per batch): lane (i, h)
// evaluates source h's weights at offset i of the sub-block, keeps wx[i] F in
// registers and forms P[h][4 i + k] = wy[i] wz[k] in LDS; the lane at point
// (i, j, k) then adds, per source, P[h][4 j + k] times wx F of lane h of its
// own row, taken by a DPP row broadcast folded into the multiply-add.  FP64,
// 32 per batch: one lane per source evaluates the sub-block's 4+4+4 weights
// and stores P[j][k] = wy[j] wz[k] and wx[i] F in LDS.
//
// Work: a region whose candidates are few is one work item.  Dense regions
// (clustered sources) are split into chunks of C candidates, each a work item,
// so that no workgroup serializes a cluster: a count pass numbers each
// region's candidates and gives each split region consecutive partial slots,
// the spreading workgroups take items (the chunks first, then the whole
// regions) from an atomic counter, and a reduction pass adds a split region's
// chunk partials in chunk order.  Every grid point is written exactly once and
// every sum has a fixed order: the result does not depend on scheduling.
constexpr int kSpreadWaves = 4;

template <typename T> struct spread_cfg;
template <> struct spread_cfg<float> { static constexpr int CAP = 256, HB = 16; };
template <> struct spread_cfg<double> { static constexpr int CAP = 256, HB = 32; };

// Per-batch arrays.  FP64: one lane per source writes P[h][4 j + k] = wy[j]
// wz[k] and wx[i] F.  FP32: P[h][4 i + k] = wy[i] wz[k] (wx F stays in the
// registers of the lane that formed it).
template <typename T> struct spread_batch_lds {
    T P[kSpreadWaves][spread_cfg<T>::HB][17];                        // wy[j] wz[k] at 4 j + k (odd stride)
    typename cplx<T>::type4 wxF[kSpreadWaves][spread_cfg<T>::HB][4]; // wx[i] (fx, fy, fz)
};
template <> struct alignas(16) spread_batch_lds<float> {
    alignas(16) float Wz[kSpreadWaves][16][4]; // wz[k] of source h (weights evaluated per batch)
    alignas(16) float P[kSpreadWaves][16][16]; // wy[j] wz[k] of source h at 4 j + k
};

// Staged sources.  FP32 with one force set: one record per source of WS
// floats: the weights at the region's 8 points on each axis (24), the scaled
// force, and the mask of the 4-point halves its window touches (bit 2 a + h);
// the weights are evaluated once per source.  Otherwise (FP64; FP32 with
// several force sets, whose records would not fit at this occupancy) the
// wrapped position, shifted window start, image code and scaled forces, and
// the batches evaluate the weights.
template <typename T, int NS> struct spread_stage {
    static constexpr bool PREW = false;
    static constexpr int CAP = spread_cfg<T>::CAP;
    double sx[3][CAP];
    int sf0[3][CAP];
    unsigned scode[CAP];
    T sF[3 * NS][CAP];
};
template <> struct spread_stage<float, 1> {
    static constexpr bool PREW = true;
    static constexpr int CAP = spread_cfg<float>::CAP;
    static constexpr int WS = 28, MASK = 27;
    alignas(16) float st[CAP][WS];
};

// NS > 1 (FP32 only): NS force sets spread together from one staging of the
// sources (the many-right-hand-side k space).
// The per-batch arrays of the FP64 and several-force-set forms (the one-set
// FP32 form reads the staged records directly).
struct spread_no_batch_lds {};
template <typename T, int NS> struct spread_batch_for {
    using type = spread_batch_lds<T>;
};
template <> struct spread_batch_for<float, 1> {
    using type = spread_no_batch_lds;
};

template <typename T, int NS = 1>
struct spread_lds : spread_batch_for<T, NS>::type, spread_stage<T, NS> {
    static constexpr int CAP = spread_stage<T, NS>::CAP, HB = spread_cfg<T>::HB;
    static_assert(CAP == 256, "a chunk of 256 candidates fits the staging; staged indices are bytes");
    int run_j[256];         // first source of each candidate run of the chunk
    int run_pre[257];       // exclusive prefix of the run lengths
    unsigned run_code[256]; // image shifts (l0 + 4) | (l1 + 4) << 4 | (l2 + 4) << 8
    int wcount[kSpreadWaves];
    // staged indices (< CAP) touching the wavefront's columns: fewer than a
    // batch pending plus 64 appended
    unsigned char ring[kSpreadWaves][128];
};

// Region rid: its first and last grid points per axis and its candidate runs
// (cell columns (u0, u1) x z segments of one periodic image each).
struct region_geo {
    int gA[3], gB[3], ulo[3], uhi[3];
    int ncy, lz0, nzseg;
    long long nruns;
};

__device__ __forceinline__ bool region_of(int rid, const pb_cells &cells, const pb_kgrid &g, region_geo &R)
{
    const int nr0 = (g.n[0] + 7) / 8, nr1 = (g.n[1] + 7) / 8, nr2 = (g.n[2] + 7) / 8;
    if (rid < 0 || rid >= nr0 * nr1 * nr2) return false;
    const int rr[3] = {rid / (nr1 * nr2), (rid / nr2) % nr1, rid % nr2};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        R.gA[a] = 8 * rr[a];
        R.gB[a] = min(R.gA[a] + 7, g.n[a] - 1);
        const double lo = (R.gA[a] - 0.5 * g.P[a]) * g.h[a], hi = (R.gB[a] + 0.5 * g.P[a]) * g.h[a];
        R.ulo[a] = (int)floor(lo * cells.inv_h[a] - 1e-9);
        R.uhi[a] = (int)floor(hi * cells.inv_h[a] + 1e-9);
    }
    const int ncx = R.uhi[0] - R.ulo[0] + 1;
    R.ncy = R.uhi[1] - R.ulo[1] + 1;
    R.lz0 = floordiv(R.ulo[2], cells.nc[2]);
    R.nzseg = floordiv(R.uhi[2], cells.nc[2]) - R.lz0 + 1;
    R.nruns = (long long)ncx * R.ncy * R.nzseg;
    return true;
}

// Run r of the region: sources j .. j + len - 1 of one cell column segment
// and its image code.
__device__ __forceinline__ void region_run(const region_geo &R, long long r, const pb_cells &cells,
                                           const int *__restrict__ cstart, int &j, int &len, unsigned &code)
{
    const int nc0 = cells.nc[0], nc1 = cells.nc[1], nc2 = cells.nc[2];
    const int seg = (int)(r % R.nzseg);
    const long long col = r / R.nzseg;
    const int u0 = R.ulo[0] + (int)(col / R.ncy), u1 = R.ulo[1] + (int)(col % R.ncy);
    const int l0 = floordiv(u0, nc0), l1 = floordiv(u1, nc1), l2 = R.lz0 + seg;
    const int zs = max(R.ulo[2], l2 * nc2) - l2 * nc2, ze = min(R.uhi[2], (l2 + 1) * nc2 - 1) - l2 * nc2;
    const int colbase = ((u0 - l0 * nc0) * nc1 + (u1 - l1 * nc1)) * nc2;
    j = cstart[colbase + zs];
    len = cstart[colbase + ze + 1] - j;
    code = (unsigned)(l0 + 4) | ((unsigned)(l1 + 4) << 4) | ((unsigned)(l2 + 4) << 8);
}

// FP64 weights of staged source e on the sub-block at sb (lane = batch slot).
template <typename T, int NS>
__device__ __forceinline__ void spread_weights(spread_lds<T, NS> &S, int wave, int lane, int e, const int *sb,
                                               const pb_kgrid &g)
{
    T w[3][4];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int l = (int)((S.scode[e] >> (4 * a)) & 15) - 4;
        const double x = S.sx[a][e];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int gq = sb[a] + q;
            const int m = gq - S.sf0[a][e];
            T v = (T)0;
            if (m >= 0 && m < g.P[a]) {
                const double dx = (double)(gq - l * g.n[a]) * g.h[a] - x;
                v = (T)exp(-dx * dx / g.eta);
            }
            w[a][q] = v;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) S.P[wave][lane][4 * j + k] = w[1][j] * w[2][k];
    const T f0 = S.sF[0][e], f1 = S.sF[1][e], f2 = S.sF[2][e];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        typename cplx<T>::type4 v;
        v.x = w[0][i] * f0;
        v.y = w[0][i] * f1;
        v.z = w[0][i] * f2;
        v.w = (T)0;
        S.wxF[wave][lane][i] = v;
    }
}

// Value v of lane H of this lane's row of 16 (DPP row_newbcast).
template <int H> __device__ __forceinline__ float row_bcast(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + H, 0xf,
                                                                 0xf, true));
}

template <int H>
__device__ __forceinline__ void spread_steps16(const float (&P)[16][16], int jk, float f0, float f1, float f2,
                                               float &a0, float &a1, float &a2)
{
    const float p = P[H][jk];
    a0 = fmaf(row_bcast<H>(f0), p, a0);
    a1 = fmaf(row_bcast<H>(f1), p, a1);
    a2 = fmaf(row_bcast<H>(f2), p, a2);
    if constexpr (H < 15) spread_steps16<H + 1>(P, jk, f0, f1, f2, a0, a1, a2);
}

// The 48 multiply-adds of a batch of 16 sources per force set, with the row
// broadcast folded into each (v_fmac_f32 with a DPP source): the lane at point
// (i, j, k) adds, for each source h, P[h][4 j + k] times wx F of lane h of its
// own row.  The leading s_nop 4 covers the wait states a DPP read needs after
// a VALU write of its source (two) or of the execution mask (five).  Build
// with PB_SPREAD_DPP_MOV for the compiler's form (a DPP move before each
// multiply-add).
// p[h] = P[h][4 j + k] of this lane's point, read once per batch.
__device__ __forceinline__ void spread_p16(const float (&P)[16][16], int lane, float (&p)[16])
{
#pragma unroll
    for (int q = 0; q < 16; q++) p[q] = P[q][lane & 15];
}

__device__ __forceinline__ void spread_fma48(const float (&P)[16][16], const float (&p)[16], int lane, float f0,
                                             float f1, float f2, float (&acc)[3])
{
#ifndef PB_SPREAD_DPP_MOV
    (void)P;
    (void)lane;
#define PB_FD(H, A, F, Q) "v_fmac_f32_dpp %" #A ", %" #F ", %" #Q " row_newbcast:" #H " row_mask:0xf bank_mask:0xf\n"
#define PB_F3(H, Q) PB_FD(H, 0, 3, Q) PB_FD(H, 1, 4, Q) PB_FD(H, 2, 5, Q)
    asm volatile("s_nop 4\n" PB_F3(0, 6) PB_F3(1, 7) PB_F3(2, 8) PB_F3(3, 9) PB_F3(4, 10) PB_F3(5, 11) PB_F3(6, 12)
                     PB_F3(7, 13) PB_F3(8, 14) PB_F3(9, 15) PB_F3(10, 16) PB_F3(11, 17) PB_F3(12, 18) PB_F3(13, 19)
                         PB_F3(14, 20) PB_F3(15, 21)
                 : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2])
                 : "v"(f0), "v"(f1), "v"(f2), "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]),
                   "v"(p[6]), "v"(p[7]), "v"(p[8]), "v"(p[9]), "v"(p[10]), "v"(p[11]), "v"(p[12]), "v"(p[13]),
                   "v"(p[14]), "v"(p[15]));
#undef PB_F3
#undef PB_FD
#else
    (void)p;
    spread_steps16<0>(P, lane & 15, f0, f1, f2, acc[0], acc[1], acc[2]);
#endif
}

// FP32 batch of up to 16 staged sources, NS force sets, weights evaluated
// here: lane (i, h) evaluates source h's weight at offset i of the sub-block
// on each axis (zero outside the window or for h >= cnt), keeps wx[i] F,
// publishes wz and forms P[h][4 i + k] = wy[i] wz[k]; the 48 multiply-adds
// then run once per force set with the same weights.
template <int NS>
__device__ __forceinline__ void spread_batch16r(spread_lds<float, NS> &S, int wave, int lane, int head, int cnt,
                                                const int *sb, const pb_kgrid &g, float (&acc)[NS][3])
{
    const int i = lane >> 4, h = lane & 15;
    __builtin_amdgcn_wave_barrier();
    const int e = S.ring[wave][(head + (h < cnt ? h : 0)) & 127];
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int l = (int)((S.scode[e] >> (4 * a)) & 15) - 4;
        const int gq = sb[a] + i;
        const int m = gq - S.sf0[a][e];
        float v = 0.0f;
        if (h < cnt && m >= 0 && m < g.P[a]) {
            const double dx = (double)(gq - l * g.n[a]) * g.h[a] - S.sx[a][e];
            v = __expf(-(float)(dx * dx) * (float)(1.0 / g.eta));
        }
        w[a] = v;
    }
    S.Wz[wave][h][i] = w[2];
    __builtin_amdgcn_wave_barrier();
    const float4 z = *reinterpret_cast<const float4 *>(&S.Wz[wave][h][0]);
    *reinterpret_cast<float4 *>(&S.P[wave][h][4 * i]) = make_float4(w[1] * z.x, w[1] * z.y, w[1] * z.z, w[1] * z.w);
    __builtin_amdgcn_wave_barrier();
    float p[16];
    spread_p16(S.P[wave], lane, p);
#pragma unroll
    for (int q = 0; q < NS; q++)
        spread_fma48(S.P[wave], p, lane, w[0] * S.sF[3 * q][e], w[0] * S.sF[3 * q + 1][e], w[0] * S.sF[3 * q + 2][e],
                     acc[q]);
    __builtin_amdgcn_wave_barrier();
}

template <typename T, int NS>
__device__ __forceinline__ void spread_batch(spread_lds<T, NS> &S, int wave, int lane, int head, int cnt,
                                             const int *sb, const int *off, const pb_kgrid &g, int li, int lj, int lk,
                                             T (&acc)[NS][3])
{
    if constexpr (sizeof(T) == 4) {
        spread_batch16r<NS>(S, wave, lane, head, cnt, sb, g, acc);
    } else {
        static_assert(NS == 1, "FP64 spreads one force set");
        T &a0 = acc[0][0], &a1 = acc[0][1], &a2 = acc[0][2];
        __builtin_amdgcn_wave_barrier();
        if (lane < cnt) spread_weights<T, NS>(S, wave, lane, S.ring[wave][(head + lane) & 127], sb, g);
        __builtin_amdgcn_wave_barrier();
        for (int h = 0; h < cnt; h++) {
            const T p = S.P[wave][h][4 * lj + lk];
            const typename cplx<T>::type4 v = S.wxF[wave][h][li];
            a0 += p * v.x;
            a1 += p * v.y;
            a2 += p * v.z;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// FP32, one force set, both z halves of the wavefront's 4 x 4 x 8 column at
// once: up to 16 NB staged records as 4 NB steps of two
// v_mfma_f32_16x16x4_f32 each, one per z half h:
//   D_h[4 i + c][4 j + k] += sum over the step's four records r of
//                            A[4 i + c][r] B_h[r][4 j + k],
//   A = wx[i] F[c] (x offset off0; row c = 3 zero), B_h = wy[j] wz[4 h + k]
// (y offset off1).  Lane l supplies A[l % 16][l / 16] and B_h[l / 16][l % 16]
// of record 4 s + l / 16 and holds D_h[4 (l / 16) + c][l % 16] in d[h][c]:
// the point (i, j, k) = (l / 16, (l / 4) % 4, l % 4) of the sub-block,
// component c, as in the rest of the spread.  Records past cnt add zero.
template <int NB>
__device__ __forceinline__ void spread_mfma_col(spread_lds<float, 1> &S, int wave, int lane, int head, int cnt,
                                                int off0, int off1, floatx4 (&d)[2])
{
    constexpr int NSTEP = 4 * NB;
    const int g = lane >> 4, hi = (lane >> 2) & 3, lo = lane & 3;
    __builtin_amdgcn_wave_barrier();
    int e[NSTEP];
#pragma unroll
    for (int s = 0; s < NSTEP; s++) e[s] = S.ring[wave][(head + (4 * s + g < cnt ? 4 * s + g : 0)) & 127];
    float wx[NSTEP], fc[NSTEP], wy[NSTEP], wz0[NSTEP], wz1[NSTEP];
#pragma unroll
    for (int s = 0; s < NSTEP; s++) {
        const float *rec = S.st[e[s]];
        wx[s] = rec[off0 + hi];
        fc[s] = rec[24 + lo]; // lo = 3: the mask word, not used
        wy[s] = rec[8 + off1 + hi];
        wz0[s] = rec[16 + lo];
        wz1[s] = rec[20 + lo];
    }
#pragma unroll
    for (int s = 0; s < NSTEP; s++) {
        const float a = 4 * s + g < cnt && lo < 3 ? wx[s] * fc[s] : 0.0f;
        d[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wy[s] * wz0[s], d[0], 0, 0, 0);
        d[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wy[s] * wz1[s], d[1], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();
}

// The FP32 one-force-set form of spread_process: the wavefront takes each
// staged source whose window touches its 4 x 4 column (either z half) once.
__device__ __forceinline__ void spread_process_col(spread_lds<float, 1> &S, int wave, int lane, int nst,
                                                   const region_geo &R, const pb_kgrid &g, float (&acc)[2][1][3])
{
    constexpr int NB = 2;
    const int hx = wave >> 1, hy = wave & 1;
    if (R.gA[0] + 4 * hx >= g.n[0] || R.gA[1] + 4 * hy >= g.n[1]) return; // uniform per wavefront
    const unsigned need = (1u << hx) | (1u << (2 + hy));
    floatx4 d[2];
#pragma unroll
    for (int sz = 0; sz < 2; sz++) d[sz] = floatx4{acc[sz][0][0], acc[sz][0][1], acc[sz][0][2], 0.0f};
    int head = 0, pend = 0;
    for (int eb = 0; eb < nst; eb += 64) {
        const int e = eb + lane;
        bool in = e < nst;
        if (in) {
            const unsigned m = __float_as_uint(S.st[e][spread_stage<float, 1>::MASK]);
            in = (m & need) == need && (m & 0x30u) != 0;
        }
        const unsigned long long m = __ballot(in);
        if (in) {
            const unsigned pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            S.ring[wave][(head + pend + pre) & 127] = e;
        }
        pend += __popcll(m);
        while (pend >= 16 * NB) {
            spread_mfma_col<NB>(S, wave, lane, head, 16 * NB, 4 * hx, 4 * hy, d);
            head = (head + 16 * NB) & 127;
            pend -= 16 * NB;
        }
    }
    if (pend > 16) spread_mfma_col<2>(S, wave, lane, head, pend, 4 * hx, 4 * hy, d);
    else if (pend > 0) spread_mfma_col<1>(S, wave, lane, head, pend, 4 * hx, 4 * hy, d);
#pragma unroll
    for (int sz = 0; sz < 2; sz++) {
        acc[sz][0][0] = d[sz][0];
        acc[sz][0][1] = d[sz][1];
        acc[sz][0][2] = d[sz][2];
    }
}

// The wavefront's two sub-blocks (z halves of its 4 x 4 column of the region)
// gather, one after the other, from the nst staged sources (FP64; FP32 with
// several force sets).
template <typename T, int NS>
__device__ __forceinline__ void spread_process_sub(spread_lds<T, NS> &S, int wave, int lane, int nst,
                                                   const region_geo &R, const pb_kgrid &g, T (&acc)[2][NS][3])
{
    constexpr int HB = spread_cfg<T>::HB;
    const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
    const int hx = wave >> 1, hy = wave & 1;
#pragma unroll
    for (int sz = 0; sz < 2; sz++) {
        const int off[3] = {4 * hx, 4 * hy, 4 * sz};
        const int sb[3] = {R.gA[0] + off[0], R.gA[1] + off[1], R.gA[2] + off[2]};
        if (sb[0] >= g.n[0] || sb[1] >= g.n[1] || sb[2] >= g.n[2]) continue; // uniform per wavefront
        int head = 0, pend = 0;
        for (int eb = 0; eb < nst; eb += 64) {
            const int e = eb + lane;
            bool in = e < nst;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int f0 = in ? S.sf0[a][e] : 0;
                in = in && f0 > sb[a] - g.P[a] && f0 <= sb[a] + 3;
            }
            const unsigned long long m = __ballot(in);
            if (in) {
                const unsigned pre = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                S.ring[wave][(head + pend + pre) & 127] = e;
            }
            pend += __popcll(m);
            while (pend >= HB) {
                spread_batch<T, NS>(S, wave, lane, head, HB, sb, off, g, li, lj, lk, acc[sz]);
                head = (head + HB) & 127;
                pend -= HB;
            }
        }
        if (pend > 0) spread_batch<T, NS>(S, wave, lane, head, pend, sb, off, g, li, lj, lk, acc[sz]);
    }
}

// The wavefronts' share of the nst staged sources onto their columns of the
// region (acc: each wavefront's two 4 x 4 x 4 sub-blocks, z halves).
template <typename T, int NS>
__device__ __forceinline__ void spread_process(spread_lds<T, NS> &S, int wave, int lane, int nst, const region_geo &R,
                                               const pb_kgrid &g, T (&acc)[2][NS][3])
{
    if constexpr (spread_stage<T, NS>::PREW)
        spread_process_col(S, wave, lane, nst, R, g, acc);
    else
        spread_process_sub(S, wave, lane, nst, R, g, acc);
}

// Candidates q0 .. q1 - 1 of region R (numbered in run order) spread onto the
// region's points in acc (the calling workgroup's four wavefronts).
template <typename T, int NS>
__device__ __forceinline__ void spread_chunk(spread_lds<T, NS> &S, const region_geo &R, long long q0, long long q1,
                                             const int *__restrict__ cstart, const double4 *__restrict__ spos,
                                             const double4 *const (&sfrc)[NS], const pb_cells &cells,
                                             const pb_kgrid &g, T (&acc)[2][NS][3])
{
    constexpr int CAP = spread_lds<T, NS>::CAP;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int sz = 0; sz < 2; sz++)
#pragma unroll
        for (int q = 0; q < NS; q++) acc[sz][q][0] = acc[sz][q][1] = acc[sz][q][2] = (T)0;
    int nst = 0;
    long long base = 0; // candidates of the runs before this chunk of runs
    for (long long rb = 0; rb < R.nruns && base < q1; rb += 256) {
        // one candidate run per thread: column (u0, u1), z segment with one image
        const long long r = rb + tid;
        int j = 0, len = 0;
        unsigned code = 0;
        if (r < R.nruns) region_run(R, r, cells, cstart, j, len, code);
        // exclusive scan of the run lengths: within each wavefront by
        // shuffles, then the wavefront totals (one block barrier)
        int incl = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane == 63) S.wcount[wave] = incl;
        S.run_j[tid] = j;
        S.run_code[tid] = code;
        __syncthreads();
        int wpre = 0;
#pragma unroll
        for (int w = 0; w < kSpreadWaves; w++) wpre += w < wave ? S.wcount[w] : 0;
        S.run_pre[tid + 1] = wpre + incl;
        if (tid == 0) S.run_pre[0] = 0;
        __syncthreads();
        const int total = S.run_pre[256];
        // this chunk's share of [q0, q1) (uniform)
        const long long lo = q0 > base ? q0 - base : 0, hi = q1 - base < total ? q1 - base : total;
        for (long long qb = lo; qb < hi; qb += 256) {
            const long long q = qb + tid;
            bool hit = false;
            int src = 0, f0[3] = {0, 0, 0};
            unsigned cq = 0;
            double xw[3] = {0, 0, 0};
            double4 fw = {0, 0, 0, 0};
            if (q < hi) {
                const int qi = (int)q;
                int lo2 = 0, hi2 = 255; // last run with run_pre <= q
                while (lo2 < hi2) {
                    const int mid = (lo2 + hi2 + 1) >> 1;
                    if (S.run_pre[mid] <= qi)
                        lo2 = mid;
                    else
                        hi2 = mid - 1;
                }
                src = S.run_j[lo2] + (qi - S.run_pre[lo2]);
                cq = S.run_code[lo2];
                if constexpr (spread_stage<T, NS>::PREW) fw = sfrc[0][src]; // loaded beside the position
                const double4 x = spos[src];
                xw[0] = x.x;
                xw[1] = x.y;
                xw[2] = x.z;
                hit = true;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const int l = (int)((cq >> (4 * a)) & 15) - 4;
                    f0[a] = (int)floor(xw[a] * g.inv_h[a] - 0.5 * g.P[a]) + 1 + l * g.n[a];
                    hit = hit && f0[a] > R.gA[a] - g.P[a] && f0[a] <= R.gB[a];
                }
            }
            // staged in a fixed order: wavefront by wavefront, lanes in order
            const unsigned long long m = __ballot(hit);
            if (lane == 0) S.wcount[wave] = __popcll(m);
            __syncthreads();
            int wbase = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < kSpreadWaves; w++) {
                wbase += w < wave ? S.wcount[w] : 0;
                tot += S.wcount[w];
            }
            if (nst + tot > CAP) { // no room: gather from what is staged, then restart
                spread_process<T, NS>(S, wave, lane, nst, R, g, acc);
                nst = 0;
                __syncthreads();
            }
            if (hit) {
                const int e = nst + wbase +
                              (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                if constexpr (spread_stage<T, NS>::PREW) {
                    // the weights at the region's 8 points per axis, once per source
                    // (branch-free: each weight evaluated, zero outside the window)
                    float *rec = S.st[e];
                    unsigned mask = 0;
                    const float c2 = (float)(-1.4426950408889634 / g.eta); // e^{-t} = 2^{-t log2 e}
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const int l = (int)((cq >> (4 * a)) & 15) - 4;
                        double dx = (double)(R.gA[a] - l * g.n[a]) * g.h[a] - xw[a];
                        const int m0 = R.gA[a] - f0[a];
                        float w8[8];
#pragma unroll
                        for (int p = 0; p < 8; p++) {
                            const float v = __builtin_amdgcn_exp2f((float)(dx * dx) * c2);
                            w8[p] = (unsigned)(m0 + p) < (unsigned)g.P[a] ? v : 0.0f;
                            dx += g.h[a];
                        }
                        *reinterpret_cast<float4 *>(rec + 8 * a) = make_float4(w8[0], w8[1], w8[2], w8[3]);
                        *reinterpret_cast<float4 *>(rec + 8 * a + 4) = make_float4(w8[4], w8[5], w8[6], w8[7]);
                        mask |= (f0[a] > R.gA[a] - g.P[a] && f0[a] <= R.gA[a] + 3 ? 1u : 0u) << (2 * a);
                        mask |= (f0[a] > R.gA[a] + 4 - g.P[a] && f0[a] <= R.gA[a] + 7 ? 1u : 0u) << (2 * a + 1);
                    }
                    const double4 f = fw;
                    rec[24] = (float)(f.x * g.scale);
                    rec[25] = (float)(f.y * g.scale);
                    rec[26] = (float)(f.z * g.scale);
                    rec[spread_stage<T, NS>::MASK] = __uint_as_float(mask);
                } else {
                    S.sx[0][e] = xw[0];
                    S.sx[1][e] = xw[1];
                    S.sx[2][e] = xw[2];
                    S.sf0[0][e] = f0[0];
                    S.sf0[1][e] = f0[1];
                    S.sf0[2][e] = f0[2];
                    S.scode[e] = cq;
#pragma unroll
                    for (int q = 0; q < NS; q++) {
                        const double4 f = sfrc[q][src];
                        S.sF[3 * q][e] = (T)(f.x * g.scale);
                        S.sF[3 * q + 1][e] = (T)(f.y * g.scale);
                        S.sF[3 * q + 2][e] = (T)(f.z * g.scale);
                    }
                }
            }
            nst += tot;
            __syncthreads();
        }
        base += total;
        __syncthreads(); // the run arrays are rewritten by the next chunk of runs
    }
    spread_process<T, NS>(S, wave, lane, nst, R, g, acc);
}

// The region's points of acc to the grid (every point of the region inside
// the grid, exactly once).
template <typename T, int NS>
__device__ __forceinline__ void spread_store(const region_geo &R, const pb_kgrid &g, const T (&acc)[2][NS][3],
                                             typename cplx<T>::type4 *const (&grid4)[NS])
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane >> 4, lj = (lane >> 2) & 3, lk = lane & 3;
    const int sbx = R.gA[0] + 4 * (wave >> 1), sby = R.gA[1] + 4 * (wave & 1);
#pragma unroll
    for (int sz = 0; sz < 2; sz++) {
        const int p0 = sbx + li, p1 = sby + lj, p2 = R.gA[2] + 4 * sz + lk;
        if (sbx >= g.n[0] || sby >= g.n[1] || R.gA[2] + 4 * sz >= g.n[2]) continue;
#pragma unroll
        for (int q = 0; q < NS; q++) {
            typename cplx<T>::type4 o;
            o.x = acc[sz][q][0];
            o.y = acc[sz][q][1];
            o.z = acc[sz][q][2];
            o.w = (T)0;
            grid4[q][((size_t)p0 * g.n[1] + p1) * g.n[2] + p2] = o;
        }
    }
}

// Local index of a region point: ((x * 8) + y) * 8 + z.
__device__ __forceinline__ int region_local(int wave, int lane, int sz)
{
    const int px = 4 * (wave >> 1) + (lane >> 4), py = 4 * (wave & 1) + ((lane >> 2) & 3), pz = 4 * sz + (lane & 3);
    return (px * 8 + py) * 8 + pz;
}

// A chunk's partial sums to its slot (512 points).
template <typename T>
__device__ __forceinline__ void spread_store_part(const T (&acc)[2][1][3], typename cplx<T>::type4 *__restrict__ slot)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int sz = 0; sz < 2; sz++) {
        typename cplx<T>::type4 o;
        o.x = acc[sz][0][0];
        o.y = acc[sz][0][1];
        o.z = acc[sz][0][2];
        o.w = (T)0;
        slot[region_local(wave, lane, sz)] = o;
    }
}

// Whole regions, one per workgroup: single system (block = region) and
// grouped systems (map entry = (system, region)).
template <typename T>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_spread_regions(
    const int *__restrict__ cstart, const double4 *__restrict__ spos,
    const double4 *__restrict__ sfrc, pb_cells cells, pb_kgrid g,
    typename cplx<T>::type4 *__restrict__ grid4)
{
    __shared__ spread_lds<T, 1> S;
    const double4 *const fr[1] = {sfrc};
    typename cplx<T>::type4 *const gr[1] = {grid4};
    region_geo R;
    if (!region_of((int)blockIdx.x, cells, g, R)) return; // whole workgroup
    T acc[2][1][3];
    spread_chunk<T, 1>(S, R, 0, LLONG_MAX, cstart, spos, fr, cells, g, acc);
    spread_store<T, 1>(R, g, acc, gr);
}

template <typename T>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_gspread(const pb_sys *__restrict__ sys,
                                                               const int2 *__restrict__ map)
{
    __shared__ spread_lds<T, 1> S;
    const int2 m = pb_map_at(map, blockIdx.x);
    const auto &D = ((pb_const_sys)sys)[m.x];
    const double4 *const fr[1] = {D.sfrc};
    typename cplx<T>::type4 *const gr[1] = {(typename cplx<T>::type4 *)D.grid4};
    const pb_cells cells = D.cells;
    const pb_kgrid g = D.g;
    region_geo R;
    if (!region_of(m.y, cells, g, R)) return;
    T acc[2][1][3];
    spread_chunk<T, 1>(S, R, 0, LLONG_MAX, D.sstart, D.spos, fr, cells, g, acc);
    spread_store<T, 1>(R, g, acc, gr);
}

// Grouped, NS force sets per workgroup: block b spreads sets NS map[b].x ..
// NS map[b].x + NS - 1 of region map[b].y (the sets share particles and grid
// geometry, the first one's descriptor gives them).
template <int NS>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_gspread_sets(const pb_sys *__restrict__ sys,
                                                                    const int2 *__restrict__ map)
{
    __shared__ spread_lds<float, NS> S;
    const int2 m = pb_map_at(map, blockIdx.x);
    const pb_const_sys D = (pb_const_sys)sys + NS * m.x;
    const double4 *fr[NS];
    float4 *gr[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
        fr[q] = D[q].sfrc;
        gr[q] = (float4 *)D[q].grid4;
    }
    const double4 *const(&frc)[NS] = fr;
    float4 *const(&grd)[NS] = gr;
    const pb_cells cells = D[0].cells;
    const pb_kgrid g = D[0].g;
    region_geo R;
    if (!region_of(m.y, cells, g, R)) return;
    float acc[2][NS][3];
    spread_chunk<float, NS>(S, R, 0, LLONG_MAX, D[0].sstart, D[0].spos, frc, cells, g, acc);
    spread_store<float, NS>(R, g, acc, grd);
}

// ---- balanced spreading of one system: count, items, reduction -----------------
// Counters: tot[0] partial slots handed out, tot[1] split regions, tot[2] the
// work counter (all zero before the count pass).
__global__ void k_spread_reset(int *__restrict__ tot)
{
    if (threadIdx.x < 4) tot[threadIdx.x] = 0;
}

// Count: one wavefront per region adds up the lengths of its candidate runs.
// A region of more than C candidates is split into nch = ceil(cand / C)
// chunks: it takes nch consecutive partial slots (base = pbase[r], chunk c in
// slot base + c, owner[base + c] = (r, c)) and joins the list of split regions.
// Slots and list positions follow the order of the atomics, but each region's
// chunks keep their order, so the sums do not depend on it.
__global__ __launch_bounds__(256) void k_spread_count(const int *__restrict__ cstart, pb_cells cells, pb_kgrid g,
                                                      int nreg, int C, int *__restrict__ cand, int *__restrict__ pbase,
                                                      int *__restrict__ mlist, int2 *__restrict__ owner, int part_cap,
                                                      int *__restrict__ tot)
{
    const int e = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    region_geo R;
    if (e >= nreg || !region_of(e, cells, g, R)) return; // whole wavefront
    int sum = 0;
    for (long long r = lane; r < R.nruns; r += 64) {
        int j, len;
        unsigned code;
        region_run(R, r, cells, cstart, j, len, code);
        sum += len;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o, 64);
    const int nch = sum > C ? (sum - 1) / C + 1 : 1;
    int base = 0;
    if (lane == 0) {
        cand[e] = sum;
        if (nch > 1) {
            base = atomicAdd(&tot[0], nch);
            pbase[e] = base;
            if (base + nch <= part_cap) mlist[atomicAdd(&tot[1], 1)] = e;
        }
    }
    if (nch > 1) { // (uniform: sum is the same in every lane)
        base = __shfl(base, 0, 64);
        for (int c = lane; c < nch; c += 64)
            if (base + c < part_cap) owner[base + c] = make_int2(e, c);
    }
}

// Items: each workgroup takes the next item from the counter until none is
// left.  Items 0 .. nslots - 1 are the chunks of split regions (owner[i]; the
// partial goes to slot i), then item nslots + r is region r whole (skipped if
// it was split).
template <typename T>
__global__ __launch_bounds__(64 * kSpreadWaves) void k_spread_items(
    const int *__restrict__ cstart, const double4 *__restrict__ spos, const double4 *__restrict__ sfrc,
    pb_cells cells, pb_kgrid g, typename cplx<T>::type4 *__restrict__ grid4, int nreg, int C,
    const int *__restrict__ cand, const int2 *__restrict__ owner, int part_cap,
    typename cplx<T>::type4 *__restrict__ part, int *__restrict__ tot)
{
    __shared__ spread_lds<T, 1> S;
    __shared__ int s_item;
    const double4 *const fr[1] = {sfrc};
    typename cplx<T>::type4 *const gr[1] = {grid4};
    const int nslots = min(tot[0], part_cap);
    const int nitems = nslots + nreg;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&tot[2], 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads(); // every thread has read s_item before it is written again
        if (it >= nitems) break;
        const bool chunk = it < nslots;
        const int2 oc = chunk ? owner[it] : make_int2(it - nslots, 0);
        const int ce = cand[oc.x];
        if (!chunk && ce > C) continue; // split: its chunks are items of their own
        region_geo R;
        if (!region_of(oc.x, cells, g, R)) continue; // (cannot happen: items name regions)
        T acc[2][1][3];
        const long long q0 = chunk ? (long long)oc.y * C : 0;
        const long long q1 = chunk ? (q0 + C < ce ? q0 + C : (long long)ce) : (long long)ce;
        spread_chunk<T, 1>(S, R, q0, q1, cstart, spos, fr, cells, g, acc);
        if (chunk)
            spread_store_part<T>(acc, part + (size_t)it * 512);
        else
            spread_store<T, 1>(R, g, acc, gr);
        __syncthreads(); // LDS is reused by the next item
    }
}

// Reduction: the chunks of each split region added in chunk order.
template <typename T>
__global__ __launch_bounds__(256) void k_spread_reduce(const int *__restrict__ tot, const int *__restrict__ mlist,
                                                       const int *__restrict__ pbase, const int *__restrict__ cand,
                                                       int C, const typename cplx<T>::type4 *__restrict__ part,
                                                       pb_cells cells, pb_kgrid g,
                                                       typename cplx<T>::type4 *__restrict__ grid4)
{
    const int nm = tot[1];
    for (int k = blockIdx.x; k < nm; k += gridDim.x) {
        const int e = mlist[k];
        const int ce = cand[e];
        const int nch = ce > C ? (ce - 1) / C + 1 : 1;
        const int s0 = pbase[e];
        region_geo R;
        if (!region_of(e, cells, g, R)) continue;
        for (int lp = threadIdx.x; lp < 512; lp += 256) {
            const int p0 = R.gA[0] + (lp >> 6), p1 = R.gA[1] + ((lp >> 3) & 7), p2 = R.gA[2] + (lp & 7);
            if (p0 >= g.n[0] || p1 >= g.n[1] || p2 >= g.n[2]) continue;
            T sx = 0, sy = 0, sz = 0;
            for (int c = 0; c < nch; c++) {
                const typename cplx<T>::type4 v = part[(size_t)(s0 + c) * 512 + lp];
                sx += v.x;
                sy += v.y;
                sz += v.z;
            }
            typename cplx<T>::type4 o;
            o.x = sx;
            o.y = sy;
            o.z = sz;
            o.w = (T)0;
            grid4[((size_t)p0 * g.n[1] + p1) * g.n[2] + p2] = o;
        }
    }
}

// ---- interpolation (gather) --------------------------------------------------
// In FP32 the inverse transform leaves the velocity grid with the y rows
// paired: point (x, y, z) at ((x n1/2 + y/2) n2 + z) 2 + (y & 1) (n1 is a
// power of two, so even), the order the last pass produces its pencil pairs
// in.  A window's z runs of two neighbouring y rows then share cache lines (one
// 288-byte run instead of two 144-byte ones at P = 12).  The FP64 grid keeps
// the plain (x, y, z) order, which its forms run faster with.  The offset is a
// sum of one term per axis: this returns the axis-a term for index j.
template <typename T> __device__ __forceinline__ int pb_grid_offset(int a, int j, int n1, int n2)
{
    if constexpr (sizeof(T) == 4)
        return a == 0 ? j * n1 * n2 : (a == 1 ? (j >> 1) * 2 * n2 + (j & 1) : 2 * j);
    else
        return a == 0 ? j * n1 * n2 : (a == 1 ? j * n2 : j);
}

// u(y) = h^3 (pi eta)^{-3/2} sum_window e^{-|x_g - y|^2/eta} G(x_g), G = (ux,
// uy, uz) per grid point.  One wavefront per target (targets in their sorted,
// spatially compact order, each XCD taking one contiguous range).  The 3P
// window weights and wrapped offsets go to LDS; each lane owns (y, z) pairs
// of the window, forms wy wz and the row offset once, and sweeps the P
// x-planes (weights and offsets the same for the whole wavefront): one vector
// load and three FMAs per point.  The partial sums meet in a wave reduction.
constexpr int kInterpWaves = 4;
static_assert(kInterpWaves == pb_interp_waves, "gather workgroup size");

template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The gather: one wavefront per TW = pb_interp_tw(P) consecutive targets.
// Lane (target w, py, pz) of the TW P^2 window columns sums its P points along
// x; with P^2 not a multiple of 64, TW = 4 targets fill the rounds of 64
// lanes (P = 12: 9 rounds instead of 4 x 3 with a quarter of the lanes idle).
// For P >= 8 a round spans at most two targets, so a lane's partial sums
// go to one of two accumulators fixed at compile time.
// GR: pb_kgrid, or its constant-address-space form in the grouped kernels.
template <typename T, int P, typename GR>
__device__ __forceinline__ void interp_wave(const int i0, int nt, const double4 *__restrict__ tpos,
                                            const int *__restrict__ tperm, const GR &g,
                                            const trip<T> *__restrict__ U3, double *__restrict__ u, int accumulate)
{
    constexpr int TW = pb_interp_tw(P), NQ = P * P, NR = (TW * NQ + 63) / 64;
    static_assert(TW == 1 || NQ >= 64, "a round spans at most two targets");
    __shared__ T wts[kInterpWaves][TW][3][P];
    __shared__ unsigned off[kInterpWaves][TW][3][P]; // wrapped index times its stride, in bytes
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (i0 >= nt) return; // whole wavefront; no block barriers below
    const int nv = min(TW, nt - i0); // targets of this wavefront (the others repeat the last)
    for (int k = lane; k < TW * 3 * P; k += 64) {
        const int w = k / (3 * P), a = (k / P) % 3, p = k % P;
        const double4 y = tpos[i0 + min(w, nv - 1)];
        const double yc = a == 0 ? y.x : (a == 1 ? y.y : y.z);
        const int f0 = (int)floor(yc * g.inv_h[a] - 0.5 * P) + 1;
        const double dx = (double)(f0 + p) * g.h[a] - yc;
        if (sizeof(T) == 8)
            wts[wave][w][a][p] = (T)exp(-dx * dx / g.eta);
        else
            wts[wave][w][a][p] = (T)__expf(-(float)(dx * dx) * (float)(1.0 / g.eta));
        int j = f0 + p;
        j = j < 0 ? j + g.n[a] : (j >= g.n[a] ? j - g.n[a] : j);
        off[wave][w][a][p] = (unsigned)sizeof(trip<T>) * (unsigned)pb_grid_offset<T>(a, j, g.n[1], g.n[2]);
    }
    __builtin_amdgcn_wave_barrier();
    T acc[TW][3];
#pragma unroll
    for (int w = 0; w < TW; w++) acc[w][0] = acc[w][1] = acc[w][2] = (T)0;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int q = r * 64 + lane;
        const int wlo = (r * 64) / NQ;                   // first target of the round (compile time)
        const int whi = wlo + 1 < TW ? wlo + 1 : wlo;    // the next one, if any
        if ((TW * NQ) % 64 == 0 || q < TW * NQ) {
            const bool hi = q >= (wlo + 1) * NQ;
            const int w = hi ? whi : wlo, qq = q - w * NQ;
            const int py = qq / P, pz = qq % P;
            const T wyz = wts[wave][w][1][py] * wts[wave][w][2][pz];
            const unsigned oyz = off[wave][w][1][py] + off[wave][w][2][pz];
            // the column's x sum with the x weights, then times wy wz: one
            // add (a 32-bit byte offset on the grid's base) and three FMAs
            // per point
            const char *col = reinterpret_cast<const char *>(U3);
            T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
            for (int px = 0; px < P; px++) {
                const T wx = wts[wave][w][0][px];
                const trip<T> v = *reinterpret_cast<const trip<T> *>(col + (off[wave][w][0][px] + oyz));
                s0 += wx * v.x;
                s1 += wx * v.y;
                s2 += wx * v.z;
            }
            s0 *= wyz;
            s1 *= wyz;
            s2 *= wyz;
            if (hi) {
                acc[whi][0] += s0;
                acc[whi][1] += s1;
                acc[whi][2] += s2;
            } else {
                acc[wlo][0] += s0;
                acc[wlo][1] += s1;
                acc[wlo][2] += s2;
            }
        }
    }
    const double sc = g.scale * g.vol;
#pragma unroll
    for (int w = 0; w < TW; w++) {
        const T a0 = wave_sum(acc[w][0]), a1 = wave_sum(acc[w][1]), a2 = wave_sum(acc[w][2]);
        if (lane == 0 && w < nv) {
            const size_t o = 3 * (size_t)tperm[i0 + w];
            if (accumulate == 2) { // the pair kernel may be adding to u at the same time
                atomicAdd(&u[o], sc * (double)a0);
                atomicAdd(&u[o + 1], sc * (double)a1);
                atomicAdd(&u[o + 2], sc * (double)a2);
            } else if (accumulate) {
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
}

template <typename T, int P>
__global__ __launch_bounds__(64 * kInterpWaves) void k_interp(int nt, int nblk8,
                                                            const double4 *__restrict__ tpos,
                                                            const int *__restrict__ tperm, pb_kgrid g,
                                                            const trip<T> *__restrict__ U3,
                                                            double *__restrict__ u, int accumulate)
{
    constexpr int TW = pb_interp_tw(P);
    interp_wave<T, P>((pb_xcd_block(blockIdx.x, nblk8) * kInterpWaves + (int)(threadIdx.x >> 6)) * TW, nt, tpos,
                      tperm, g, U3, u, accumulate);
}

// Grouped: blocks renumbered per XCD before the map lookup; the gather adds
// to u with atomics (acc_mode 2: the pair kernel may run at the same time)
// or overwrites it (0).
template <typename T, int P>
__global__ __launch_bounds__(64 * kInterpWaves) void k_ginterp(const pb_sys *__restrict__ sys,
                                                             const int2 *__restrict__ map, int nblk8)
{
    constexpr int TW = pb_interp_tw(P);
    const int2 m = pb_map_at(map, pb_xcd_block(blockIdx.x, nblk8));
    if (m.x < 0) return;
    const auto &S = ((pb_const_sys)sys)[m.x];
    interp_wave<T, P>((m.y * kInterpWaves + (int)(threadIdx.x >> 6)) * TW, S.nt, S.tpos, S.tperm, S.g,
                      (const trip<T> *)S.grid4, S.u, S.acc_mode);
}

// Per-axis window sizes that differ (fftBrinkman's P[3]) take this path.
template <typename T, typename GR>
__device__ __forceinline__ void interp_any_target(const int i, int nt, const double4 *__restrict__ tpos,
                                                  const int *__restrict__ tperm, const GR &g,
                                                  const trip<T> *__restrict__ U3, double *__restrict__ u,
                                                  int accumulate)
{
    if (i >= nt) return;
    const double4 y = tpos[i];
    const double yc[3] = {y.x, y.y, y.z};
    int f0[3];
    for (int a = 0; a < 3; a++) f0[a] = (int)floor(yc[a] * g.inv_h[a] - 0.5 * g.P[a]) + 1;
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
                const trip<T> v = U3[(size_t)pb_grid_offset<T>(0, jx, g.n[1], g.n[2]) + pb_grid_offset<T>(1, jy, g.n[1], g.n[2]) +
                                     pb_grid_offset<T>(2, jz, g.n[1], g.n[2])];
                a0 += wt * (double)v.x;
                a1 += wt * (double)v.y;
                a2 += wt * (double)v.z;
            }
        }
    }
    const double sc = g.scale * g.vol;
    const size_t o = 3 * (size_t)tperm[i];
    if (accumulate == 2) {
        atomicAdd(&u[o], sc * a0);
        atomicAdd(&u[o + 1], sc * a1);
        atomicAdd(&u[o + 2], sc * a2);
    } else if (accumulate) {
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
__global__ __launch_bounds__(256) void k_interp_any(int nt, const double4 *__restrict__ tpos,
                                                    const int *__restrict__ tperm, pb_kgrid g,
                                                    const trip<T> *__restrict__ U3,
                                                    double *__restrict__ u, int accumulate)
{
    interp_any_target<T>((int)(blockIdx.x * blockDim.x + threadIdx.x), nt, tpos, tperm, g, U3, u, accumulate);
}

template <typename T>
__global__ __launch_bounds__(256) void k_ginterp_any(const pb_sys *__restrict__ sys, const int2 *__restrict__ map)
{
    const int2 m = pb_map_at(map, blockIdx.x);
    if (m.x < 0) return;
    const auto &S = ((pb_const_sys)sys)[m.x];
    interp_any_target<T>(m.y * 256 + (int)threadIdx.x, S.nt, S.tpos, S.tperm, S.g, (const trip<T> *)S.grid4, S.u,
                         S.acc_mode);
}

template <typename T>
void launch_interp(hipStream_t s, int nt, const pb_kgrid &g, const void *tpos, const void *tperm,
                   const trip<T> *U3, double *u, int acc)
{
    const bool same = g.P[0] == g.P[1] && g.P[1] == g.P[2];
    const int tpb = kInterpWaves * (same ? pb_interp_tw(g.P[0]) : 1); // targets per block
    const int nblk8 = ((nt + tpb - 1) / tpb + 7) & ~7;
    dim3 gd(nblk8), bd(64 * kInterpWaves);
    dim3 gd1((nt + 255) / 256), bd1(256);
    const double4 *tp = (const double4 *)tpos;
    const int *pm = (const int *)tperm;
#define PB_INTERP(PP) hipLaunchKernelGGL((k_interp<T, PP>), gd, bd, 0, s, nt, nblk8, tp, pm, g, U3, u, acc)
    switch (same ? g.P[0] : -1) {
    case 4: PB_INTERP(4); break;
    case 6: PB_INTERP(6); break;
    case 8: PB_INTERP(8); break;
    case 10: PB_INTERP(10); break;
    case 12: PB_INTERP(12); break;
    case 16: PB_INTERP(16); break;
    case 20: PB_INTERP(20); break;
    case 24: PB_INTERP(24); break;
    default: hipLaunchKernelGGL(k_interp_any<T>, gd1, bd1, 0, s, nt, tp, pm, g, U3, u, acc); break;
    }
#undef PB_INTERP
}

// ---- FFT ---------------------------------------------------------------------
// Pencils of length N held in LDS (pencil stride N + 1 against bank
// conflicts) are transformed by Stockham stages of radix 8, 4, 2 (largest
// first).  A stage reads all its butterfly inputs into registers, waits for
// the workgroup, and writes the outputs in place, so one LDS buffer suffices.
// Element budget per workgroup: E complex values (FP32 per pass, measured on
// MI355X at 128^3: z passes 2048, y 4096, x with the multiply 3072; FP64 2048
// throughout).  A single system's FP32 passes on axes of <= 128 points take
// half (z 1024, y 2048, x 1536): they run beside the pair kernel, and more,
// smaller workgroups interleave with its (the headline step 0.479 -> 0.471
// ms); the grouped passes and longer axes keep the full budget (the 64 sets
// of the multi-RHS call and the 256^3 clustered system measured slower with
// the half).
template <typename T, int PASS> struct pbudget { static constexpr int E = 2048; };
template <> struct pbudget<float, 0> { static constexpr int E = 2048; };
template <> struct pbudget<float, 1> { static constexpr int E = 4096; };
template <> struct pbudget<float, 2> { static constexpr int E = 3072; };
template <typename T, int PASS, int N> struct sbudget {
    static constexpr int E = sizeof(T) == 4 && N <= 128 ? pbudget<T, PASS>::E / 2 : pbudget<T, PASS>::E;
};

template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b)
{
    C r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename C> __device__ __forceinline__ C csub(C a, C b)
{
    C r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
// multiply by -i (forward) or +i (inverse)
template <typename C> __device__ __forceinline__ C mul_mi(C a, bool inv)
{
    C r;
    if (inv) {
        r.x = -a.y;
        r.y = a.x;
    } else {
        r.x = a.y;
        r.y = -a.x;
    }
    return r;
}

// In-register DFT of length R (2, 4, 8): V[q] = sum_m v[m] e^{-+2 pi i m q / R}.
template <typename C, int R> __device__ __forceinline__ void dft(C *v, bool inv)
{
    using T = decltype(v[0].x);
    if constexpr (R == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const C s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]), inv);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else {
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft<C, 4>(e, inv);
        dft<C, 4>(o, inv);
        const T h = (T)0.70710678118654752440084436210484904;
        C w1, w3;
        // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
        w1.x = h * (o[1].x + (inv ? -o[1].y : o[1].y));
        w1.y = h * (o[1].y + (inv ? o[1].x : -o[1].x));
        const C w2 = mul_mi(o[2], inv);
        w3.x = h * (-o[3].x + (inv ? -o[3].y : o[3].y));
        w3.y = h * (-o[3].y + (inv ? o[3].x : -o[3].x));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1);
        v[5] = csub(e[1], w1);
        v[2] = cadd(e[2], w2);
        v[6] = csub(e[2], w2);
        v[3] = cadd(e[3], w3);
        v[7] = csub(e[3], w3);
    }
}

// One Stockham stage of radix R after NS points are combined, for the E / N
// pencils of the workgroup.  tw[q] = e^{-2 pi i q / N} (LDS).
template <typename C, int N, int R, int NS, int E>
__device__ __forceinline__ void fft_stage(C *buf, bool inv, const C *tw)
{
    constexpr int NP = N + 1;
    constexpr int NB = E / R;
    constexpr int KB = (NB + 255) / 256;
    C v[KB][R];
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
        const int b = threadIdx.x + 256 * kb;
        if (NB % 256 == 0 || b < NB) {
            const int p = b / (N / R), t = b % (N / R);
#pragma unroll
            for (int m = 0; m < R; m++) v[kb][m] = buf[p * NP + t + m * (N / R)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
        const int b = threadIdx.x + 256 * kb;
        if (NB % 256 == 0 || b < NB) {
            const int p = b / (N / R), t = b % (N / R);
            const int j = t % NS;
            if (NS > 1) {
#pragma unroll
                for (int m = 1; m < R; m++) {
                    C w = tw[(m * j * (N / (NS * R))) % N];
                    if (inv) w.y = -w.y;
                    v[kb][m] = cmul(v[kb][m], w);
                }
            }
            dft<C, R>(v[kb], inv);
            const int o = p * NP + (t / NS) * NS * R + j;
#pragma unroll
            for (int q = 0; q < R; q++) buf[o + q * NS] = v[kb][q];
        }
    }
    __syncthreads();
}

template <typename C, int N, int NS, int E>
__device__ __forceinline__ void fft_stages(C *buf, bool inv, const C *tw)
{
    if constexpr (NS < N) {
        constexpr int REM = N / NS;
        constexpr int R = (REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2);
        fft_stage<C, N, R, NS, E>(buf, inv, tw);
        fft_stages<C, N, NS * R, E>(buf, inv, tw);
    }
}

template <typename C, int N>
__device__ __forceinline__ void load_twiddles(C *tw, const C *__restrict__ gtw)
{
    for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = gtw[k];
}

// Forward z pass, real to complex: pencils a = (i0, i1) and b = (i0, i1 + 1)
// of the real grid R (T4: fx, fy, fz) are transformed as the three complex
// pencils fx_a + i fy_a, fx_b + i fy_b, fz_a + i fz_b; the half spectra
// (kz = 0 .. N/2) of the six real pencils are written to S.
template <typename T, int N, int E_ = pbudget<T, 0>::E>
__device__ __forceinline__ void fft_z_r2c(const int blk, const typename cplx<T>::type4 *__restrict__ R4,
                                          typename cplx<T>::type *__restrict__ S, int n0, int n1,
                                          const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    constexpr int NP = N + 1, H = N / 2 + 1;
    constexpr int PP = E_ / (3 * N) > 0 ? E_ / (3 * N) : 1;
    constexpr int E = 3 * PP * N;
    __shared__ C buf[3 * PP * NP];
    __shared__ C tw[N];
    load_twiddles<C, N>(tw, gtw);
    const int npair = n0 * (n1 / 2);
    const int pair0 = blk * PP;
    for (int idx = threadIdx.x; idx < 2 * PP * N; idx += 256) {
        const int pp = idx / (2 * N), r = idx % (2 * N), side = r / N, z = r % N;
        const int pr = pair0 + pp;
        typename cplx<T>::type4 f;
        if (pr < npair) {
            const size_t pos = (size_t)(2 * pr + side) * N + z; // pencil 2 pr + side = (i0, i1)
            f = R4[pos];
        } else {
            f.x = f.y = f.z = f.w = (T)0;
        }
        C a;
        a.x = f.x;
        a.y = f.y;
        buf[(3 * pp + side) * NP + z] = a;
        if (side == 0)
            buf[(3 * pp + 2) * NP + z].x = f.z;
        else
            buf[(3 * pp + 2) * NP + z].y = f.z;
    }
    __syncthreads();
    fft_stages<C, N, 1, E>(buf, false, tw);
    const size_t fs = (size_t)n0 * n1 * H;
    for (int idx = threadIdx.x; idx < PP * 2 * H; idx += 256) {
        const int pp = idx / (2 * H), r = idx % (2 * H), side = r / H, k = r % H;
        const int pr = pair0 + pp;
        if (pr >= npair) continue;
        const int km = (N - k) & (N - 1);
        const C z1 = buf[(3 * pp + side) * NP + k], z1m = buf[(3 * pp + side) * NP + km];
        const C z3 = buf[(3 * pp + 2) * NP + k], z3m = buf[(3 * pp + 2) * NP + km];
        // X = (Z(k) + conj Z(-k)) / 2, Y = (Z(k) - conj Z(-k)) / (2i)
        C fx, fy, fz;
        fx.x = (T)0.5 * (z1.x + z1m.x);
        fx.y = (T)0.5 * (z1.y - z1m.y);
        fy.x = (T)0.5 * (z1.y + z1m.y);
        fy.y = (T)0.5 * (z1m.x - z1.x);
        if (side == 0) {
            fz.x = (T)0.5 * (z3.x + z3m.x);
            fz.y = (T)0.5 * (z3.y - z3m.y);
        } else {
            fz.x = (T)0.5 * (z3.y + z3m.y);
            fz.y = (T)0.5 * (z3m.x - z3.x);
        }
        const size_t o = (size_t)(2 * pr + side) * H + k;
        S[o] = fx;
        S[fs + o] = fy;
        S[2 * fs + o] = fz;
    }
}

// y pass (forward or inverse) on S: pencils along i1 of length N = n1 for
// every (field, i0, kz), B consecutive (field, i0, kz) per workgroup.
template <typename T, int N, int E_ = pbudget<T, 1>::E>
__device__ __forceinline__ void fft_y(const int blk, typename cplx<T>::type *__restrict__ S, int n0, int h2,
                                      int inverse, const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    constexpr int NP = N + 1;
    constexpr int B = E_ / N > 0 ? E_ / N : 1;
    constexpr int E = B * N;
    __shared__ C buf[B * NP];
    __shared__ C tw[N];
    load_twiddles<C, N>(tw, gtw);
    const int npen = 3 * n0 * h2;
    const int q0 = blk * B;
    // pencil q = (f i0) * h2 + kz; element i1 at ((f i0) * N + i1) * h2 + kz.
    // Element idx of the workgroup is (i1, g) = (idx / B, idx % B): a thread
    // meets G distinct pencils g (one when B divides 256), whose offsets it
    // computes once.
    static_assert(E % 256 == 0 && (B % 256 == 0 || 256 % B == 0), "y-pass workgroup shape");
    constexpr int G = B >= 256 ? B / 256 : 1;
    size_t base[G];
    bool ok[G];
#pragma unroll
    for (int c = 0; c < G; c++) {
        const int g = B >= 256 ? (int)threadIdx.x + 256 * c : (int)threadIdx.x % B;
        const int q = q0 + g;
        ok[c] = q < npen;
        const int fi = q / h2, kz = q - fi * h2;
        base[c] = (size_t)fi * N * h2 + kz;
    }
#pragma unroll
    for (int m = 0; m < E / 256; m++) {
        const int idx = (int)threadIdx.x + 256 * m, i1 = idx / B, g = idx % B, c = B >= 256 ? m % G : 0;
        C v;
        v.x = v.y = (T)0;
        if (ok[c]) v = S[base[c] + (size_t)i1 * h2];
        buf[g * NP + i1] = v;
    }
    __syncthreads();
    fft_stages<C, N, 1, E>(buf, inverse != 0, tw);
#pragma unroll
    for (int m = 0; m < E / 256; m++) {
        const int idx = (int)threadIdx.x + 256 * m, i1 = idx / B, g = idx % B, c = B >= 256 ? m % G : 0;
        if (ok[c]) S[base[c] + (size_t)i1 * h2] = buf[g * NP + i1];
    }
}

using kmult = pb_kmult;

// x pass with the Brinkman projection between the forward and inverse
// transforms: pencils along i0 of the three fields for B consecutive
// (i1, kz).  Each mode gets
//   U = b(k) (|k|^2 F - (k (k.F) + k' (k'.F)) / 2),
//   b = (1 + c |k|^2) e^{-beta |k|^2} / (|k|^2 (|k|^2 + a^2) n0 n1 n2),
// with Nyquist frequencies taken as -n/2 and k' = k with its Nyquist
// components negated: averaging over their sign keeps the spectrum Hermitian,
// which equals the real part of the full complex transform (the
// reference's convention).  k = 0 is set to zero.
template <typename T, int N, int E_ = pbudget<T, 2>::E>
__device__ __forceinline__ void fft_x_mult(const int blk, typename cplx<T>::type *__restrict__ S, int n1, int h2,
                                           const kmult &km, const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    constexpr int NP = N + 1;
    constexpr int B = E_ / (3 * N) > 0 ? E_ / (3 * N) : 1;
    constexpr int E = 3 * B * N;
    __shared__ C buf[3 * B * NP];
    __shared__ C tw[N];
    load_twiddles<C, N>(tw, gtw);
    const long long npen = (long long)n1 * h2;
    const long long q0 = (long long)blk * B;
    const size_t fs = (size_t)N * npen;
    // element (field f, i0) of pencil q = i1 * h2 + kz at (f * N + i0) * npen + q
    for (int idx = threadIdx.x; idx < E; idx += 256) {
        const int f = idx / (B * N), r = idx % (B * N), i0 = r / B, g = r % B;
        const long long q = q0 + g;
        C v;
        v.x = v.y = (T)0;
        if (q < npen) v = S[f * fs + (size_t)i0 * npen + q];
        buf[(3 * g + f) * NP + i0] = v;
    }
    __syncthreads();
    fft_stages<C, N, 1, E>(buf, false, tw);
    for (int idx = threadIdx.x; idx < B * N; idx += 256) {
        const int g = idx / N, i0 = idx % N;
        const long long q = q0 + g;
        if (q >= npen) continue;
        const int i1 = (int)q / h2, kz = (int)q - i1 * h2;
        const int n2 = km.n[2];
        const int f0 = (2 * i0 >= N) ? i0 - N : i0;
        const int f1 = (2 * i1 >= n1) ? i1 - n1 : i1;
        const int f2 = (2 * kz >= n2) ? kz - n2 : kz;
        const double kx = km.k[0] * f0, ky = km.k[1] * f1, kz_ = km.k[2] * f2;
        const double kn = kx * kx + ky * ky + kz_ * kz_;
        C *F0 = &buf[(3 * g) * NP + i0], *F1 = F0 + NP, *F2 = F1 + NP;
        if (kn == 0.0) {
            C z;
            z.x = z.y = (T)0;
            *F0 = z;
            *F1 = z;
            *F2 = z;
            continue;
        }
        const double b = (1.0 + km.c * kn) * exp(-km.beta * kn) * km.inv_ng / (kn * (kn + km.a2));
        const T tx = (T)kx, ty = (T)ky, tz = (T)kz_;
        const T sx = (2 * i0 == N) ? -tx : tx, sy = (2 * i1 == n1) ? -ty : ty, sz = (2 * kz == n2) ? -tz : tz;
        const C v0 = *F0, v1 = *F1, v2 = *F2;
        const T kr = tx * v0.x + ty * v1.x + tz * v2.x, ki = tx * v0.y + ty * v1.y + tz * v2.y;
        const T sr = sx * v0.x + sy * v1.x + sz * v2.x, si = sx * v0.y + sy * v1.y + sz * v2.y;
        const T bk = (T)b, tkn = (T)kn, hf = (T)0.5;
        C o0, o1, o2;
        o0.x = bk * (tkn * v0.x - hf * (tx * kr + sx * sr));
        o0.y = bk * (tkn * v0.y - hf * (tx * ki + sx * si));
        o1.x = bk * (tkn * v1.x - hf * (ty * kr + sy * sr));
        o1.y = bk * (tkn * v1.y - hf * (ty * ki + sy * si));
        o2.x = bk * (tkn * v2.x - hf * (tz * kr + sz * sr));
        o2.y = bk * (tkn * v2.y - hf * (tz * ki + sz * si));
        *F0 = o0;
        *F1 = o1;
        *F2 = o2;
    }
    __syncthreads();
    fft_stages<C, N, 1, E>(buf, true, tw);
    for (int idx = threadIdx.x; idx < E; idx += 256) {
        const int f = idx / (B * N), r = idx % (B * N), i0 = r / B, g = r % B;
        const long long q = q0 + g;
        if (q < npen) S[f * fs + (size_t)i0 * npen + q] = buf[(3 * g + f) * NP + i0];
    }
}

// Inverse z pass, complex to real: the half spectra of pencils a = (i0, i1)
// and b = (i0, i1 + 1) are extended by U(N - k) = conj U(k) and transformed
// as ux_a + i uy_a, ux_b + i uy_b, uz_a + i uz_b into the real grid (T4).
template <typename T, int N, int E_ = pbudget<T, 0>::E>
__device__ __forceinline__ void fft_z_c2r(const int blk, const typename cplx<T>::type *__restrict__ S,
                                          trip<T> *__restrict__ U3, int n0, int n1,
                                          const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    constexpr int NP = N + 1, H = N / 2 + 1;
    constexpr int PP = E_ / (3 * N) > 0 ? E_ / (3 * N) : 1;
    constexpr int E = 3 * PP * N;
    __shared__ C buf[3 * PP * NP];
    __shared__ C tw[N];
    load_twiddles<C, N>(tw, gtw);
    const int npair = n0 * (n1 / 2);
    const int pair0 = blk * PP;
    const size_t fs = (size_t)n0 * n1 * H;
    for (int idx = threadIdx.x; idx < PP * 2 * N; idx += 256) {
        const int pp = idx / (2 * N), r = idx % (2 * N), side = r / N, k = r % N;
        const int pr = pair0 + pp;
        C ux, uy, uz;
        ux.x = ux.y = uy.x = uy.y = uz.x = uz.y = (T)0;
        if (pr < npair) {
            const bool up = k >= H;
            const int kk = up ? N - k : k;
            const size_t o = (size_t)(2 * pr + side) * H + kk;
            ux = S[o];
            uy = S[fs + o];
            uz = S[2 * fs + o];
            if (up) {
                ux.y = -ux.y;
                uy.y = -uy.y;
                uz.y = -uz.y;
            }
        }
        C z;
        z.x = ux.x - uy.y; // ux + i uy
        z.y = ux.y + uy.x;
        buf[(3 * pp + side) * NP + k] = z;
        if (side == 0) { // uz_a enters the real part of the third pencil
            buf[(3 * pp + 2) * NP + k].x = uz.x;
            buf[(3 * pp + 2) * NP + k].y = uz.y;
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < PP * N; idx += 256) { // + i uz_b
        const int pp = idx / N, k = idx % N;
        const int pr = pair0 + pp;
        if (pr >= npair) continue;
        const bool up = k >= H;
        const int kk = up ? N - k : k;
        C uz = S[2 * fs + (size_t)(2 * pr + 1) * H + kk];
        if (up) uz.y = -uz.y;
        buf[(3 * pp + 2) * NP + k].x -= uz.y;
        buf[(3 * pp + 2) * NP + k].y += uz.x;
    }
    __syncthreads();
    fft_stages<C, N, 1, E>(buf, true, tw);
    for (int idx = threadIdx.x; idx < 2 * PP * N; idx += 256) {
        const int pp = idx / (2 * N), r = idx % (2 * N), side = r / N, z = r % N;
        const int pr = pair0 + pp;
        if (pr >= npair) continue;
        const C a = buf[(3 * pp + side) * NP + z], c3 = buf[(3 * pp + 2) * NP + z];
        const size_t o = sizeof(T) == 4 ? ((size_t)pr * N + z) * 2 + side : (size_t)(2 * pr + side) * N + z;
        U3[o] = trip<T>{a.x, a.y, side == 0 ? c3.x : c3.y}; // the layout of pb_grid_offset
    }
}

// ---- fused (y, z) plane passes ----------------------------------------------
// Where a whole x-plane of the three half spectra fits in LDS (N x N planes,
// N <= kPlaneMax<T>), the z and y transforms of a plane run in one workgroup:
// the forward pass reads the real plane once and writes its spectrum once, the
// inverse pass the reverse, so the transform makes three trips through HBM
// (plane, x with the multiply, plane) instead of five.  The arithmetic is that
// of the separate passes: z pencils two at a time as one complex transform,
// the half spectra (kz = 0 .. N/2) transposed into y pencils in LDS.
template <typename T> struct plane_max { static constexpr int N = 64; };
template <> struct plane_max<double> { static constexpr int N = 32; };

template <typename T, int N> struct plane_lds {
    static constexpr int NP = N + 1, H = N / 2 + 1;
    typename cplx<T>::type buf[3 * H * NP]; // z pencils [3 (N/2)][NP], then y pencils [3 H][NP]
    typename cplx<T>::type tw[N];
};

template <typename T, int N>
__device__ __forceinline__ void fft_yz_fwd(const int i0, const typename cplx<T>::type4 *__restrict__ R4,
                                           typename cplx<T>::type *__restrict__ S, int n0,
                                           const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    using C4 = typename cplx<T>::type4;
    constexpr int NP = N + 1, H = N / 2 + 1;
    __shared__ plane_lds<T, N> L;
    C *buf = L.buf;
    load_twiddles<C, N>(L.tw, gtw);
    const C4 *P4 = R4 + (size_t)i0 * N * N;
    for (int idx = threadIdx.x; idx < N * N; idx += 256) {
        const int i1 = idx / N, z = idx % N, pp = i1 >> 1, side = i1 & 1;
        const C4 f = P4[idx];
        C a;
        a.x = f.x;
        a.y = f.y;
        buf[(3 * pp + side) * NP + z] = a;
        if (side == 0)
            buf[(3 * pp + 2) * NP + z].x = f.z;
        else
            buf[(3 * pp + 2) * NP + z].y = f.z;
    }
    __syncthreads();
    fft_stages<C, N, 1, 3 * (N / 2) * N>(buf, false, L.tw);
    // half spectra of pencil i1 at kz = k, moved to y pencil (f, k) in place
    constexpr int NI = N * H, KI = (NI + 255) / 256;
    C o[KI][3];
#pragma unroll
    for (int kb = 0; kb < KI; kb++) {
        const int idx = (int)threadIdx.x + 256 * kb;
        if (NI % 256 == 0 || idx < NI) {
            const int i1 = idx / H, k = idx % H, pp = i1 >> 1, side = i1 & 1, km = (N - k) & (N - 1);
            const C z1 = buf[(3 * pp + side) * NP + k], z1m = buf[(3 * pp + side) * NP + km];
            const C z3 = buf[(3 * pp + 2) * NP + k], z3m = buf[(3 * pp + 2) * NP + km];
            o[kb][0].x = (T)0.5 * (z1.x + z1m.x);
            o[kb][0].y = (T)0.5 * (z1.y - z1m.y);
            o[kb][1].x = (T)0.5 * (z1.y + z1m.y);
            o[kb][1].y = (T)0.5 * (z1m.x - z1.x);
            if (side == 0) {
                o[kb][2].x = (T)0.5 * (z3.x + z3m.x);
                o[kb][2].y = (T)0.5 * (z3.y - z3m.y);
            } else {
                o[kb][2].x = (T)0.5 * (z3.y + z3m.y);
                o[kb][2].y = (T)0.5 * (z3m.x - z3.x);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KI; kb++) {
        const int idx = (int)threadIdx.x + 256 * kb;
        if (NI % 256 == 0 || idx < NI) {
            const int i1 = idx / H, k = idx % H;
#pragma unroll
            for (int f = 0; f < 3; f++) buf[(f * H + k) * NP + i1] = o[kb][f];
        }
    }
    __syncthreads();
    fft_stages<C, N, 1, 3 * H * N>(buf, false, L.tw);
    const size_t fs = (size_t)n0 * N * H;
    for (int idx = threadIdx.x; idx < 3 * N * H; idx += 256) {
        const int f = idx / (N * H), r = idx % (N * H), i1 = r / H, kz = r % H;
        S[f * fs + ((size_t)i0 * N + i1) * H + kz] = buf[(f * H + kz) * NP + i1];
    }
}

template <typename T, int N>
__device__ __forceinline__ void fft_yz_inv(const int i0, const typename cplx<T>::type *__restrict__ S,
                                           trip<T> *__restrict__ U3, int n0,
                                           const typename cplx<T>::type *__restrict__ gtw)
{
    using C = typename cplx<T>::type;
    constexpr int NP = N + 1, H = N / 2 + 1;
    __shared__ plane_lds<T, N> L;
    C *buf = L.buf;
    load_twiddles<C, N>(L.tw, gtw);
    const size_t fs = (size_t)n0 * N * H;
    for (int idx = threadIdx.x; idx < 3 * N * H; idx += 256) {
        const int f = idx / (N * H), r = idx % (N * H), i1 = r / H, kz = r % H;
        buf[(f * H + kz) * NP + i1] = S[f * fs + ((size_t)i0 * N + i1) * H + kz];
    }
    __syncthreads();
    fft_stages<C, N, 1, 3 * H * N>(buf, true, L.tw);
    // z pencils of the pair (2 pp, 2 pp + 1) at k: U(N - k) = conj U(k)
    constexpr int NI = (N / 2) * N, KI = (NI + 255) / 256;
    C o[KI][3];
#pragma unroll
    for (int kb = 0; kb < KI; kb++) {
        const int idx = (int)threadIdx.x + 256 * kb;
        if (NI % 256 == 0 || idx < NI) {
            const int pp = idx / N, k = idx % N;
            const bool up = k >= H;
            const int kk = up ? N - k : k;
            C v[3][2];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int sd = 0; sd < 2; sd++) {
                    v[f][sd] = buf[(f * H + kk) * NP + 2 * pp + sd];
                    if (up) v[f][sd].y = -v[f][sd].y;
                }
#pragma unroll
            for (int sd = 0; sd < 2; sd++) { // ux + i uy
                o[kb][sd].x = v[0][sd].x - v[1][sd].y;
                o[kb][sd].y = v[0][sd].y + v[1][sd].x;
            }
            o[kb][2].x = v[2][0].x - v[2][1].y; // uz_a + i uz_b
            o[kb][2].y = v[2][0].y + v[2][1].x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KI; kb++) {
        const int idx = (int)threadIdx.x + 256 * kb;
        if (NI % 256 == 0 || idx < NI) {
            const int pp = idx / N, k = idx % N;
#pragma unroll
            for (int q = 0; q < 3; q++) buf[(3 * pp + q) * NP + k] = o[kb][q];
        }
    }
    __syncthreads();
    fft_stages<C, N, 1, 3 * (N / 2) * N>(buf, true, L.tw);
    trip<T> *P3 = U3 + (size_t)i0 * N * N;
    for (int idx = threadIdx.x; idx < N * N; idx += 256) {
        const int i1 = idx / N, z = idx % N, pp = i1 >> 1, side = i1 & 1;
        const C a = buf[(3 * pp + side) * NP + z], c3 = buf[(3 * pp + 2) * NP + z];
        P3[sizeof(T) == 4 ? (pp * N + z) * 2 + side : idx] = trip<T>{a.x, a.y, side == 0 ? c3.x : c3.y};
    }
}

template <typename T, int N, bool INV>
__global__ __launch_bounds__(256) void k_fft_yz(const typename cplx<T>::type4 *__restrict__ R4,
                                                typename cplx<T>::type *__restrict__ S, int n0,
                                                const typename cplx<T>::type *__restrict__ gtw)
{
    if constexpr (N <= plane_max<T>::N) {
        if (INV)
            fft_yz_inv<T, N>((int)blockIdx.x, S, (trip<T> *)R4, n0, gtw);
        else
            fft_yz_fwd<T, N>((int)blockIdx.x, R4, S, n0, gtw);
    }
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_fft_z_r2c(const typename cplx<T>::type4 *__restrict__ R4,
                                                   typename cplx<T>::type *__restrict__ S, int n0,
                                                   int n1, const typename cplx<T>::type *__restrict__ gtw)
{
    fft_z_r2c<T, N, sbudget<T, 0, N>::E>((int)blockIdx.x, R4, S, n0, n1, gtw);
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_fft_y(typename cplx<T>::type *__restrict__ S, int n0, int h2,
                                               int inverse, const typename cplx<T>::type *__restrict__ gtw)
{
    fft_y<T, N, sbudget<T, 1, N>::E>((int)blockIdx.x, S, n0, h2, inverse, gtw);
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_fft_x_mult(typename cplx<T>::type *__restrict__ S, int n1, int h2,
                                                    kmult km, const typename cplx<T>::type *__restrict__ gtw)
{
    fft_x_mult<T, N, sbudget<T, 2, N>::E>((int)blockIdx.x, S, n1, h2, km, gtw);
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_fft_z_c2r(const typename cplx<T>::type *__restrict__ S,
                                                   trip<T> *__restrict__ U3, int n0,
                                                   int n1, const typename cplx<T>::type *__restrict__ gtw)
{
    fft_z_c2r<T, N, sbudget<T, 0, N>::E>((int)blockIdx.x, S, U3, n0, n1, gtw);
}

// Grouped FFT pass PASS (0 z forward, 1 y forward, 2 x with the multiply,
// 3 y inverse, 4 z inverse) for the systems whose transformed axis has
// length N; tw: the twiddles of length N.
template <typename T, int N, int PASS>
__global__ __launch_bounds__(256) void k_gfft(const pb_sys *__restrict__ sys, const int2 *__restrict__ map,
                                              const typename cplx<T>::type *__restrict__ tw)
{
    using C = typename cplx<T>::type;
    const int2 m = pb_map_at(map, blockIdx.x);
    const auto &S = ((pb_const_sys)sys)[m.x];
    const int h2 = S.g.n[2] / 2 + 1;
    if constexpr (PASS == 5 || PASS == 6) {
        if constexpr (N <= plane_max<T>::N) {
            if (PASS == 5)
                fft_yz_fwd<T, N>(m.y, (const typename cplx<T>::type4 *)S.grid4, (C *)S.spec, S.g.n[0], tw);
            else
                fft_yz_inv<T, N>(m.y, (const C *)S.spec, (trip<T> *)S.grid4, S.g.n[0], tw);
        }
    } else if constexpr (PASS == 0)
        fft_z_r2c<T, N>(m.y, (const typename cplx<T>::type4 *)S.grid4, (C *)S.spec, S.g.n[0], S.g.n[1], tw);
    else if constexpr (PASS == 1 || PASS == 3)
        fft_y<T, N>(m.y, (C *)S.spec, S.g.n[0], h2, PASS == 3, tw);
    else if constexpr (PASS == 2) {
        const kmult km = S.km;
        fft_x_mult<T, N>(m.y, (C *)S.spec, S.g.n[1], h2, km, tw);
    } else
        fft_z_c2r<T, N>(m.y, (const C *)S.spec, (trip<T> *)S.grid4, S.g.n[0], S.g.n[1], tw);
}

template <typename T, int N>
void launch_axis(hipStream_t s, int pass, const pb_kgrid &g, void *grid4, void *spec, const kmult &km,
                 const void *tw)
{
    using C = typename cplx<T>::type;
    using C4 = typename cplx<T>::type4;
    const int h2 = g.n[2] / 2 + 1;
    const C *t = (const C *)tw;
    if (pass == 5 || pass == 6) { // (y, z) planes, one workgroup per x-plane
        if constexpr (N <= plane_max<T>::N) {
            if (pass == 5)
                hipLaunchKernelGGL((k_fft_yz<T, N, false>), dim3(g.n[0]), dim3(256), 0, s, (const C4 *)grid4,
                                   (C *)spec, g.n[0], t);
            else
                hipLaunchKernelGGL((k_fft_yz<T, N, true>), dim3(g.n[0]), dim3(256), 0, s, (const C4 *)grid4,
                                   (C *)spec, g.n[0], t);
        }
    } else if (pass == 0 || pass == 4) { // z
        constexpr int PP = sbudget<T, 0, N>::E / (3 * N) > 0 ? sbudget<T, 0, N>::E / (3 * N) : 1;
        const int npair = g.n[0] * (g.n[1] / 2);
        dim3 gd((npair + PP - 1) / PP);
        if (pass == 0)
            hipLaunchKernelGGL((k_fft_z_r2c<T, N>), gd, dim3(256), 0, s, (const C4 *)grid4, (C *)spec,
                               g.n[0], g.n[1], t);
        else
            hipLaunchKernelGGL((k_fft_z_c2r<T, N>), gd, dim3(256), 0, s, (const C *)spec, (trip<T> *)grid4,
                               g.n[0], g.n[1], t);
    } else if (pass == 1 || pass == 3) { // y
        constexpr int B = sbudget<T, 1, N>::E / N > 0 ? sbudget<T, 1, N>::E / N : 1;
        const long long npen = 3LL * g.n[0] * h2;
        hipLaunchKernelGGL((k_fft_y<T, N>), dim3((unsigned)((npen + B - 1) / B)), dim3(256), 0, s, (C *)spec,
                           g.n[0], h2, pass == 3 ? 1 : 0, t);
    } else { // x with the multiply
        constexpr int B = sbudget<T, 2, N>::E / (3 * N) > 0 ? sbudget<T, 2, N>::E / (3 * N) : 1;
        const long long npen = (long long)g.n[1] * h2;
        hipLaunchKernelGGL((k_fft_x_mult<T, N>), dim3((unsigned)((npen + B - 1) / B)), dim3(256), 0, s,
                           (C *)spec, g.n[1], h2, km, t);
    }
}

// The passes of a transform in order: 0 z forward (r2c), 1 y forward, 2 x
// forward + multiply + inverse, 3 y inverse, 4 z inverse (c2r); or, for
// n1 = n2 planes that fit in LDS, 5 (y, z) forward, 2, 6 (y, z) inverse.
template <typename T>
int fft_passes_t(const pb_kgrid &g, int *passes)
{
    if (g.n[1] == g.n[2] && g.n[2] <= plane_max<T>::N) {
        passes[0] = 5;
        passes[1] = 2;
        passes[2] = 6;
        return 3;
    }
    for (int k = 0; k < 5; k++) passes[k] = k;
    return 5;
}

template <typename T>
int fft_pass(hipStream_t s, int pass, const pb_kgrid &g, void *grid4, void *spec, const kmult &km,
             const void *const *tws)
{
    const int ax = pb_fft_axis(pass);
    const void *tw = tws[ax];
    switch (g.n[ax]) {
    case 4: launch_axis<T, 4>(s, pass, g, grid4, spec, km, tw); break;
    case 8: launch_axis<T, 8>(s, pass, g, grid4, spec, km, tw); break;
    case 16: launch_axis<T, 16>(s, pass, g, grid4, spec, km, tw); break;
    case 32: launch_axis<T, 32>(s, pass, g, grid4, spec, km, tw); break;
    case 64: launch_axis<T, 64>(s, pass, g, grid4, spec, km, tw); break;
    case 128: launch_axis<T, 128>(s, pass, g, grid4, spec, km, tw); break;
    case 256: launch_axis<T, 256>(s, pass, g, grid4, spec, km, tw); break;
    case 512: launch_axis<T, 512>(s, pass, g, grid4, spec, km, tw); break;
    default: return PB_ERR_UNSUPPORTED;
    }
    return PB_OK;
}

// Twiddles e^{-2 pi i q/n}, q < n, for every supported length n = 4 .. 512
// (table n at offset n - 4), computed in long double and uploaded once per
// precision, so calls with different grids need no further upload.
template <typename T>
int upload_twiddles(pb_context *c, hipStream_t s)
{
    using C = typename cplx<T>::type;
    const int prec = sizeof(T) == 8 ? 0 : 1;
    if (c->twid_ok[prec]) return PB_OK;
    const size_t tot = 1020; // 4 + 8 + ... + 512
    std::vector<C> h(tot);
    for (int n = 4; n <= 512; n *= 2)
        for (int k = 0; k < n; k++) {
            long double ang = -2.0L * 3.141592653589793238462643383279502884L * k / n;
            h[n - 4 + k].x = (T)cosl(ang);
            h[n - 4 + k].y = (T)sinl(ang);
        }
    int st;
    if ((st = pb_reserve(c, &c->twid[prec], tot * sizeof(C)))) return st;
    if ((st = pb_hip_check(c, hipMemcpyAsync(c->twid[prec].ptr, h.data(), tot * sizeof(C),
                                             hipMemcpyHostToDevice, s), "twiddle upload")))
        return st;
    if ((st = pb_hip_check(c, hipStreamSynchronize(s), "twiddle sync"))) return st;
    c->twid_ok[prec] = 1;
    return PB_OK;
}

// Upper bound on the candidate runs a source can appear in over all regions:
// per axis, the largest number of (region, unwrapped cell) pairs that fall on
// one cell (one cell of slack on each side for rounding), multiplied.
long long spread_multiplicity(const pb_cells &cells, const pb_kgrid &g)
{
    long long K = 1;
    for (int a = 0; a < 3; a++) {
        const int nr = (g.n[a] + 7) / 8, nc = cells.nc[a];
        std::vector<int> cnt(nc, 0);
        for (int r = 0; r < nr; r++) {
            const int gA = 8 * r, gB = std::min(gA + 7, g.n[a] - 1);
            const double lo = (gA - 0.5 * g.P[a]) * g.h[a], hi = (gB + 0.5 * g.P[a]) * g.h[a];
            const int ulo = (int)std::floor(lo * cells.inv_h[a] - 1e-9) - 1;
            const int uhi = (int)std::floor(hi * cells.inv_h[a] + 1e-9) + 1;
            for (int u = ulo; u <= uhi; u++) cnt[((u % nc) + nc) % nc]++;
        }
        K *= *std::max_element(cnt.begin(), cnt.end());
    }
    return K;
}

// Spreading of one system.  Large spreading work (ns P^3 >= 1e9 window
// points, e.g. 1e6 sources at P = 16) splits regions of more than C
// candidates into chunks of C (count, items, reduction), C doubled until the
// partial slots any split could need fit 256 MiB (the split regions' chunks
// number at most 2 cand_bound / C).  Smaller work, and the allocation-free
// entry point (whose scratch holds no partials), take one workgroup per
// region: beside the pair kernel, the three extra launches cost some 20-50 us
// (measured at 1e5 sources, 128^3, P 12: 0.668 against 0.618 ms per step),
// which only pays off against long spreading.
template <typename T>
int spread_balanced(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells, int ns,
                    typename cplx<T>::type4 *grid4)
{
    using C4 = typename cplx<T>::type4;
    const int nreg = ((g.n[0] + 7) / 8) * ((g.n[1] + 7) / 8) * ((g.n[2] + 7) / 8);
    const int *cstart = (const int *)c->cell_start.ptr;
    const double4 *spos = (const double4 *)c->spos.ptr, *sfrc = (const double4 *)c->sfrc.ptr;
    if (!pb_spread_split(c, g, ns)) {
        hipLaunchKernelGGL(k_spread_regions<T>, dim3((unsigned)nreg), dim3(64 * kSpreadWaves), 0, s, cstart, spos,
                           sfrc, cells, g, grid4);
        return PB_OK;
    }
    const long long cand_bound = (long long)std::max(ns, 1) * spread_multiplicity(cells, g);
    const size_t slot = 512 * sizeof(C4);
    long long C = 4096;
    while ((2 * (cand_bound / C) + 1) * (long long)slot > (256LL << 20) && C < (1LL << 24)) C *= 2;
    const long long part_cap = 2 * (cand_bound / C) + 1;
    if (part_cap + nreg > INT_MAX / 2)
        return pb_fail(c, PB_ERR_UNSUPPORTED, "too many sources for the spreading plan");
    int st;
    if ((st = pb_reserve(c, &c->sp_int, sizeof(int) * (3 * (size_t)nreg + 4)))) return st;
    if ((st = pb_reserve(c, &c->sp_items, sizeof(int2) * (size_t)part_cap))) return st;
    if ((st = pb_reserve(c, &c->sp_part, slot * (size_t)part_cap))) return st;
    int *cand = (int *)c->sp_int.ptr, *mlist = cand + nreg, *pbase = mlist + nreg, *tot = pbase + nreg;
    int2 *owner = (int2 *)c->sp_items.ptr;
    C4 *part = (C4 *)c->sp_part.ptr;
    static int per_cu = 0; // resident item workgroups per compute unit
    if (!per_cu) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spread_items<T>, 64 * kSpreadWaves, 0) != hipSuccess ||
            nb < 1)
            nb = 2;
        per_cu = nb;
    }
    const int nblk = (int)std::min<long long>(nreg + part_cap, (long long)per_cu * c->cu_count);
    hipLaunchKernelGGL(k_spread_reset, dim3(1), dim3(64), 0, s, tot);
    hipLaunchKernelGGL(k_spread_count, dim3((unsigned)((nreg + 3) / 4)), dim3(256), 0, s, cstart, cells, g, nreg,
                       (int)C, cand, pbase, mlist, owner, (int)part_cap, tot);
    hipLaunchKernelGGL(k_spread_items<T>, dim3((unsigned)nblk), dim3(64 * kSpreadWaves), 0, s, cstart, spos, sfrc,
                       cells, g, grid4, nreg, (int)C, (const int *)cand, (const int2 *)owner, (int)part_cap, part, tot);
    hipLaunchKernelGGL(k_spread_reduce<T>, dim3((unsigned)std::min(nreg, 4 * c->cu_count)), dim3(256), 0, s,
                       (const int *)tot, (const int *)mlist, (const int *)pbase, (const int *)cand, (int)C,
                       (const C4 *)part, cells, g, grid4);
    return PB_OK;
}

template <typename T>
int run_kspace_t(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells,
                 const pb_box &box, double a, double d, int ns, int nt, double *u, int accumulate,
                 const std::function<int(hipStream_t &)> *fork, int fork_at)
{
    using C = typename cplx<T>::type;
    using C4 = typename cplx<T>::type4;
    int st;
    for (int ax = 0; ax < 3; ax++) {
        int n = g.n[ax];
        if (n < 4 || (n & (n - 1)) || n > 512)
            return pb_fail(c, PB_ERR_UNSUPPORTED, "FFT grid sizes must be powers of two in [4, 512]");
        if (g.P[ax] < 1 || g.P[ax] > kPmax || g.P[ax] > n)
            return pb_fail(c, PB_ERR_INVALID, "window points P must be in [1, min(32, grid)]");
    }
    const size_t h2 = (size_t)g.n[2] / 2 + 1;
    // real grid: (fx, fy, fz, 0) per point after the spread, packed
    // (ux, uy, uz) after the inverse transform
    if ((st = pb_reserve(c, &c->grid, (size_t)g.ng * sizeof(C4)))) return st;
    if ((st = pb_reserve(c, &c->spec, 3 * (size_t)g.n[0] * g.n[1] * h2 * sizeof(C)))) return st;
    if ((st = upload_twiddles<T>(c, s))) return st;
    const kmult km = pb_make_kmult(g, box, a, d);
    void *grid4 = c->grid.ptr, *spec = c->spec.ptr;
    // the caller's fork (the pair kernel beside the k space) at point
    // fork_at: 0 before the spread, 1 after it, 2 after the forward
    // transform, 3 before the gather; it may move the rest of the k space to
    // another stream (s)
    auto fork_here = [&](int at) { return fork && fork_at == at ? (*fork)(s) : PB_OK; };
    if ((st = fork_here(0))) return st;
    if ((st = spread_balanced<T>(c, s, g, cells, ns, (C4 *)grid4))) return st;
    if ((st = fork_here(1))) return st;
    const C *tw = (const C *)c->twid[sizeof(T) == 8 ? 0 : 1].ptr;
    const void *tws[3] = {tw + (g.n[0] - 4), tw + (g.n[1] - 4), tw + (g.n[2] - 4)};
    pb_mark(c, s, 2);
    int passes[5];
    const int np = fft_passes_t<T>(g, passes);
    for (int k = 0; k < np; k++) {
        const int pass = passes[k];
        if (pass == 2) {
            if ((st = fork_here(2))) return st;
            pb_mark(c, s, 3);
        }
        if (k > 0 && passes[k - 1] == 2) pb_mark(c, s, 4);
        if ((st = fft_pass<T>(s, pass, g, grid4, spec, km, tws))) return pb_fail(c, st, "FFT size");
    }
    pb_mark(c, s, 5);
    if ((st = fork_here(3))) return st;
    if (nt > 0) launch_interp<T>(s, nt, g, c->tpos.ptr, c->tperm.ptr, (const trip<T> *)grid4, u, accumulate);
    if (c->timing != 2) pb_mark(c, s, 6); // (mode 2: the pair kernel records boundary 6 itself)
    return pb_hip_check(c, hipGetLastError(), "k-space launch");
}

} // namespace

int pb_run_kspace(pb_context *c, hipStream_t s, const pb_kgrid &g, const pb_cells &cells,
                  const pb_box &box, double a, double d, int ns, int nt, double *u, int accumulate,
                  int precision, const std::function<int(hipStream_t &)> *fork, int fork_at)
{
    if (precision == PB_PREC_MIXED)
        return run_kspace_t<float>(c, s, g, cells, box, a, d, ns, nt, u, accumulate, fork, fork_at);
    return run_kspace_t<double>(c, s, g, cells, box, a, d, ns, nt, u, accumulate, fork, fork_at);
}

bool pb_spread_split(const pb_context *c, const pb_kgrid &g, int ns)
{
    return !c->arena && (double)std::max(ns, 0) * g.P[0] * g.P[1] * g.P[2] >= 1e9;
}

pb_kmult pb_make_kmult(const pb_kgrid &g, const pb_box &box, double a, double d)
{
    pb_kmult km;
    for (int ax = 0; ax < 3; ax++) {
        km.k[ax] = 2 * M_PI / box.L[ax];
        km.n[ax] = g.n[ax];
    }
    km.c = -std::expm1(-a * a * d * d / 4) / (a * a);
    km.a2 = a * a;
    km.inv_ng = 1.0 / (double)g.ng;
    km.beta = d * d / 4 - g.eta / 2;
    return km;
}

int pb_twiddles(pb_context *c, hipStream_t s, int precision, const void **tw)
{
    int st = precision == PB_PREC_MIXED ? upload_twiddles<float>(c, s) : upload_twiddles<double>(c, s);
    *tw = c->twid[precision == PB_PREC_MIXED ? 1 : 0].ptr;
    return st;
}

namespace {
template <typename T>
int fft_blocks_t(int pass, const pb_kgrid &g)
{
    const long long h2 = g.n[2] / 2 + 1;
    long long work, per;
    if (pass == 5 || pass == 6) {
        per = 1;
        work = g.n[0];
    } else if (pass == 0 || pass == 4) {
        per = std::max(1, pbudget<T, 0>::E / (3 * g.n[2]));
        work = (long long)g.n[0] * (g.n[1] / 2);
    } else if (pass == 1 || pass == 3) {
        per = std::max(1, pbudget<T, 1>::E / g.n[1]);
        work = 3LL * g.n[0] * h2;
    } else {
        per = std::max(1, pbudget<T, 2>::E / (3 * g.n[0]));
        work = (long long)g.n[1] * h2;
    }
    return (int)((work + per - 1) / per);
}

template <typename T, int N>
void launch_gfft(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int pass, const void *tw)
{
    using C = typename cplx<T>::type;
    const C *t = (const C *)tw + (N - 4); // table of length N
    dim3 gd(nblk), bd(256);
    switch (pass) {
    case 0: hipLaunchKernelGGL((k_gfft<T, N, 0>), gd, bd, 0, s, sys, map, t); break;
    case 1: hipLaunchKernelGGL((k_gfft<T, N, 1>), gd, bd, 0, s, sys, map, t); break;
    case 2: hipLaunchKernelGGL((k_gfft<T, N, 2>), gd, bd, 0, s, sys, map, t); break;
    case 3: hipLaunchKernelGGL((k_gfft<T, N, 3>), gd, bd, 0, s, sys, map, t); break;
    case 4: hipLaunchKernelGGL((k_gfft<T, N, 4>), gd, bd, 0, s, sys, map, t); break;
    default:
        if constexpr (N <= plane_max<T>::N) {
            if (pass == 5) hipLaunchKernelGGL((k_gfft<T, N, 5>), gd, bd, 0, s, sys, map, t);
            else hipLaunchKernelGGL((k_gfft<T, N, 6>), gd, bd, 0, s, sys, map, t);
        }
        break;
    }
}

template <typename T>
int group_fft_t(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int pass, int n, const void *tw)
{
    switch (n) {
    case 4: launch_gfft<T, 4>(s, sys, map, nblk, pass, tw); break;
    case 8: launch_gfft<T, 8>(s, sys, map, nblk, pass, tw); break;
    case 16: launch_gfft<T, 16>(s, sys, map, nblk, pass, tw); break;
    case 32: launch_gfft<T, 32>(s, sys, map, nblk, pass, tw); break;
    case 64: launch_gfft<T, 64>(s, sys, map, nblk, pass, tw); break;
    case 128: launch_gfft<T, 128>(s, sys, map, nblk, pass, tw); break;
    case 256: launch_gfft<T, 256>(s, sys, map, nblk, pass, tw); break;
    case 512: launch_gfft<T, 512>(s, sys, map, nblk, pass, tw); break;
    default: return PB_ERR_UNSUPPORTED;
    }
    return PB_OK;
}

template <typename T>
int group_interp_t(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int P)
{
    dim3 gd(nblk), bd(64 * kInterpWaves);
#define PB_GINTERP(PP) hipLaunchKernelGGL((k_ginterp<T, PP>), gd, bd, 0, s, sys, map, nblk)
    switch (P) {
    case 4: PB_GINTERP(4); break;
    case 6: PB_GINTERP(6); break;
    case 8: PB_GINTERP(8); break;
    case 10: PB_GINTERP(10); break;
    case 12: PB_GINTERP(12); break;
    case 16: PB_GINTERP(16); break;
    case 20: PB_GINTERP(20); break;
    case 24: PB_GINTERP(24); break;
    case 0: hipLaunchKernelGGL(k_ginterp_any<T>, gd, dim3(256), 0, s, sys, map); break;
    default: return PB_ERR_UNSUPPORTED;
    }
#undef PB_GINTERP
    return PB_OK;
}
} // namespace

int pb_fft_axis(int pass)
{
    return pass == 1 || pass == 3 ? 1 : (pass == 2 ? 0 : 2);
}

int pb_fft_passes(int precision, const pb_kgrid &g, int *passes)
{
    return precision == PB_PREC_MIXED ? fft_passes_t<float>(g, passes) : fft_passes_t<double>(g, passes);
}

int pb_group_fft_blocks(int precision, int pass, const pb_kgrid &g)
{
    return precision == PB_PREC_MIXED ? fft_blocks_t<float>(pass, g) : fft_blocks_t<double>(pass, g);
}

void pb_group_spread(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision)
{
    if (nblk <= 0) return;
    if (precision == PB_PREC_MIXED)
        hipLaunchKernelGGL(k_gspread<float>, dim3(nblk), dim3(64 * kSpreadWaves), 0, s, sys, map);
    else
        hipLaunchKernelGGL(k_gspread<double>, dim3(nblk), dim3(64 * kSpreadWaves), 0, s, sys, map);
}

void pb_group_spread_sets(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk)
{
    if (nblk > 0)
        hipLaunchKernelGGL(k_gspread_sets<pb_spread_sets>, dim3(nblk), dim3(64 * kSpreadWaves), 0, s, sys, map);
}

int pb_group_fft(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision, int pass, int n,
                 const void *twiddles)
{
    if (nblk <= 0) return PB_OK;
    return precision == PB_PREC_MIXED ? group_fft_t<float>(s, sys, map, nblk, pass, n, twiddles)
                                      : group_fft_t<double>(s, sys, map, nblk, pass, n, twiddles);
}

int pb_group_interp(hipStream_t s, const pb_sys *sys, const int2 *map, int nblk, int precision, int P)
{
    if (nblk <= 0) return PB_OK;
    return precision == PB_PREC_MIXED ? group_interp_t<float>(s, sys, map, nblk, P)
                                      : group_interp_t<double>(s, sys, map, nblk, P);
}
