// pb_api.cpp -- the C ABI (include/periodicbrinkman.h): contexts, argument
// checks turned into status codes, and the host orchestration of the kernels.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pb_internal.h"

#include <algorithm>
#include <functional>
#include <cstdint>
#include <mutex>
#include <thread>

// ---- context helpers --------------------------------------------------------
int pb_fail(pb_context *c, int status, const std::string &msg)
{
    if (c) c->err = msg;
    return status;
}

int pb_hip_check(pb_context *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return PB_OK;
    return pb_fail(c, e == hipErrorOutOfMemory ? PB_ERR_NOMEM : PB_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

int pb_reserve(pb_context *c, pb_buffer *b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b->bytes >= bytes) return PB_OK;
    if (c->arena) { // a slice of the device scratch; nothing is allocated
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (need > c->arena_end - c->arena_used)
            return pb_fail(c, PB_ERR_UNSUPPORTED, "the workload exceeds the no-allocation scratch");
        b->ptr = c->arena + c->arena_used;
        b->bytes = need;
        c->arena_used += need;
        return PB_OK;
    }
    if (b->ptr) {
        hipError_t e = hipFree(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
        if (e != hipSuccess) return pb_hip_check(c, e, "hipFree");
    }
    size_t want = bytes + bytes / 8; // headroom for slowly growing sizes
    hipError_t e = hipMalloc(&b->ptr, want);
    if (e != hipSuccess) {
        b->ptr = nullptr;
        return pb_hip_check(c, e, "hipMalloc");
    }
    b->bytes = want;
    return PB_OK;
}

// Each recorded event drains the queue (about 6 us on the device timeline),
// so only the boundaries that a call reaches are recorded, and mode 2 records
// only the two around the pair kernel.
void pb_mark(pb_context *c, hipStream_t s, int i)
{
    if (!c || !c->timing || !c->ev_ready || i < 0 || i > PB_NPHASE) return;
    if (c->timing == 2 && i < PB_NPHASE - 1) return;
    (void)hipEventRecord(c->ev[c->cur][i], s);
    c->recorded[c->cur] |= 1u << i;
}

void pb_mark_end(pb_context *c, hipStream_t s)
{
    if (c && c->timing && c->ev_ready && !(c->recorded[c->cur] & (1u << PB_NPHASE))) pb_mark(c, s, PB_NPHASE);
}

namespace {

// Fold a finished slot's phase durations into the running sums.
// (A phase counts only when the call recorded both its boundaries: phases
// that did not run have zero length.)
int timing_collect(pb_context *c, int slot)
{
    if (!c->slot_used[slot]) return PB_OK;
    c->slot_used[slot] = 0;
    const unsigned m = c->recorded[slot];
    if (!(m & (1u << PB_NPHASE))) return PB_OK; // the call stopped early: no sample
    for (int i = 0; i <= PB_NPHASE; i++) // (events of one call can sit on two streams)
        if (((m >> i) & 1u) && hipEventSynchronize(c->ev[slot][i]) != hipSuccess) return PB_ERR_HIP;
    for (int i = 0; i < PB_NPHASE; i++) {
        float t = 0;
        if (((m >> i) & 3u) == 3u && hipEventElapsedTime(&t, c->ev[slot][i], c->ev[slot][i + 1]) != hipSuccess)
            return PB_ERR_HIP;
        c->tsum[i] += t;
    }
    c->nsummed++;
    return PB_OK;
}

// Start the timing slot of a new call: every boundary is recorded once here so
// phases that do not run keep a zero-length interval.
int timing_begin(pb_context *c, hipStream_t s)
{
    if (!c->timing) return PB_OK;
    int slot = (int)(c->ncalls % PB_TSLOTS);
    int st = timing_collect(c, slot);
    if (st) return st;
    c->cur = slot;
    c->slot_used[slot] = 1;
    c->recorded[slot] = 0;
    c->ncalls++;
    pb_mark(c, s, 0);
    return PB_OK;
}

} // namespace

namespace {

bool finite_pos(double v) { return std::isfinite(v) && v > 0; }

int check_params(pb_context *c, const pb_params *p)
{
    if (!p) return pb_fail(c, PB_ERR_INVALID, "null params");
    for (int i = 0; i < 3; i++)
        if (!finite_pos(p->L[i])) return pb_fail(c, PB_ERR_INVALID, "box lengths must be finite and > 0");
    if (!finite_pos(p->alpha)) return pb_fail(c, PB_ERR_INVALID, "alpha must be finite and > 0");
    if (!(std::isfinite(p->eps) && p->eps >= 0)) return pb_fail(c, PB_ERR_INVALID, "eps must be >= 0");
    if (!finite_pos(p->d)) return pb_fail(c, PB_ERR_INVALID, "split d must be > 0 (run pb_tune)");
    if (!finite_pos(p->rc)) return pb_fail(c, PB_ERR_INVALID, "cutoff rc must be > 0 (run pb_tune)");
    if (p->precision != PB_PREC_FP64 && p->precision != PB_PREC_MIXED)
        return pb_fail(c, PB_ERR_INVALID, "precision must be PB_PREC_FP64 or PB_PREC_MIXED");
    if (p->kernel != PB_KERNEL_TABLE && p->kernel != PB_KERNEL_DIRECT && p->kernel != PB_KERNEL_TABLE_FAR32)
        return pb_fail(c, PB_ERR_INVALID,
                       "kernel must be PB_KERNEL_TABLE, PB_KERNEL_DIRECT or PB_KERNEL_TABLE_FAR32");
    return PB_OK;
}

int make_cells(pb_context *c, const pb_params *p, pb_cells *cells)
{
    std::memset(cells, 0, sizeof(*cells));
    long ncell = 1, nbrick = 1;
    for (int i = 0; i < 3; i++) {
        int n = (int)std::floor(3.0 * p->L[i] / p->rc);
        if (n < 1) n = 1;
        if (n > 160) n = 160;
        // an even count (cells only grow, so 3 cells still span rc): target
        // tiles are 2 x 2 cells across, and an odd count leaves a last column
        // one cell wide whose tiles pay a full candidate walk for half the
        // targets (1.1e5 targets, rc = 0.089: 33 -> 32 cells, 2% faster)
        if (n >= 8) n -= n % 2;
        cells->nc[i] = n;
        cells->nb[i] = (n + 3) / 4;
        cells->h[i] = p->L[i] / n;
        cells->inv_h[i] = n / p->L[i];
        cells->st[i] = (int)std::ceil(p->rc / cells->h[i]);
        if (p->rc > 3 * p->L[i])
            return pb_fail(c, PB_ERR_UNSUPPORTED, "cutoff rc larger than 3 box lengths is not supported");
        ncell *= n;
        nbrick *= cells->nb[i];
    }
    cells->ncell = (int)ncell;
    cells->nkey_t = (int)(nbrick * 64);
    cells->rc = p->rc;
    cells->rc2 = p->rc * p->rc;
    return PB_OK;
}

// The FFT supports powers of two in [4, 512] per axis, the window 1 <= P <=
// min(32, n) points: checked before anything is enqueued (the gather wraps a
// window index once, so a window wider than the grid would read outside it).
int check_kgrid(pb_context *c, const pb_kgrid &g)
{
    for (int ax = 0; ax < 3; ax++) {
        const int m = g.n[ax];
        if (m < 4 || (m & (m - 1)) || m > 512)
            return pb_fail(c, PB_ERR_UNSUPPORTED, "FFT grid sizes must be powers of two in [4, 512]");
        if (g.P[ax] < 1 || g.P[ax] > 32 || g.P[ax] > m)
            return pb_fail(c, PB_ERR_INVALID, "window points P must be in [1, min(32, grid)]");
    }
    return PB_OK;
}

// Window points per axis: P on every axis, or Paxis[i] (fftBrinkman's P[3]).
int make_kgrid(pb_context *c, const pb_params *p, pb_kgrid *g, const int *Paxis = nullptr)
{
    std::memset(g, 0, sizeof(*g));
    g->ng = 1;
    for (int i = 0; i < 3; i++) {
        if (p->grid[i] < 2) return pb_fail(c, PB_ERR_INVALID, "grid sizes must be >= 2 (run pb_tune)");
        g->n[i] = p->grid[i];
        g->P[i] = Paxis ? Paxis[i] : p->P;
        g->h[i] = p->L[i] / p->grid[i];
        g->inv_h[i] = (double)p->grid[i] / p->L[i];
        g->ng *= p->grid[i];
    }
    int st;
    if ((st = check_kgrid(c, *g))) return st;
    double m = p->m > 0 ? p->m : pb_window_m(p->P);
    double eta = p->P * g->h[0] / m; // fftBrinkman.h:113-114
    g->eta = eta * eta / 2;
    double sc = 1.0 / (g->eta * M_PI);
    g->scale = sc * std::sqrt(sc);
    g->vol = g->h[0] * g->h[1] * g->h[2];
    return PB_OK;
}

struct kparams_cache {
    double a = -1, d = -1, e = -1, rfar = -1;
    pb_kparams kp;
};

const pb_kparams &cached_kparams(double a, double d, double e, double rfar)
{
    static thread_local kparams_cache cache[2];
    static thread_local int next = 0;
    for (auto &k : cache)
        if (k.a == a && k.d == d && k.e == e && k.rfar == rfar) return k.kp;
    kparams_cache &k = cache[next];
    next ^= 1;
    pb_fill_kparams(&k.kp, a, d, e, rfar);
    k.a = a;
    k.d = d;
    k.e = e;
    k.rfar = rfar;
    return k.kp;
}

// Host copy of the kernel constants and (table kernels) the radial table on
// [rcommon, rc]: coefficients in coef (empty without a table); the table
// pointers of kp are left for the caller.
// Degree of the radial table for a precision: the FP64 form keeps the table
// at FP64 rounding (4e-15), the mixed form needs PB_TAB_TOL_MIXED and takes degree 6
// (112 bytes of coefficients per pair instead of 208, 14 FMAs instead of 26)
// where that fits the LDS budget (large alpha d needs short intervals), else 8,
// else 12.
int table_degree(const pb_params *p)
{
    return p->precision == PB_PREC_MIXED ? PB_TAB_DEG_MIXED : PB_TAB_DEG;
}

void host_kparams(const pb_params *p, double rfar, pb_kparams *kp, std::vector<double> &coef)
{
    *kp = cached_kparams(p->alpha, p->d, p->eps, rfar);
    kp->tn = 0;
    kp->tab64 = nullptr;
    kp->tab32 = nullptr;
    coef.clear();
    if (p->kernel != PB_KERNEL_DIRECT && p->rc > kp->rcommon) {
        const bool mixed = p->precision == PB_PREC_MIXED;
        const double z0 = kp->rcommon / p->d, z1 = p->rc / p->d * (1 + 1e-9);
        pb_build_table(kp, z0, z1, coef, table_degree(p), mixed ? PB_TAB_TOL_MIXED : 4e-15);
        if (kp->tn == 0 && mixed) pb_build_table(kp, z0, z1, coef, PB_TAB_DEG_MID, PB_TAB_TOL_MIXED);
        if (kp->tn == 0 && mixed) pb_build_table(kp, z0, z1, coef, PB_TAB_DEG, PB_TAB_TOL_MIXED);
        if (kp->tn == 0) coef.clear();
    }
}

// Device copy of the kernel constants and table for the given parameters,
// rebuilt on change.
int device_kparams(pb_context *c, hipStream_t s, const pb_params *p, double rfar,
                   const pb_kparams **out, size_t *table_doubles)
{
    int st;
    const double key[7] = {p->alpha, p->d, p->eps, rfar, p->rc, (double)p->kernel, (double)table_degree(p)};
    if ((st = pb_reserve(c, &c->kparams, sizeof(pb_kparams)))) return st;
    bool same = true;
    for (int i = 0; i < 7; i++) same = same && c->kp_key[i] == key[i];
    if (!same) {
        pb_kparams kp;
        std::vector<double> coef;
        host_kparams(p, rfar, &kp, coef);
        c->table_doubles = 0;
        if (!coef.empty()) {
            size_t n = coef.size();
            std::vector<float> c32(coef.begin(), coef.end());
            if ((st = pb_reserve(c, &c->table, n * (sizeof(double) + sizeof(float))))) return st;
            kp.tab64 = (const double *)c->table.ptr;
            kp.tab32 = (const float *)((char *)c->table.ptr + n * sizeof(double));
            if ((st = pb_hip_check(c, hipMemcpyAsync(c->table.ptr, coef.data(), n * sizeof(double),
                                                     hipMemcpyHostToDevice, s), "table upload")))
                return st;
            if ((st = pb_hip_check(c, hipMemcpyAsync((void *)kp.tab32, c32.data(), n * sizeof(float),
                                                     hipMemcpyHostToDevice, s), "table upload")))
                return st;
            c->table_doubles = n;
        }
        c->table_deg = kp.tdeg;
        if ((st = pb_hip_check(c, hipMemcpyAsync(c->kparams.ptr, &kp, sizeof(kp), hipMemcpyHostToDevice, s),
                               "kparams upload")))
            return st;
        if ((st = pb_hip_check(c, hipStreamSynchronize(s), "kparams upload"))) return st;
        for (int i = 0; i < 7; i++) c->kp_key[i] = key[i];
    }
    *out = (const pb_kparams *)c->kparams.ptr;
    *table_doubles = c->table_doubles;
    return PB_OK;
}

// FP32 far-field pairs: the closed-form kernel in mixed precision, or the
// table on request (on gfx950 FP64 FMA runs at the non-packed FP32 rate, so
// the FP64 table is the faster default).
bool far32_of(const pb_params *p, size_t ntab)
{
    return p->precision == PB_PREC_MIXED &&
           (p->kernel == PB_KERNEL_DIRECT || (p->kernel == PB_KERNEL_TABLE_FAR32 && ntab));
}

// The device work of one evaluation on stream s with kernel constants already
// on the device (kp, ntab): binning, k space (overwrite u), real space (add).
int enqueue_velocity(pb_context *c, hipStream_t s, const pb_params *p, const pb_kparams *kp, size_t ntab,
                     int tdeg, int ns, const double *xn, const double *f, int nt, const double *y, double *u,
                     bool do_real, bool do_k, bool overlap)
{
    int st;
    pb_box box;
    pb_fill_box(&box, p->L);
    pb_cells cells;
    if ((st = make_cells(c, p, &cells))) return st;
    pb_kgrid g;
    if (do_k && (st = make_kgrid(c, p, &g))) return st; // (validated before anything is enqueued)
    overlap = overlap && do_real && do_k && nt > 0 && c->timing != 1;
    // (the overlapped form's output zeroed by the binning's first launch)
    if ((st = pb_bin_all(c, s, cells, box, ns, xn, f, nt, y, overlap ? u : nullptr))) return st;
    pb_mark(c, s, 1);
    // Both halves: the pair kernel on a stream of its own, forked from the
    // k-space pipeline and concurrent with the rest of it (each adds to a
    // zeroed u with atomics; the pair kernel's latency-bound waves and the
    // k-space kernels' matrix-core-, memory- and texture-path-bound ones fill
    // each other's idle issue slots).  Not while every phase is timed (the
    // phases would overlap).
    if (overlap) {
        if (!c->pair_stream &&
            (st = pb_hip_check(c, hipStreamCreateWithFlags(&c->pair_stream, hipStreamNonBlocking), "stream")))
            return st;
        for (int k = 0; k < 2; k++)
            if (!c->pair_event[k] &&
                (st = pb_hip_check(c, hipEventCreateWithFlags(&c->pair_event[k], hipEventDisableTiming), "event")))
                return st;
        bool forked = false;
        int jst = 0;
        // Fork point: before the spread, whose FP32 form runs on the matrix
        // cores and leaves the vector units to the pair kernel (at 1e5 -> 1e5,
        // 128^3, P 12: mixed 0.476 against 0.487 ms with the fork after the
        // spread, closed forms 1.106 against 1.178 ms); after the spread for
        // the FP64 table form, whose FP64 spread contends with the pair kernel
        // for the vector units (1.063 against 1.081 ms), and for the balanced
        // spread of large work (its item workgroups fill the device: 1e6
        // clustered sources, 256^3, P 16, 44.4-44.8 against 45.2-45.8 ms).
        // Which of the two goes first onto the device after the fork decides
        // which gets the compute units: the launch on the stream that waits
        // for the fork event starts a few microseconds later.  The pair
        // kernel goes to the pair stream (the k space first) except with the
        // balanced spread, where the pair kernel is the longer path (the
        // clustered system: 42.1 against 45.1 ms per call with the k space
        // first) and stays on s, the rest of the k space moving to the pair
        // stream.
        const bool split = pb_spread_split(c, g, ns);
        const bool fp64_table = p->precision != PB_PREC_MIXED && ntab > 0;
        const double rounds = pb_pair_rounds(p->precision, ntab);
        const std::function<int(hipStream_t &)> fork = [&](hipStream_t &ks) -> int {
            int q;
            if ((q = pb_hip_check(c, hipEventRecord(c->pair_event[0], s), "fork"))) return q;
            if ((q = pb_hip_check(c, hipStreamWaitEvent(c->pair_stream, c->pair_event[0], 0), "fork"))) return q;
            // from here every return joins the pair stream back into s, so no
            // work of this call is left running beside the caller's next call
            forked = true;
            if (split) {
                ks = c->pair_stream;
                return pb_run_real(c, s, cells, kp, ntab, tdeg, box, ns, nt, u, true, far32_of(p, ntab), rounds);
            }
            q = pb_run_real(c, c->pair_stream, cells, kp, ntab, tdeg, box, ns, nt, u, true, far32_of(p, ntab),
                            rounds);
            jst = pb_hip_check(c, hipEventRecord(c->pair_event[1], c->pair_stream), "join");
            return q ? q : jst;
        };
        st = pb_run_kspace(c, s, g, cells, box, p->alpha, p->d, ns, nt, u, 2, p->precision, &fork,
                           fp64_table || split ? 1 : 0);
        if (forked && split) // (the k space's last launch is on the pair stream)
            jst = pb_hip_check(c, hipEventRecord(c->pair_event[1], c->pair_stream), "join");
        const int wst = !forked ? 0 : jst ? jst : pb_hip_check(c, hipStreamWaitEvent(s, c->pair_event[1], 0), "join");
        if (!forked && !st) st = pb_fail(c, PB_ERR_INVALID, "real space not launched");
        return st ? st : wst;
    }
    bool acc = false;
    if (do_k) {
        if ((st = pb_run_kspace(c, s, g, cells, box, p->alpha, p->d, ns, nt, u, 0, p->precision))) return st;
        acc = true;
    }
    if (do_real &&
        (st = pb_run_real(c, s, cells, kp, ntab, tdeg, box, ns, nt, u, acc, far32_of(p, ntab),
                           pb_pair_rounds(p->precision, ntab))))
        return st;
    return PB_OK;
}

int velocity_impl(pb_context *c, const pb_params *p, int ns, const double *xn, const double *f,
                  int nt, const double *y, double *u, void *stream, bool do_real, bool do_k)
{
    if (!c) return PB_ERR_INVALID;
    int st;
    if ((st = check_params(c, p))) return st;
    if (ns < 0 || nt < 0) return pb_fail(c, PB_ERR_INVALID, "negative point count");
    if ((ns > 0 && (!xn || !f)) || (nt > 0 && (!y || !u)))
        return pb_fail(c, PB_ERR_INVALID, "null point array");
    if ((st = pb_hip_check(c, hipSetDevice(c->device), "hipSetDevice"))) return st;
    hipStream_t s = (hipStream_t)stream;
    const pb_kparams *kp = nullptr;
    size_t ntab = 0;
    if (do_real && (st = device_kparams(c, s, p, std::fmax(3 * p->d, 8 * p->eps), &kp, &ntab))) return st;
    if ((st = timing_begin(c, s))) return pb_fail(c, st, "timing events");
    if ((st = enqueue_velocity(c, s, p, kp, ntab, c->table_deg, ns, xn, f, nt, y, u, do_real, do_k, true)))
        return st;
    pb_mark_end(c, s);
    return PB_OK;
}

// A default context per host thread and device for the reference-signature
// entry points (created on the thread's current device at first use,
// destroyed when the thread ends).
struct default_ctx_holder {
    std::vector<pb_context *> ctx; // by device ordinal
    ~default_ctx_holder()
    {
        for (pb_context *c : ctx)
            if (c) pb_destroy(c);
    }
};

pb_context *default_ctx()
{
    static thread_local default_ctx_holder h;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    if ((int)h.ctx.size() <= dev) h.ctx.resize(dev + 1, nullptr);
    if (!h.ctx[dev] && pb_create(&h.ctx[dev], dev) != PB_OK) h.ctx[dev] = nullptr;
    return h.ctx[dev];
}

// hipMemcpy that folds failures into a status
int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (bytes == 0) return PB_OK;
    return hipMemcpy(dst, src, bytes, kind) == hipSuccess ? PB_OK : PB_ERR_HIP;
}

struct dev_tmp {
    std::vector<void *> ptrs;
    ~dev_tmp()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    double *alloc(size_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(double)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (double *)p;
    }
};

int shells_impl(double *A, double *absA, int lda, const double *x, int nt, const double *x0,
                int ns, const int *newM, const int *oldM, const double *L, double a, double d,
                double e, bool fourier)
{
    if (!A || !absA || !x || !x0 || !newM || !oldM || !L) return PB_ERR_INVALID;
    if (nt < 0 || ns < 0 || lda < 3 * nt) return PB_ERR_INVALID;
    if (!(a > 0) || !(d > 0) || !(e >= 0)) return PB_ERR_INVALID;
    for (int i = 0; i < 3; i++)
        if (!(L[i] > 0) || newM[i] < 0 || oldM[i] < 0) return PB_ERR_INVALID;
    bool home = newM[0] == oldM[0] && newM[1] == oldM[1];
    if (fourier && home) return PB_OK; // periodicBrinkman.h:719-721
    if (nt == 0 || ns == 0) return PB_OK;
    dev_tmp t;
    size_t na = (size_t)lda * 3 * ns;
    double *dA = t.alloc(na), *dB = t.alloc(na), *dx = t.alloc(3 * (size_t)nt), *dx0 = t.alloc(3 * (size_t)ns);
    if (!dA || !dB || !dx || !dx0) return PB_ERR_NOMEM;
    int st = copy(dA, A, na * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(dB, absA, na * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(dx, x, 3 * (size_t)nt * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(dx0, x0, 3 * (size_t)ns * sizeof(double), hipMemcpyHostToDevice);
    if (st) return st;
    pb_box box;
    pb_fill_box(&box, L);
    const pb_kparams &kp = cached_kparams(a, d, e, 0.0);
    pb_context *c = default_ctx();
    if (!c) return PB_ERR_HIP;
    if (fourier)
        st = pb_run_dense(c, nullptr, dA, lda, dx, nt, dx0, ns, kp, box, nullptr, nullptr, 0, oldM, newM, 1, dB);
    else
        st = pb_run_dense(c, nullptr, dA, lda, dx, nt, dx0, ns, kp, box, oldM, newM, home ? 1 : 2, nullptr,
                          nullptr, 0, dB);
    if (st) return st;
    if (hipDeviceSynchronize() != hipSuccess) return PB_ERR_HIP;
    st = copy(A, dA, na * sizeof(double), hipMemcpyDeviceToHost);
    if (!st) st = copy(absA, dB, na * sizeof(double), hipMemcpyDeviceToHost);
    return st;
}

// ---- the allocation-free entry point ----------------------------------------
// One context per device works in arena mode on the device's scratch array
// (pb_noalloc.hip): the constants at fixed offsets of its first
// PB_NOALLOC_FIXED bytes (kept across calls, re-uploaded when the parameters
// change), every other buffer carved anew by each call.  A mutex per device
// serializes the calls' host-side set-up; the device work of calls on one
// stream is ordered by that stream, calls on different streams race.
constexpr int kNaDevices = 64;
constexpr size_t kNaTable = 0, kNaKparams = 32u << 10, kNaTwid64 = 36u << 10, kNaTwid32 = 52u << 10;
static_assert(PB_TAB_LDS_DOUBLES * (sizeof(double) + sizeof(float)) <= kNaKparams - kNaTable, "table slot");
static_assert(sizeof(pb_kparams) <= kNaTwid64 - kNaKparams, "kernel-constant slot");
static_assert(1020 * sizeof(double2) <= kNaTwid32 - kNaTwid64, "FP64 twiddle slot");
static_assert(kNaTwid32 + 1020 * sizeof(float2) <= PB_NOALLOC_FIXED, "FP32 twiddle slot");
pb_context *na_ctx[kNaDevices] = {};
std::mutex na_mutex[kNaDevices];
std::mutex na_setup_mutex;

void noalloc_setup(int device, int cu_count)
{
    if (device < 0 || device >= kNaDevices) return;
    std::lock_guard<std::mutex> lock(na_setup_mutex);
    if (na_ctx[device] || pb_noalloc_scratch_init(device) != PB_OK) return;
    char *base = (char *)pb_noalloc_scratch(device);
    if (!base) return;
    pb_context *na = new pb_context();
    na->device = device;
    na->cu_count = cu_count;
    na->arena = base;
    na->arena_end = PB_NOALLOC_BYTES;
    na->arena_used = PB_NOALLOC_FIXED;
    na->table = {base + kNaTable, kNaKparams - kNaTable};
    na->kparams = {base + kNaKparams, kNaTwid64 - kNaKparams};
    na->twid[0] = {base + kNaTwid64, kNaTwid32 - kNaTwid64};
    na->twid[1] = {base + kNaTwid32, PB_NOALLOC_FIXED - kNaTwid32};
    na_ctx[device] = na;
}

// Device bytes the arena hands out for one evaluation (the pb_reserve calls of
// the fused binning, the k space and the pair kernel, 256-byte slices).
size_t noalloc_bytes(const pb_context *na, const pb_cells &cells, const pb_kgrid &g, int precision, int ns,
                     int nt)
{
    auto sl = [](size_t b) { return ((b ? b : 16) + 255) & ~(size_t)255; };
    const size_t nsn = (size_t)(ns > 0 ? ns : 1), ntn = (size_t)(nt > 0 ? nt : 1);
    const size_t ncell = (size_t)cells.ncell, nkt = (size_t)cells.nkey_t, nq = nkt / 16;
    const size_t mt = std::min<size_t>(nq, (size_t)nt) + (size_t)nt / 64 + 1;
    size_t b = 2 * sl(4 * (ncell + 1)) + 2 * sl(4 * nsn) + 2 * sl(32 * nsn) + sl(4 * nsn);  // sources
    b += 2 * sl(4 * (nkt + 1)) + 2 * sl(4 * ntn) + sl(32 * ntn) + sl(4 * ntn);            // targets
    b += sl(4 * (nq + 1)) + sl(8 * mt);                                                    // tiles
    const size_t cb = precision == PB_PREC_MIXED ? 8 : 16;                                 // complex bytes
    b += sl((size_t)g.ng * 2 * cb) + sl(3 * (size_t)g.n[0] * g.n[1] * (g.n[2] / 2 + 1) * cb); // grid, spectrum
    if (nt > 0) { // pair-kernel statistics (pb_run_real's launch size)
        // (the most wavefronts per tile of any form)
        const int nsplit = std::max(pb_pair_split(na->cu_count, nt, pb_pair_rounds(PB_PREC_FP64, 0)),
                                    std::max(pb_pair_split(na->cu_count, nt, pb_pair_rounds(PB_PREC_FP64, 1)),
                                             pb_pair_split(na->cu_count, nt, pb_pair_rounds(PB_PREC_MIXED, 1))));
        const size_t blocks = ((size_t)mt * nsplit + pb_pair_waves - 1) / pb_pair_waves;
        b += sl(sizeof(uint4) * blocks * pb_pair_waves);
    }
    return b;
}

// Every buffer of a context (the constants' slots of the arena context are
// kept, the others carved anew per call).
template <typename F> void for_each_buffer(pb_context *c, F fn)
{
    pb_buffer *bufs[] = {&c->cell_count, &c->cell_start, &c->cell_of_s, &c->rank_s, &c->cell_of_t, &c->rank_t,
                         &c->tcell_count, &c->tcell_start, &c->spos, &c->sfrc, &c->tpos, &c->tperm, &c->counters,
                         &c->scan_tmp, &c->qcount, &c->qoff, &c->tiles, &c->grid, &c->spec, &c->sp_int, &c->sp_items, &c->sp_part, &c->sperm, &c->mrhs_f, &c->mrhs_u,
                         &c->twid[0], &c->twid[1], &c->batch_kp, &c->batch_tab, &c->dense_tmp, &c->kparams,
                         &c->table, &c->gb_int, &c->gb_d4, &c->gb_tiles, &c->gb_grid, &c->gb_spec, &c->gb_sys,
                         &c->gb_map, &c->gb_stats, &c->mr_sys, &c->mr_map, &c->mr_frc, &c->mr_grid, &c->mr_spec};
    for (pb_buffer *b : bufs) fn(c, b);
}

} // namespace

extern "C" {

const char *pb_status_string(int status)
{
    switch (status) {
    case PB_OK: return "ok";
    case PB_ERR_INVALID: return "invalid argument";
    case PB_ERR_HIP: return "HIP runtime error";
    case PB_ERR_NOMEM: return "device out of memory";
    case PB_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

int pb_create(pb_context **ctx, int device)
{
    if (!ctx) return PB_ERR_INVALID;
    *ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PB_ERR_HIP;
    if (device < 0 || device >= n) return PB_ERR_INVALID;
    pb_context *c = new pb_context();
    c->device = device;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cu_count = cus;
    // set-up of the allocation-free entry point on this device (once per
    // device; a failure only disables that entry point)
    noalloc_setup(device, c->cu_count);
    *ctx = c;
    return PB_OK;
}

int pb_destroy(pb_context *c)
{
    if (!c) return PB_OK;
    (void)hipSetDevice(c->device);
    for (int k = 0; k < 4; k++) {
        if (c->sub[k]) pb_destroy(c->sub[k]);
        if (c->sub_stream[k]) (void)hipStreamDestroy(c->sub_stream[k]);
    }
    for (int k = 0; k < 7; k++)
        if (c->sub_event[k]) (void)hipEventDestroy(c->sub_event[k]);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
    if (c->pair_stream) (void)hipStreamDestroy(c->pair_stream);
    if (c->pair_stream2) (void)hipStreamDestroy(c->pair_stream2);
    if (c->pair_event2) (void)hipEventDestroy(c->pair_event2);
    for (int k = 0; k < 2; k++)
        if (c->pair_event[k]) (void)hipEventDestroy(c->pair_event[k]);
    if (!c->arena) // (an arena context's buffers are slices of the device scratch)
        for_each_buffer(c, [](pb_context *, pb_buffer *b) {
            if (b->ptr) (void)hipFree(b->ptr);
        });
    if (c->ev_ready)
        for (int k = 0; k < PB_TSLOTS; k++)
            for (int i = 0; i <= PB_NPHASE; i++) (void)hipEventDestroy(c->ev[k][i]);
    delete c;
    return PB_OK;
}

const char *pb_last_error(const pb_context *c) { return c ? c->err.c_str() : "null context"; }

int pb_set_timing(pb_context *c, int enable)
{
    if (!c) return PB_ERR_INVALID;
    if (enable && !c->ev_ready) {
        if (hipSetDevice(c->device) != hipSuccess) return PB_ERR_HIP;
        for (int k = 0; k < PB_TSLOTS; k++)
            for (int i = 0; i <= PB_NPHASE; i++)
                if (hipEventCreate(&c->ev[k][i]) != hipSuccess) return PB_ERR_HIP;
        c->ev_ready = 1;
    }
    // drain outstanding slots, then restart the averages
    for (int k = 0; k < PB_TSLOTS; k++)
        if (c->slot_used[k] && hipEventSynchronize(c->ev[k][PB_NPHASE]) != hipSuccess) return PB_ERR_HIP;
    for (int k = 0; k < PB_TSLOTS; k++) c->slot_used[k] = 0;
    for (int i = 0; i < PB_NPHASE; i++) c->tsum[i] = 0;
    c->ncalls = c->nsummed = 0;
    c->timing = enable == 2 ? 2 : (enable ? 1 : 0);
    return PB_OK;
}

int pb_last_timings(pb_context *c, float *ms, int n)
{
    if (!c || !ms || n < 0 || n > PB_NPHASE) return PB_ERR_INVALID;
    if (!c->ev_ready) return pb_fail(c, PB_ERR_INVALID, "timing was never enabled");
    for (int k = 0; k < PB_TSLOTS; k++) {
        int st = timing_collect(c, k);
        if (st) return pb_fail(c, st, "timing events");
    }
    for (int i = 0; i < n; i++) ms[i] = c->nsummed ? (float)(c->tsum[i] / c->nsummed) : 0.0f;
    return PB_OK;
}

int pb_velocity(pb_context *ctx, const pb_params *p, int ns, const double *xn, const double *f,
                int nt, const double *y, double *u, void *stream)
{
    return velocity_impl(ctx, p, ns, xn, f, nt, y, u, stream, true, true);
}

int pb_velocity_real(pb_context *ctx, const pb_params *p, int ns, const double *xn,
                     const double *f, int nt, const double *y, double *u, void *stream)
{
    return velocity_impl(ctx, p, ns, xn, f, nt, y, u, stream, true, false);
}

int pb_velocity_kspace(pb_context *ctx, const pb_params *p, int ns, const double *xn,
                       const double *f, int nt, const double *y, double *u, void *stream)
{
    return velocity_impl(ctx, p, ns, xn, f, nt, y, u, stream, false, true);
}

int pb_velocity_noalloc(pb_context *c, const pb_params *p, int ns, const double *xn, const double *f, int nt,
                        const double *y, double *u, void *stream)
{
    if (!c) return PB_ERR_INVALID;
    int st;
    if ((st = check_params(c, p))) return st;
    if (ns < 0 || nt < 0) return pb_fail(c, PB_ERR_INVALID, "negative point count");
    if ((ns > 0 && (!xn || !f)) || (nt > 0 && (!y || !u))) return pb_fail(c, PB_ERR_INVALID, "null point array");
    pb_context *na = c->device >= 0 && c->device < kNaDevices ? na_ctx[c->device] : nullptr;
    if (!na) return pb_fail(c, PB_ERR_HIP, "the device scratch of the no-allocation entry point is unavailable");
    pb_cells cells;
    pb_kgrid g;
    if ((st = make_cells(c, p, &cells))) return st;
    if ((st = make_kgrid(c, p, &g))) return st;
    if (!pb_bin_fused_ok(cells))
        return pb_fail(c, PB_ERR_UNSUPPORTED, "the no-allocation entry point needs cell keys + target keys <= 2^18 "
                                              "(a larger cutoff rc)");
    const size_t need = noalloc_bytes(na, cells, g, p->precision, ns, nt);
    if (need > PB_NOALLOC_BYTES - PB_NOALLOC_FIXED) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "the workload needs %.1f MiB of scratch; the no-allocation entry point "
                      "holds at most %.1f MiB", need / 1048576.0, (PB_NOALLOC_BYTES - PB_NOALLOC_FIXED) / 1048576.0);
        return pb_fail(c, PB_ERR_UNSUPPORTED, msg);
    }
    if ((st = pb_hip_check(c, hipSetDevice(c->device), "hipSetDevice"))) return st;
    std::lock_guard<std::mutex> lock(na_mutex[c->device]);
    // this call's buffers are carved from the scratch anew (the constants'
    // slots at its start are kept)
    for_each_buffer(na, [](pb_context *a, pb_buffer *b) {
        if (b != &a->table && b != &a->kparams && b != &a->twid[0] && b != &a->twid[1]) *b = pb_buffer{};
    });
    na->arena_used = PB_NOALLOC_FIXED;
    na->err.clear();
    hipStream_t s = (hipStream_t)stream;
    const pb_kparams *kp = nullptr;
    size_t ntab = 0;
    // the binning's count arrays are cleared by a fill kernel, the grid is
    // written point by point, u overwritten by the gather: nothing in the
    // scratch needs to start at zero
    st = device_kparams(na, s, p, std::fmax(3 * p->d, 8 * p->eps), &kp, &ntab);
    if (!st)
        st = enqueue_velocity(na, s, p, kp, ntab, na->table_deg, ns, xn, f, nt, y, u, true, true, false);
    if (st) return pb_fail(c, st, na->err);
    return PB_OK;
}

namespace {

size_t pad16(size_t n) { return (n + 15) & ~(size_t)15; }
size_t pad256(size_t n) { return (n + 255) & ~(size_t)255; }

// k space of R force sets (FP32 pipeline) on the binned particle set: up to
// 64 sets per grouped launch (fewer when their grids would pass 4 GB), each
// set a system of the group with the shared cells, sorted points and k grid
// and its own forces, grid, spectrum and block of u (overwritten).
// fork is called once, after the first chunk's maps and force sets are
// enqueued: work the caller puts beside the k space from there on does not
// hold up these small launches.
int mrhs_kspace(pb_context *c, hipStream_t s, const pb_params *p, const pb_cells &cells, const pb_box &box,
                const pb_kgrid &g, int ns, int nt, int R, int RP, double *u, const std::function<int()> &fork)
{
    int st;
    const size_t gbytes = pad256((size_t)g.ng * 16);                                   // float4 per point
    const size_t sbytes = pad256(3 * (size_t)g.n[0] * g.n[1] * (g.n[2] / 2 + 1) * 8); // complex float
    const size_t fbytes = (size_t)std::max(ns, 1) * sizeof(double4);
    const int CH = (int)std::max<size_t>(1, std::min<size_t>(64, ((size_t)4 << 30) / (gbytes + sbytes + fbytes)));
    const int nc = std::min(R, CH);
    const void *tw;
    if ((st = pb_twiddles(c, s, PB_PREC_MIXED, &tw))) return st;
    if ((st = pb_reserve(c, &c->mr_grid, gbytes * nc))) return st;
    if ((st = pb_reserve(c, &c->mr_spec, sbytes * nc))) return st;
    if ((st = pb_reserve(c, &c->mr_frc, fbytes * nc))) return st;
    // one descriptor per set, uploaded once per call
    pb_sys base;
    std::memset(&base, 0, sizeof base);
    base.ns = ns;
    base.nt = nt;
    base.box = box;
    base.cells = cells;
    base.g = g;
    base.km = pb_make_kmult(g, box, p->alpha, p->d);
    base.sstart = (int *)c->cell_start.ptr;
    base.spos = (double4 *)c->spos.ptr;
    base.tpos = (double4 *)c->tpos.ptr;
    base.tperm = (int *)c->tperm.ptr;
    base.acc_mode = 0;
    std::vector<pb_sys> sys(R, base);
    for (int r = 0; r < R; r++) {
        const int q = r % CH;
        sys[r].sfrc = (double4 *)((char *)c->mr_frc.ptr + fbytes * q);
        sys[r].grid4 = (char *)c->mr_grid.ptr + gbytes * q;
        sys[r].spec = (char *)c->mr_spec.ptr + sbytes * q;
        sys[r].u = u + (size_t)r * 3 * nt;
    }
    // uploaded (with a wait for the copy) only when they differ from the last
    // call's: a repeated call enqueues its work without blocking the host
    if (sys.size() != c->mr_sys_host.size() ||
        std::memcmp(sys.data(), c->mr_sys_host.data(), sizeof(pb_sys) * sys.size()) != 0) {
        c->mr_sys_host.clear();
        if ((st = pb_reserve(c, &c->mr_sys, sizeof(pb_sys) * R))) return st;
        if ((st = pb_hip_check(c, hipMemcpyAsync(c->mr_sys.ptr, sys.data(), sizeof(pb_sys) * R,
                                                 hipMemcpyHostToDevice, s), "descriptor upload")))
            return st;
        if ((st = pb_hip_check(c, hipStreamSynchronize(s), "descriptor upload"))) return st;
        c->mr_sys_host = sys;
    }
    // maps: spread, the FFT passes (five, or three with the fused plane
    // passes), the gather (XCD-renumbered: a multiple of 8)
    const int nreg = ((g.n[0] + 7) / 8) * ((g.n[1] + 7) / 8) * ((g.n[2] + 7) / 8);
    int passes[5];
    const int npass = pb_fft_passes(PB_PREC_MIXED, g, passes);
    int per[7] = {0, 0, 0, 0, 0, 0, 0};
    per[0] = nreg;
    for (int k = 0; k < npass; k++) per[1 + k] = pb_group_fft_blocks(PB_PREC_MIXED, passes[k], g);
    const bool same = g.P[0] == g.P[1] && g.P[1] == g.P[2];
    const int Ps[8] = {4, 6, 8, 10, 12, 16, 20, 24};
    const int P = same && std::find(Ps, Ps + 8, g.P[0]) != Ps + 8 ? g.P[0] : 0; // 0: per-axis windows
    const int tpb = P ? pb_interp_waves * pb_interp_tw(P) : 256; // targets per gather block
    per[6] = (nt + tpb - 1) / tpb;
    size_t off[8] = {0};
    for (int k = 0; k < 7; k++) {
        size_t n = (size_t)per[k] * nc;
        if (k == 6) n = 8 * (size_t)nc * ((per[6] + 7) / 8);
        off[k + 1] = off[k] + n;
    }
    if ((st = pb_reserve(c, &c->mr_map, sizeof(int2) * off[7]))) return st;
    int2 *map = (int2 *)c->mr_map.ptr;
    int filled = -1; // chunk size the maps were filled for
    long long ngather = 0;
    for (int r0 = 0; r0 < R; r0 += CH) {
        const int n = std::min(CH, R - r0);
        if (n != filled) {
            for (int k = 0; k < 6; k++) {
                const long long valid = (long long)per[k] * n;
                pb_fill_map(s, map + off[k], valid, per[k], valid);
            }
            ngather = P ? pb_fill_map_xcd(s, map + off[6], n, per[6]) : (long long)per[6] * n;
            if (!P) pb_fill_map(s, map + off[6], ngather, per[6], ngather);
            filled = n;
        }
        const pb_sys *sd = (const pb_sys *)c->mr_sys.ptr + r0;
        if ((st = pb_force_sets(c, s, ns, r0, n, RP, (double4 *)c->mr_frc.ptr))) return st;
        if (r0 == 0 && (st = fork())) return st;
        // spreading: groups of pb_spread_sets sets per workgroup (one staging
        // and one set of window weights for all of them), then the rest one by one
        const int n8 = n / pb_spread_sets, rest = n - n8 * pb_spread_sets;
        pb_group_spread_sets(s, sd, map + off[0], per[0] * n8);
        if (rest) pb_group_spread(s, sd + n8 * pb_spread_sets, map + off[0], per[0] * rest, PB_PREC_MIXED);
        for (int k = 0; k < npass; k++) {
            const int pass = passes[k];
            if ((st = pb_group_fft(s, sd, map + off[1 + k], per[1 + k] * n, PB_PREC_MIXED, pass,
                                   g.n[pb_fft_axis(pass)], tw)))
                return pb_fail(c, st, "FFT size");
        }
        if ((st = pb_group_interp(s, sd, map + off[6], (int)ngather, PB_PREC_MIXED, P)))
            return pb_fail(c, st, "gather window");
    }
    return pb_hip_check(c, hipGetLastError(), "multi-RHS k-space launch");
}

} // namespace

int pb_velocity_multirhs(pb_context *c, const pb_params *p, int ns, const double *xn, int R,
                         const float *f, int nt, const double *y, double *u, void *stream)
{
    if (!c) return PB_ERR_INVALID;
    int st;
    if ((st = check_params(c, p))) return st;
    if (ns < 0 || nt < 0 || R < 1) return pb_fail(c, PB_ERR_INVALID, "negative point count or R < 1");
    if ((ns > 0 && (!xn || !f)) || (nt > 0 && (!y || !u))) return pb_fail(c, PB_ERR_INVALID, "null point array");
    if (p->kernel == PB_KERNEL_DIRECT)
        return pb_fail(c, PB_ERR_UNSUPPORTED, "the multi-RHS path evaluates the radial table (PB_KERNEL_TABLE)");
    if ((st = pb_hip_check(c, hipSetDevice(c->device), "hipSetDevice"))) return st;
    hipStream_t s = (hipStream_t)stream;
    pb_box box;
    pb_fill_box(&box, p->L);
    pb_cells cells;
    if ((st = make_cells(c, p, &cells))) return st;
    pb_kgrid g;
    if ((st = make_kgrid(c, p, &g))) return st;
    // always the mixed form: the table at the mixed form's degree and
    // tolerance, built (and checked to exist) before anything is enqueued
    const double rfar = std::fmax(3 * p->d, 8 * p->eps);
    pb_params pt = *p;
    pt.kernel = PB_KERNEL_TABLE;
    pt.precision = PB_PREC_MIXED;
    const pb_kparams *kp = nullptr;
    size_t ntab = 0;
    if ((st = device_kparams(c, s, &pt, rfar, &kp, &ntab))) return st;
    if (nt > 0 && !ntab)
        return pb_fail(c, PB_ERR_UNSUPPORTED, "the multi-RHS path needs the radial table (cutoff rc too small)");
    const int RP = (R + 31) / 32 * 32;
    const long long nu = (long long)R * 3 * nt;
    if ((st = pb_reserve(c, &c->mrhs_u, sizeof(double) * (size_t)std::max(nu, 1LL)))) return st;
    if ((st = pb_bin_all(c, s, cells, box, ns, xn, nullptr, nt, y))) return st;
    if ((st = pb_sort_forces(c, s, ns, R, RP, f))) return st;
    if (nt == 0) return PB_OK;
    // The real-space contraction (matrix cores and FP64 radial functions) runs
    // on the pair stream beside the k space (memory- and VALU-bound spreading,
    // FFT and gather) on s, into a buffer of its own that it overwrites; the k
    // space overwrites u; after the join u += the real-space part.
    if (!c->pair_stream &&
        (st = pb_hip_check(c, hipStreamCreateWithFlags(&c->pair_stream, hipStreamNonBlocking), "stream")))
        return st;
    for (int k = 0; k < 2; k++)
        if (!c->pair_event[k] &&
            (st = pb_hip_check(c, hipEventCreateWithFlags(&c->pair_event[k], hipEventDisableTiming), "event")))
            return st;
    bool forked = false;
    int jst = 0;
    auto fork = [&]() -> int {
        int q;
        if ((q = pb_hip_check(c, hipEventRecord(c->pair_event[0], s), "fork"))) return q;
        if ((q = pb_hip_check(c, hipStreamWaitEvent(c->pair_stream, c->pair_event[0], 0), "fork"))) return q;
        forked = true; // from here every return joins the pair stream back into s
        q = pb_run_mrhs_real(c, c->pair_stream, cells, kp, ntab, c->table_deg, box, ns, nt, R, RP,
                             (double *)c->mrhs_u.ptr);
        jst = pb_hip_check(c, hipEventRecord(c->pair_event[1], c->pair_stream), "join");
        return q ? q : jst;
    };
    st = mrhs_kspace(c, s, p, cells, box, g, ns, nt, R, RP, u, fork);
    const int wst = !forked ? 0 : jst ? jst : pb_hip_check(c, hipStreamWaitEvent(s, c->pair_event[1], 0), "join");
    if (!forked && !st) st = pb_fail(c, PB_ERR_INVALID, "multi-RHS real space not launched");
    if (st || wst) return st ? st : wst;
    pb_add(s, nu, u, (const double *)c->mrhs_u.ptr);
    return pb_hip_check(c, hipGetLastError(), "multi-RHS add");
}

namespace {

// ---- grouped batched systems ------------------------------------------------
// Systems whose binning fits the fused path (one scan workgroup) run as one
// group: a launch per phase and class (kernel template) covers all of them,
// each launched block finding its system and its block of that system in a
// map, so a call of hundreds of systems is a few dozen launches that fill the
// device instead of a dozen small launches per system.

// Scratch layout, descriptors, maps and launch list of the group idx; kept
// while key (systems, parameters, arrays) is unchanged.
int group_prepare(pb_context *c, hipStream_t s, const std::vector<int> &idx, const pb_params *params,
                  const int *src_off, const int *tgt_off, const double *xn, const double *f, const double *y,
                  double *u, const std::vector<double> &key)
{
    if (key == c->group_key) return PB_OK;
    c->group_key.clear();
    c->group_launches.clear();
    int st;
    const int n = (int)idx.size();
    std::vector<pb_sys> sys(n);
    std::vector<size_t> mt(n);
    size_t ncount = 0, nint = 0, nd4 = 0, ntile = 0, ngrid = 0, nspec = 0;
    for (int q = 0; q < n; q++) {
        const int i = idx[q];
        const pb_params &p = params[i];
        pb_sys &S = sys[q];
        std::memset(&S, 0, sizeof S);
        pb_fill_box(&S.box, p.L);
        if ((st = make_cells(c, &p, &S.cells))) return st;
        if ((st = make_kgrid(c, &p, &S.g))) return st; // (checks the grid and window)
        S.km = pb_make_kmult(S.g, S.box, p.alpha, p.d);
        S.ns = src_off[i + 1] - src_off[i];
        S.nt = tgt_off[i + 1] - tgt_off[i];
        S.xn = S.ns ? xn + 3 * (size_t)src_off[i] : nullptr;
        S.f = S.ns && f ? f + 3 * (size_t)src_off[i] : nullptr;
        S.y = S.nt ? y + 3 * (size_t)tgt_off[i] : nullptr;
        S.u = S.nt ? u + 3 * (size_t)tgt_off[i] : nullptr;
        S.kp = (const pb_kparams *)c->batch_kp.ptr + i;
        S.acc_mode = 2;
        const int nc = S.cells.ncell, nk = S.cells.nkey_t;
        ncount += pad16(nc + 1) + pad16(nk + 1);
        nint += pad16(nc + 1) + 3 * pad16(S.ns) + pad16(nk + 1) + 3 * pad16(S.nt) + 16;
        nd4 += 2 * pad16(S.ns) + pad16(S.nt);
        mt[q] = (size_t)std::min(nk / 16, S.nt) + (size_t)S.nt / 64 + 1;
        ntile += pad16(mt[q]);
        const size_t cb = p.precision == PB_PREC_MIXED ? 8 : 16; // complex bytes
        ngrid += pad256((size_t)S.g.ng * 2 * cb);
        nspec += pad256(3 * (size_t)S.g.n[0] * S.g.n[1] * (S.g.n[2] / 2 + 1) * cb);
    }
    if ((st = pb_reserve(c, &c->gb_int, sizeof(int) * (ncount + nint)))) return st;
    if ((st = pb_reserve(c, &c->gb_d4, sizeof(double4) * std::max<size_t>(nd4, 1)))) return st;
    if ((st = pb_reserve(c, &c->gb_tiles, sizeof(int2) * ntile))) return st;
    if ((st = pb_reserve(c, &c->gb_grid, ngrid))) return st;
    if ((st = pb_reserve(c, &c->gb_spec, nspec))) return st;
    {
        int *cnt = (int *)c->gb_int.ptr, *ip = cnt + ncount;
        double4 *dp = (double4 *)c->gb_d4.ptr;
        int2 *tp = (int2 *)c->gb_tiles.ptr;
        char *gp = (char *)c->gb_grid.ptr, *sp = (char *)c->gb_spec.ptr;
        for (int q = 0; q < n; q++) {
            pb_sys &S = sys[q];
            const int nc = S.cells.ncell, nk = S.cells.nkey_t;
            S.scount = cnt, cnt += pad16(nc + 1);
            S.tcount = cnt, cnt += pad16(nk + 1);
            S.sstart = ip, ip += pad16(nc + 1);
            S.skey = ip, ip += pad16(S.ns);
            S.srank = ip, ip += pad16(S.ns);
            S.sperm = ip, ip += pad16(S.ns);
            S.tstart = ip, ip += pad16(nk + 1);
            S.tkey = ip, ip += pad16(S.nt);
            S.trank = ip, ip += pad16(S.nt);
            S.tperm = ip, ip += pad16(S.nt);
            S.ntiles = ip, ip += 16;
            S.spos = dp, dp += pad16(S.ns);
            S.sfrc = dp, dp += pad16(S.ns);
            S.tpos = dp, dp += pad16(S.nt);
            S.tiles = tp, tp += pad16(mt[q]);
            const size_t cb = params[idx[q]].precision == PB_PREC_MIXED ? 8 : 16;
            S.grid4 = gp, gp += pad256((size_t)S.g.ng * 2 * cb);
            S.spec = sp, sp += pad256(3 * (size_t)S.g.n[0] * S.g.n[1] * (S.g.n[2] / 2 + 1) * cb);
        }
    }
    // launches in phase order, each with its map
    std::vector<int2> map;
    auto add = [&](int kind, int a, int b, int cc, size_t lds, const std::vector<int2> &m) {
        if (m.empty()) return;
        c->group_launches.push_back({kind, a, b, cc, lds, map.size(), (int)m.size()});
        map.insert(map.end(), m.begin(), m.end());
    };
    {
        std::vector<int2> m, one;
        for (int q = 0; q < n; q++) {
            const long long npts = (long long)sys[q].ns + sys[q].nt;
            for (int b = 0; (long long)b * 256 < npts; b++) m.push_back(make_int2(q, b));
            one.push_back(make_int2(q, 0));
        }
        add(0, 0, 0, 0, 0, m);
        add(1, 0, 0, 0, 0, one);
        add(2, 0, 0, 0, 0, m);
    }
    // pair kernel: classes by (FP32 far field, table degree; 0 = closed forms)
    long long items = 0;
    {
        std::vector<std::pair<int, int>> classes;
        std::vector<int> cls(n, -1);
        for (int q = 0; q < n; q++) {
            const int i = idx[q];
            if (sys[q].nt == 0) continue;
            const int tdeg = c->batch_ntab[i] ? c->batch_tdeg[i] : 0;
            const std::pair<int, int> k(far32_of(&params[i], c->batch_ntab[i]) ? 1 : 0, tdeg);
            auto it = std::find(classes.begin(), classes.end(), k);
            cls[q] = (int)(it - classes.begin());
            if (it == classes.end()) classes.push_back(k);
        }
        std::vector<size_t> stat_off(n, 0);
        for (int k = 0; k < (int)classes.size(); k++) {
            double tiles_est = 0;
            size_t lds = 0;
            for (int q = 0; q < n; q++)
                if (cls[q] == k) {
                    tiles_est += std::max(1.0, sys[q].nt / 48.0);
                    const size_t nt_ = c->batch_ntab[idx[q]];
                    lds = std::max(lds, nt_ * (sizeof(double) + (classes[k].first ? sizeof(float) : 0)));
                }
            // wavefronts per tile: about two rounds of resident wavefronts in total
            const int nsplit = (int)std::min(16.0, std::max(1.0, std::round(2.25 * 16 * c->cu_count / tiles_est)));
            std::vector<int2> m;
            for (int q = 0; q < n; q++)
                if (cls[q] == k) {
                    const long long blocks = ((long long)mt[q] * nsplit + pb_pair_waves - 1) / pb_pair_waves;
                    stat_off[q] = (size_t)items;
                    items += blocks * pb_pair_waves;
                    for (long long b = 0; b < blocks; b++) m.push_back(make_int2(q, (int)b));
                }
            add(3, classes[k].first, classes[k].second, nsplit, lds, m);
        }
        if ((st = pb_reserve(c, &c->gb_stats, sizeof(uint4) * std::max<long long>(items, 1)))) return st;
        for (int q = 0; q < n; q++) sys[q].stats = (uint4 *)c->gb_stats.ptr + stat_off[q];
    }
    c->group_items = items;
    // k space: spread, the five FFT passes, gather; classes by precision
    // (and transformed length, window)
    auto prec_of = [&](int q) { return params[idx[q]].precision == PB_PREC_MIXED ? PB_PREC_MIXED : PB_PREC_FP64; };
    const int precs[2] = {PB_PREC_MIXED, PB_PREC_FP64};
    for (int pr : precs) {
        std::vector<int2> m;
        for (int q = 0; q < n; q++)
            if (prec_of(q) == pr) {
                const pb_kgrid &g = sys[q].g;
                const int nreg = ((g.n[0] + 7) / 8) * ((g.n[1] + 7) / 8) * ((g.n[2] + 7) / 8);
                for (int b = 0; b < nreg; b++) m.push_back(make_int2(q, b));
            }
        add(4, pr, 0, 0, 0, m);
    }
    // FFT stages: stage k runs pass k of the five-pass systems and, at stages
    // 0, 2 and 4, passes 5, 2 and 6 of the systems with fused plane passes
    std::vector<int> fused(n);
    for (int q = 0; q < n; q++) {
        int ps[5];
        fused[q] = pb_fft_passes(prec_of(q), sys[q].g, ps) == 3;
    }
    for (int stage = 0; stage < 5; stage++) {
        const int fpass = stage == 0 ? 5 : (stage == 2 ? 2 : (stage == 4 ? 6 : -1));
        for (int pr : precs)
            for (int v = 0; v < (fpass >= 0 && fpass != stage ? 2 : 1); v++) {
                const int pass = v ? fpass : stage;
                const int ax = pb_fft_axis(pass);
                for (int len = 4; len <= 512; len *= 2) {
                    std::vector<int2> m;
                    for (int q = 0; q < n; q++)
                        if (prec_of(q) == pr && sys[q].g.n[ax] == len &&
                            (fused[q] ? pass == fpass : pass == stage)) {
                            const int blocks = pb_group_fft_blocks(pr, pass, sys[q].g);
                            for (int b = 0; b < blocks; b++) m.push_back(make_int2(q, b));
                        }
                    add(5, pr, pass, len, 0, m);
                }
            }
    }
    {
        const int Ps[9] = {4, 6, 8, 10, 12, 16, 20, 24, 0};
        for (int pr : precs)
            for (int P : Ps) {
                std::vector<int2> m;
                for (int q = 0; q < n; q++) {
                    const pb_kgrid &g = sys[q].g;
                    const bool same = g.P[0] == g.P[1] && g.P[1] == g.P[2];
                    const int Pq = same && std::find(Ps, Ps + 8, g.P[0]) != Ps + 8 ? g.P[0] : 0;
                    if (prec_of(q) != pr || Pq != P || sys[q].nt == 0) continue;
                    const int per = P ? pb_interp_waves * pb_interp_tw(P) : 256;
                    for (int b = 0; b * per < sys[q].nt; b++) m.push_back(make_int2(q, b));
                }
                if (P) // the gather renumbers blocks per XCD: a multiple of 8
                    while (m.size() % 8) m.push_back(make_int2(-1, 0));
                add(6, pr, P, 0, 0, m);
            }
    }
    c->group_counts = ncount;
    if ((st = pb_reserve(c, &c->gb_sys, sizeof(pb_sys) * n))) return st;
    if ((st = pb_reserve(c, &c->gb_map, sizeof(int2) * std::max<size_t>(map.size(), 1)))) return st;
    if ((st = pb_hip_check(c, hipMemcpyAsync(c->gb_sys.ptr, sys.data(), sizeof(pb_sys) * n, hipMemcpyHostToDevice, s),
                           "group upload")))
        return st;
    if (!map.empty() &&
        (st = pb_hip_check(c, hipMemcpyAsync(c->gb_map.ptr, map.data(), sizeof(int2) * map.size(),
                                             hipMemcpyHostToDevice, s), "group upload")))
        return st;
    for (int pr : precs) {
        const void *tw;
        bool used = false;
        for (int q = 0; q < n; q++) used = used || prec_of(q) == pr;
        if (used && (st = pb_twiddles(c, s, pr, &tw))) return st;
    }
    if ((st = pb_hip_check(c, hipStreamSynchronize(s), "group upload"))) return st;
    c->group_key = key;
    return PB_OK;
}

// The group's launches on stream g (pair kernels forked to two pair streams
// before the spread and joined back); u of the group's targets must be zero.
int group_enqueue(pb_context *c, hipStream_t g)
{
    int st;
    const pb_sys *sys = (const pb_sys *)c->gb_sys.ptr;
    const int2 *map = (const int2 *)c->gb_map.ptr;
    pb_fill_zero(g, (long long)c->group_counts, (int *)c->gb_int.ptr);
    bool forked = false;
    int npair = 0;
    // the pair-kernel classes, forked to the pair streams at the fork point
    auto launch_pairs = [&]() -> int {
        int q;
        for (const auto &L : c->group_launches) {
            if (L.kind != 3) continue;
            if (!forked) {
                if (!c->pair_stream &&
                    (q = pb_hip_check(c, hipStreamCreateWithFlags(&c->pair_stream, hipStreamNonBlocking),
                                      "stream")))
                    return q;
                for (int k = 0; k < 2; k++)
                    if (!c->pair_event[k] &&
                        (q = pb_hip_check(c, hipEventCreateWithFlags(&c->pair_event[k], hipEventDisableTiming),
                                          "event")))
                        return q;
                if (!c->pair_stream2 &&
                    (q = pb_hip_check(c, hipStreamCreateWithFlags(&c->pair_stream2, hipStreamNonBlocking),
                                      "stream")))
                    return q;
                if (!c->pair_event2 &&
                    (q = pb_hip_check(c, hipEventCreateWithFlags(&c->pair_event2, hipEventDisableTiming), "event")))
                    return q;
                if ((q = pb_hip_check(c, hipEventRecord(c->pair_event[0], g), "fork"))) return q;
                if ((q = pb_hip_check(c, hipStreamWaitEvent(c->pair_stream, c->pair_event[0], 0), "fork")))
                    return q;
                if ((q = pb_hip_check(c, hipStreamWaitEvent(c->pair_stream2, c->pair_event[0], 0), "fork")))
                    return q;
                forked = true;
            }
            // the classes alternate between the two pair streams, so they run
            // side by side (and beside the k space on g)
            if ((q = pb_group_pair(npair++ % 2 ? c->pair_stream2 : c->pair_stream, sys, map + L.map, L.nblk, L.c,
                                   L.a != 0, L.b, L.lds)))
                return pb_fail(c, q, "grouped pair kernel launch");
        }
        return PB_OK;
    };
    // fork before the first spread launch (the 512-system call: mixed 8.40
    // against 8.57 ms with the fork after the last one, closed forms 18.6
    // against 20.7 ms, the FP64 table form unchanged at 16.9 ms)
    int fork_before = -1, i = 0;
    for (const auto &L : c->group_launches) {
        if (L.kind == 4) {
            fork_before = i;
            break;
        }
        i++;
    }
    i = 0;
    for (const auto &L : c->group_launches) {
        if (i++ == fork_before && (st = launch_pairs())) return st;
        const int2 *m = map + L.map;
        switch (L.kind) {
        case 0: pb_group_bin_count(g, sys, m, L.nblk); break;
        case 1: pb_group_bin_scan(g, sys, m, L.nblk); break;
        case 2: pb_group_bin_scatter(g, sys, m, L.nblk); break;
        case 3: break;
        case 4: pb_group_spread(g, sys, m, L.nblk, L.a); break;
        case 5: {
            const void *tw = c->twid[L.a == PB_PREC_MIXED ? 1 : 0].ptr;
            if ((st = pb_group_fft(g, sys, m, L.nblk, L.a, L.b, L.c, tw))) return pb_fail(c, st, "FFT size");
            break;
        }
        default:
            if ((st = pb_group_interp(g, sys, m, L.nblk, L.a, L.b))) return pb_fail(c, st, "gather window");
            break;
        }
    }
    if (!forked && (st = launch_pairs())) return st; // (no launch at or after the fork point)
    if (forked) {
        if ((st = pb_hip_check(c, hipEventRecord(c->pair_event[1], c->pair_stream), "join"))) return st;
        if ((st = pb_hip_check(c, hipStreamWaitEvent(g, c->pair_event[1], 0), "join"))) return st;
        if ((st = pb_hip_check(c, hipEventRecord(c->pair_event2, c->pair_stream2), "join"))) return st;
        if ((st = pb_hip_check(c, hipStreamWaitEvent(g, c->pair_event2, 0), "join"))) return st;
    }
    return pb_hip_check(c, hipGetLastError(), "grouped launch");
}

} // namespace

int pb_velocity_batched(pb_context *c, int nsys, pb_params *params, const int *src_off, const double *xn,
                        const double *f, const int *tgt_off, const double *y, double *u, void *stream)
{
    if (!c) return PB_ERR_INVALID;
    if (nsys < 0 || (nsys > 0 && (!params || !src_off || !tgt_off)))
        return pb_fail(c, PB_ERR_INVALID, "null system description");
    if (nsys == 0) return PB_OK;
    if (src_off[0] != 0 || tgt_off[0] != 0) return pb_fail(c, PB_ERR_INVALID, "offsets must start at 0");
    int st;
    for (int i = 0; i < nsys; i++) {
        const int ns = src_off[i + 1] - src_off[i], nt = tgt_off[i + 1] - tgt_off[i];
        if (ns < 0 || nt < 0) return pb_fail(c, PB_ERR_INVALID, "offsets must not decrease");
        pb_params &p = params[i];
        if (!(p.d > 0) || p.grid[0] <= 0 || p.grid[1] <= 0 || p.grid[2] <= 0) {
            // unset split or grid: the autotuner chooses them for this system
            if ((st = pb_autotune(&p, ns, nt, nullptr))) return pb_fail(c, st, "autotune failed for a system");
        }
        if ((st = check_params(c, &p))) return st;
        pb_kgrid g;
        if ((st = make_kgrid(c, &p, &g))) return st; // every system's grid and window, before any launch
    }
    if ((src_off[nsys] > 0 && (!xn || !f)) || (tgt_off[nsys] > 0 && (!y || !u)))
        return pb_fail(c, PB_ERR_INVALID, "null point array");
    if ((st = pb_hip_check(c, hipSetDevice(c->device), "hipSetDevice"))) return st;
    hipStream_t s = (hipStream_t)stream;
    // kernel constants and tables of every system, uploaded together; kept on
    // the device while the systems' parameters do not change
    constexpr int KW = 7; // key words per system
    std::vector<double> key((size_t)nsys * KW);
    for (int i = 0; i < nsys; i++) {
        const pb_params &p = params[i];
        const double k7[KW] = {p.alpha, p.d, p.eps, std::fmax(3 * p.d, 8 * p.eps), p.rc, (double)p.kernel,
                               (double)table_degree(&p)};
        std::copy(k7, k7 + KW, key.begin() + KW * (size_t)i);
    }
    if (key != c->batch_key) {
        std::vector<pb_kparams> kps(nsys);
        std::vector<std::vector<double>> coef(nsys);
        // distinct parameter sets, built on several host threads
        std::vector<int> uniq, rep(nsys);
        for (int i = 0; i < nsys; i++) {
            rep[i] = i;
            for (int j : uniq)
                if (std::equal(key.begin() + KW * (size_t)i, key.begin() + KW * (size_t)i + KW,
                               key.begin() + KW * (size_t)j)) {
                    rep[i] = j;
                    break;
                }
            if (rep[i] == i) uniq.push_back(i);
        }
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; t++)
            pool.emplace_back([&, t]() {
                for (size_t q = t; q < uniq.size(); q += nth) {
                    const int i = uniq[q];
                    host_kparams(&params[i], std::fmax(3 * params[i].d, 8 * params[i].eps), &kps[i], coef[i]);
                }
            });
        for (auto &th : pool) th.join();
        for (int i = 0; i < nsys; i++)
            if (rep[i] != i) {
                kps[i] = kps[rep[i]];
                coef[i] = coef[rep[i]];
            }
        std::vector<size_t> toff(nsys + 1, 0);
        for (int i = 0; i < nsys; i++) toff[i + 1] = toff[i] + coef[i].size();
        const size_t ntot = toff[nsys];
        std::vector<double> coefs;
        coefs.reserve(ntot);
        for (int i = 0; i < nsys; i++) coefs.insert(coefs.end(), coef[i].begin(), coef[i].end());
        std::vector<float> c32(coefs.begin(), coefs.end());
        if ((st = pb_reserve(c, &c->batch_kp, sizeof(pb_kparams) * nsys))) return st;
        if ((st = pb_reserve(c, &c->batch_tab, (ntot ? ntot : 1) * (sizeof(double) + sizeof(float))))) return st;
        const double *tab64 = (const double *)c->batch_tab.ptr;
        const float *tab32 = (const float *)((const char *)c->batch_tab.ptr + ntot * sizeof(double));
        for (int i = 0; i < nsys; i++)
            if (toff[i + 1] > toff[i]) {
                kps[i].tab64 = tab64 + toff[i];
                kps[i].tab32 = tab32 + toff[i];
            }
        if (ntot) {
            if ((st = pb_hip_check(c, hipMemcpyAsync(c->batch_tab.ptr, coefs.data(), ntot * sizeof(double),
                                                     hipMemcpyHostToDevice, s), "table upload")))
                return st;
            if ((st = pb_hip_check(c, hipMemcpyAsync((void *)tab32, c32.data(), ntot * sizeof(float),
                                                     hipMemcpyHostToDevice, s), "table upload")))
                return st;
        }
        if ((st = pb_hip_check(c, hipMemcpyAsync(c->batch_kp.ptr, kps.data(), sizeof(pb_kparams) * nsys,
                                                 hipMemcpyHostToDevice, s), "kparams upload")))
            return st;
        if ((st = pb_hip_check(c, hipStreamSynchronize(s), "batch upload"))) return st;
        c->batch_key = key;
        c->batch_ntab.assign(nsys, 0);
        c->batch_tdeg.assign(nsys, 0);
        for (int i = 0; i < nsys; i++) {
            c->batch_ntab[i] = toff[i + 1] - toff[i];
            c->batch_tdeg[i] = kps[i].tdeg;
        }
    }
    // the systems go to 4 streams, each with its own scratch; largest first,
    // each to the least loaded stream (by grid points plus particles)
    constexpr int K = 4;
    for (int k = 0; k < K; k++) {
        if (!c->sub[k] && (st = pb_create(&c->sub[k], c->device))) return pb_fail(c, st, "sub-context");
        if (!c->sub_stream[k] &&
            (st = pb_hip_check(c, hipStreamCreateWithFlags(&c->sub_stream[k], hipStreamNonBlocking), "stream")))
            return st;
    }
    for (int k = 0; k < K + 3; k++)
        if (!c->sub_event[k] &&
            (st = pb_hip_check(c, hipEventCreateWithFlags(&c->sub_event[k], hipEventDisableTiming), "event")))
            return st;
    auto work = [&](int i) {
        return (double)params[i].grid[0] * params[i].grid[1] * params[i].grid[2] + 4.0 * (src_off[i + 1] - src_off[i]) +
               4.0 * (tgt_off[i + 1] - tgt_off[i]);
    };
    // systems whose binning fits one scan workgroup form the group (one launch
    // per phase for all of them); the others run one by one on the streams
    std::vector<int> grouped, order;
    for (int i = 0; i < nsys; i++) {
        pb_cells cl;
        if ((st = make_cells(c, &params[i], &cl))) return st;
        (pb_bin_fused_ok(cl) ? grouped : order).push_back(i);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    double load[K] = {0, 0, 0, 0};
    std::vector<int> lane_of(nsys);
    for (int i : order) {
        int k = 0;
        for (int q = 1; q < K; q++)
            if (load[q] < load[k]) k = q;
        lane_of[i] = k;
        load[k] += work(i);
    }
    // Everything below runs on an internal stream g ordered after s (events),
    // fanned out to the four system streams and joined back.  A call that
    // repeats the previous one (same systems, parameters, arrays and stream)
    // replays the graph captured from it: one launch instead of some twenty
    // per system.  The first of two identical calls runs eagerly (it sizes
    // every buffer); the second captures, and both leave s ordered after the
    // work.
    if (!c->cap_stream &&
        (st = pb_hip_check(c, hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking), "stream")))
        return st;
    hipStream_t g = c->cap_stream;
    std::vector<double> gkey = key;
    for (int i = 0; i <= nsys; i++) {
        gkey.push_back(src_off[i]);
        gkey.push_back(tgt_off[i]);
    }
    const double ptrs[5] = {(double)(uintptr_t)xn, (double)(uintptr_t)f, (double)(uintptr_t)y,
                            (double)(uintptr_t)u, (double)(uintptr_t)s};
    gkey.insert(gkey.end(), ptrs, ptrs + 5);
    if (!grouped.empty() && (st = group_prepare(c, g, grouped, params, src_off, tgt_off, xn, f, y, u, gkey)))
        return st;
    c->stat_group = !grouped.empty();
    c->stat_items = grouped.empty() ? 0 : c->group_items;
    if ((st = pb_hip_check(c, hipEventRecord(c->sub_event[K], s), "order"))) return st;
    if ((st = pb_hip_check(c, hipStreamWaitEvent(g, c->sub_event[K], 0), "order"))) return st;
    auto finish = [&]() -> int {
        int q;
        if ((q = pb_hip_check(c, hipEventRecord(c->sub_event[K + 2], g), "order"))) return q;
        return pb_hip_check(c, hipStreamWaitEvent(s, c->sub_event[K + 2], 0), "order");
    };
    if (c->graph_exec && gkey == c->graph_key) {
        if ((st = pb_hip_check(c, hipGraphLaunch(c->graph_exec, g), "graph launch"))) return st;
        return finish();
    }
    const bool capture = gkey == c->last_batch_key;
    c->last_batch_key = gkey;
    if (c->graph_exec) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
        c->graph_key.clear();
    }
    if (capture && (st = pb_hip_check(c, hipStreamBeginCapture(g, hipStreamCaptureModeThreadLocal), "capture")))
        return st;
    // the group adds into zeroed u; the systems on the streams overwrite theirs
    if (!grouped.empty()) pb_zero_u(g, tgt_off[nsys], u);
    // fork
    if ((st = pb_hip_check(c, hipEventRecord(c->sub_event[K + 1], g), "fork"))) return st;
    for (int k = 0; k < K; k++)
        if ((st = pb_hip_check(c, hipStreamWaitEvent(c->sub_stream[k], c->sub_event[K + 1], 0), "fork")))
            return st;
    const pb_kparams *kpd = (const pb_kparams *)c->batch_kp.ptr;
    int est = PB_OK;
    for (int i : order) {
        const int ns = src_off[i + 1] - src_off[i], nt = tgt_off[i + 1] - tgt_off[i];
        pb_context *sc = c->sub[lane_of[i]];
        if ((est = enqueue_velocity(sc, c->sub_stream[lane_of[i]], &params[i], kpd + i, c->batch_ntab[i],
                                    c->batch_tdeg[i], ns,
                                    xn + 3 * (size_t)src_off[i], f + 3 * (size_t)src_off[i], nt,
                                    y + 3 * (size_t)tgt_off[i], u + 3 * (size_t)tgt_off[i], true, true, false))) {
            est = pb_fail(c, est, sc->err);
            break;
        }
    }
    if (!est && !grouped.empty()) est = group_enqueue(c, g);
    // join (also after a failure, so a capture can be closed)
    for (int k = 0; k < K; k++) {
        if ((st = pb_hip_check(c, hipEventRecord(c->sub_event[k], c->sub_stream[k]), "join"))) return st;
        if ((st = pb_hip_check(c, hipStreamWaitEvent(g, c->sub_event[k], 0), "join"))) return st;
    }
    if (capture) {
        hipGraph_t graph = nullptr;
        st = pb_hip_check(c, hipStreamEndCapture(g, &graph), "end capture");
        if (st || est) {
            if (graph) (void)hipGraphDestroy(graph);
            c->last_batch_key.clear();
            return est ? est : st;
        }
        st = pb_hip_check(c, hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0), "instantiate");
        (void)hipGraphDestroy(graph);
        if (st) {
            c->graph_exec = nullptr;
            return st;
        }
        c->graph_key = gkey;
        if ((st = pb_hip_check(c, hipGraphLaunch(c->graph_exec, g), "graph launch"))) return st;
    } else if (est) {
        return est;
    }
    return finish();
}

int pb_last_pair_stats(pb_context *c, long long out[4])
{
    if (!c || !out) return PB_ERR_INVALID;
    out[0] = out[1] = out[2] = out[3] = 0;
    const pb_buffer &b = c->stat_group ? c->gb_stats : c->counters;
    if (!b.ptr || c->stat_items <= 0) return PB_OK;
    std::vector<unsigned> h(4 * (size_t)c->stat_items);
    int st = pb_hip_check(c, hipMemcpy(h.data(), b.ptr, h.size() * sizeof(unsigned),
                                       hipMemcpyDeviceToHost), "counters");
    for (size_t i = 0; i < h.size(); i++) out[i & 3] += h[i];
    return st;
}

int pb_last_pair_counts(pb_context *c, long long *n64, long long *n32)
{
    if (!c || !n64 || !n32) return PB_ERR_INVALID;
    long long s[4];
    int st = pb_last_pair_stats(c, s);
    *n64 = s[0];
    *n32 = s[1];
    return st;
}

int periodicBrinkman(double *dA, int lda, const double *dx, int nt, const double *dx0, int ns,
                     double a, double d, double e, int M, int N, const double *L, int maxBlockSize)
{
    (void)maxBlockSize;
    if (!L || (nt > 0 && (!dA || !dx)) || (ns > 0 && !dx0)) return PB_ERR_INVALID;
    if (nt < 0 || ns < 0 || lda < 3 * nt) return PB_ERR_INVALID; // periodicBrinkman.h:443-445
    if (!(a > 0) || !(d > 0) || !(e >= 0)) return PB_ERR_INVALID;
    for (int i = 0; i < 3; i++)
        if (!(L[i] > 0)) return PB_ERR_INVALID;
    // the reference's shell counts: sort L, real shells ceil(M Lmin/Li),
    // Fourier shells ceil(N Lmax/Li) (periodicBrinkman.h:457-487)
    int idL[3] = {0, 1, 2}, tmp;
    if (L[idL[0]] > L[idL[1]]) { tmp = idL[0]; idL[0] = idL[1]; idL[1] = tmp; }
    if (L[idL[2]] <= L[idL[0]]) {
        tmp = idL[2]; idL[2] = idL[1]; idL[1] = idL[0]; idL[0] = tmp;
    } else if (L[idL[2]] < L[idL[1]]) {
        tmp = idL[1]; idL[1] = idL[2]; idL[2] = tmp;
    }
    int Mx[3] = {M, M, M}, Nx[3] = {N, N, N};
    Mx[idL[2]] = (int)std::ceil(M * L[idL[0]] / L[idL[2]]);
    Mx[idL[1]] = (int)std::ceil(M * L[idL[0]] / L[idL[1]]);
    Nx[idL[0]] = (int)std::ceil(N * L[idL[2]] / L[idL[0]]);
    Nx[idL[1]] = (int)std::ceil(N * L[idL[2]] / L[idL[1]]);
    pb_box box;
    pb_fill_box(&box, L);
    const pb_kparams &kp = cached_kparams(a, d, e, 0.0);
    int zero[3] = {0, 0, 0};
    int st = PB_OK;
    if (nt == 0 || ns == 0) return PB_OK;
    pb_context *c = default_ctx();
    if (!c) return PB_ERR_HIP;
    // both launches on the null stream with no host synchronization, as the
    // reference's (periodicBrinkman.h:500,511)
    if (N > 0) st = pb_run_dense(c, nullptr, dA, lda, dx, nt, dx0, ns, kp, box, nullptr, nullptr, 0, zero, Nx, 1, nullptr);
    if (st) return st;
    if (M >= 0) st = pb_run_dense(c, nullptr, dA, lda, dx, nt, dx0, ns, kp, box, zero, Mx, 3, nullptr, nullptr, 0, nullptr);
    return st;
}

int realShells(double *A, double *absA, int lda, const double *x, int nt, const double *x0, int ns,
               const int *newM, const int *oldM, const double *L, double a, double d, double e)
{
    return shells_impl(A, absA, lda, x, nt, x0, ns, newM, oldM, L, a, d, e, false);
}

int fourierShells(double *A, double *absA, int lda, const double *x, int nt, const double *x0,
                  int ns, const int *newM, const int *oldM, const double *L, double a, double d,
                  double e)
{
    return shells_impl(A, absA, lda, x, nt, x0, ns, newM, oldM, L, a, d, e, true);
}

int pb_split_sweep(const pb_driver_params *p, const double *x, int nt, double dval[8], double err[8])
{
    if (!p || !dval || !err || nt < 1 || !x || p->M < 0 || !(p->d > 0) || !(p->alpha > 0) || !(p->e >= 0))
        return PB_ERR_INVALID;
    const double L[3] = {p->Lx, p->Ly, p->Lz};
    const double x0[3] = {0.5, 0.5, 0.5}; // the driver's single source (periodicBrinkman.cu:114)
    const int lda = 3 * nt;
    const size_t n = (size_t)lda * 3;
    const int zero[3] = {0, 0, 0}, Ms[3] = {p->M, p->M, p->M};
    // A = realShells(home) + realShells(layers 1..M) + fourierShells(1..M)
    auto matrix = [&](double d, std::vector<double> &A) -> int {
        std::vector<double> absA(n, 0.0);
        A.assign(n, 0.0);
        int st;
        if ((st = realShells(A.data(), absA.data(), lda, x, nt, x0, 1, zero, zero, L, p->alpha, d, p->e))) return st;
        if ((st = realShells(A.data(), absA.data(), lda, x, nt, x0, 1, Ms, zero, L, p->alpha, d, p->e))) return st;
        return fourierShells(A.data(), absA.data(), lda, x, nt, x0, 1, Ms, zero, L, p->alpha, d, p->e);
    };
    std::vector<double> A0, A;
    int st;
    if ((st = matrix(p->d, A0))) return st;
    for (int i = 0; i < 8; i++) {
        dval[i] = (0.1 + i * 0.95) * p->d;
        if ((st = matrix(dval[i], A))) return st;
        double m = 0;
        for (size_t q = 0; q < n; q++) m = std::fmax(m, std::fabs(A[q] - A0[q]));
        err[i] = m;
    }
    return PB_OK;
}

int pb_vR_batch(double *A, const double *xh, int n, double a, double d, double e, void *stream)
{
    if (n < 0 || (n > 0 && (!A || !xh)) || !(a > 0) || !(d > 0) || !(e >= 0)) return PB_ERR_INVALID;
    return pb_run_vR_batch((hipStream_t)stream, A, xh, n, cached_kparams(a, d, e, 0.0));
}

int pb_vF_batch(double *A, const double *xh, const double *k, int n, double a, double d, void *stream)
{
    if (n < 0 || (n > 0 && (!A || !xh || !k)) || !(a > 0) || !(d > 0)) return PB_ERR_INVALID;
    return pb_run_vF_batch((hipStream_t)stream, A, xh, k, n, a, d);
}

int vR(double *A, const double *xh, double a, double d, double e)
{
    if (!A || !xh) return PB_ERR_INVALID;
    dev_tmp t;
    double *dA = t.alloc(9), *dx = t.alloc(3);
    if (!dA || !dx) return PB_ERR_NOMEM;
    int st = copy(dx, xh, 3 * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = pb_vR_batch(dA, dx, 1, a, d, e, nullptr);
    if (!st) st = copy(A, dA, 9 * sizeof(double), hipMemcpyDeviceToHost);
    return st;
}

int vF(double *A, const double *xh, double a, double d, const double *k)
{
    if (!A || !xh || !k) return PB_ERR_INVALID;
    dev_tmp t;
    double *dA = t.alloc(9), *dx = t.alloc(3), *dk = t.alloc(3);
    if (!dA || !dx || !dk) return PB_ERR_NOMEM;
    int st = copy(dx, xh, 3 * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(dk, k, 3 * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = pb_vF_batch(dA, dx, dk, 1, a, d, nullptr);
    if (!st) st = copy(A, dA, 9 * sizeof(double), hipMemcpyDeviceToHost);
    return st;
}

int fftBrinkman(double *U, const double *y, int m, const double *f, const double *xn, int N,
                const double *L, const int *M, const int *P, int msd, int C, double e, double tol,
                double d, double alpha)
{
    (void)C;
    (void)e;
    (void)tol;
    if (!L || !M || !P || m < 0 || N < 0 || (m > 0 && (!U || !y)) || (N > 0 && (!f || !xn)))
        return PB_ERR_INVALID;
    if (!(d > 0) || !(alpha > 0) || msd <= 0) return PB_ERR_INVALID;
    pb_context *c = default_ctx();
    if (!c) return PB_ERR_HIP;
    if (hipSetDevice(c->device) != hipSuccess) return PB_ERR_HIP;
    pb_params p;
    pb_params_default(&p);
    for (int i = 0; i < 3; i++) {
        p.L[i] = L[i];
        p.grid[i] = M[i];
        if (!(L[i] > 0) || M[i] < 2 || P[i] < 1) return PB_ERR_INVALID;
    }
    p.alpha = alpha;
    p.d = d;
    p.m = msd;
    p.P = P[0]; // eta follows axis 0 (fftBrinkman.h:113-114)
    p.rc = std::fmin(L[0], std::fmin(L[1], L[2])); // binning only
    pb_kgrid g;
    int st;
    if ((st = make_kgrid(c, &p, &g, P))) return st; // grid and windows checked before any copy
    dev_tmp t;
    double *dy = t.alloc(3 * (size_t)m), *df = t.alloc(3 * (size_t)N), *dx = t.alloc(3 * (size_t)N),
           *du = t.alloc(3 * (size_t)m);
    if (!dy || !df || !dx || !du) return PB_ERR_NOMEM;
    st = copy(dy, y, 3 * (size_t)m * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(df, f, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice);
    if (!st) st = copy(dx, xn, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice);
    if (st) return st;
    pb_box box;
    pb_fill_box(&box, p.L);
    pb_cells cells;
    std::memset(&cells, 0, sizeof(cells));
    for (int i = 0; i < 3; i++) {
        cells.nc[i] = 1;
        cells.h[i] = L[i];
        cells.inv_h[i] = 1.0 / L[i];
        cells.st[i] = 1;
    }
    cells.ncell = 1;
    cells.nb[0] = cells.nb[1] = cells.nb[2] = 1;
    cells.nkey_t = 64;
    if ((st = pb_bin_points(c, nullptr, cells, box, N, dx, df, true))) return st;
    if ((st = pb_bin_points(c, nullptr, cells, box, m, dy, nullptr, false))) return st;
    if ((st = pb_run_kspace(c, nullptr, g, cells, box, alpha, d, N, m, du, 0, PB_PREC_FP64))) return st;
    if (hipDeviceSynchronize() != hipSuccess) return PB_ERR_HIP;
    return copy(U, du, 3 * (size_t)m * sizeof(double), hipMemcpyDeviceToHost);
}

} // extern "C"
