// Host side of libdpca: context, PLINK input, Omega, the subspace-iteration
// driver and the C ABI of include/dpca.h.
//
// The driver mirrors BlockSubspaceIter_MPI_OOC_double_buffering
// (src/methods.cpp:702-1135) for one rank, with every per-pass and
// per-iteration step on the device: the fused decode + product kernels for
// A A^T Q, CholQR2 instead of dgeqrf/dorgqr, a device Jacobi eigensolver
// instead of dsyev, and the stopping test.  The host only issues launches and
// reads a few bytes of criterion state once per outer iteration.
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "dpca.h"
#include "dpca_gen.h"
#include "dpca_internal.h"
#include "dpca_model.h"

namespace dpca {

static thread_local char g_err[2048] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static int host_threads()
{
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 4;
    return h > 16 ? 16 : (int)h;
}

// ------------------------------------------------------------------ Omega --
// Park-Miller minimal standard generator with Schrage's factorisation and the
// one-sided Box-Muller transform of src/gennorm.c:21-72, driven in the order
// of src/distpca.cpp:389-394.  A positive seed restarts the state at the seed;
// a non-positive one (the reference's int seed jj*M + kk wraps for
// M >= 46341) advances it one step, so elements are independent from each
// positive seed on: the matrix is cut into chunks that start at positive
// seeds and the chunks are generated in parallel.
namespace omega {
struct Lcg {
    long x = 0;
    double next()
    {
        const long q = 127773, r = 2836, a = 16807, m = 2147483647;
        const long hi = x / q, lo = x % q;
        const long t = a * lo - r * hi;
        x = (t > 0) ? t : t + m;
        return (double)x / m;
    }
};

static inline int32_t seed_of(int64_t e, int M, int k)
{
    const uint32_t jj = (uint32_t)(e / k), kk = (uint32_t)(e % k);
    return (int32_t)(jj * (uint32_t)M + kk);
}

static void gen_range(int64_t e0, int64_t e1, int M, int k, long state, double *out)
{
    // no fused multiply-adds (the reference's IEEE operation order): g++ gets
    // -ffp-contract=off from the build, clang this pragma
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    Lcg g;
    g.x = state;
    const double pi = 3.14159265;
    for (int64_t e = e0; e < e1; e++) {
        const int32_t s = seed_of(e, M, k);
        if (s > 0)
            g.x = s;
        else
            g.next();
        double u = 0.0;
        while (u == 0.0) u = g.next();
        const double rr = sqrt(-2.0 * log(u));
        double th = 0.0;
        while (th == 0.0) th = 2.0 * pi * g.next();
        const double x = rr * cos(th);
        out[e] = (x * 1.0) + 0.0;
    }
}
}  // namespace omega

int omega_generate(int M, int k, double *out)
{
    const int64_t n = (int64_t)M * k;
    const int T = host_threads();
    if (n < 65536 || T == 1) {
        omega::gen_range(0, n, M, k, 0, out);
        return 0;
    }
    // chunk starts: first element at/after an even split whose seed is > 0
    std::vector<int64_t> starts{0};
    for (int t = 1; t < T; t++) {
        int64_t e = n * t / T;
        while (e < n && omega::seed_of(e, M, k) <= 0) e++;
        if (e > starts.back() && e < n) starts.push_back(e);
    }
    starts.push_back(n);
    std::vector<std::thread> th;
    for (size_t c = 0; c + 1 < starts.size(); c++)
        th.emplace_back([=] { omega::gen_range(starts[c], starts[c + 1], M, k, 0, out); });
    for (auto &x : th) x.join();
    // chunk 0 starts at element 0 with the program-start state 0, exactly as
    // the reference; every later chunk starts at a positive seed, so the
    // state it was given is overwritten before use.
    return 0;
}

}  // namespace dpca

using namespace dpca;

// ----------------------------------------------------------------- context --
struct Timer {
    int kind;
    hipEvent_t a, b;
};
enum { T_LOAD, T_MM_A, T_MM_AT, T_MM, T_GS, T_RR, T_OTHER, T_H2D, T_NKIND };
// a graph-replayed iteration: its time is split over the kinds above in the
// proportions measured on the eager first iteration
constexpr int T_GRAPH = T_NKIND;

// One steady-state outer iteration captured as a HIP graph.  The buffers Q,
// tmp and ARHS trade places inside an iteration, so a graph is keyed by the
// pointer state it starts from (`in`) and leaves (`out`).
struct IterGraph {
    double *in[3], *out[3];
    hipGraphExec_t exec;
    int passes;
};

struct dpca_ctx {
    dpca_params p{};
    int dev = 0;
    hipStream_t st = nullptr;
    int M = 0;
    int64_t N = 0;
    int k = 0, kp = 0, nsv = 0, bsize = 0;
    int64_t np = 0, ldp = 0, Mpad = 0, Npad = 0;
    bool incore = false;
    std::vector<std::string> iids;
    // device buffers
    uint8_t *packed = nullptr;        // Npad x ldp
    uint8_t *zt_img = nullptr;        // its tile-order image for Z^T T (mixed, resident)
    double *zval = nullptr, *mu = nullptr, *dsc = nullptr;
    int64_t *ssum = nullptr, *sctr = nullptr;
    double *Q = nullptr, *ARHS = nullptr, *RHS2 = nullptr, *RHS2old = nullptr, *tmp = nullptr;
    double *T = nullptr;              // Npad x kp
    double *part = nullptr;           // splits x Mpad x kp
    int splits = 1;
    double *gpart = nullptr, *G = nullptr, *Rinv = nullptr, *w = nullptr, *X = nullptr,
           *dots = nullptr;
    int ngparts = 256;
    // mixed form
    uint4 *UW8 = nullptr, *Q8 = nullptr;
    unsigned *smax = nullptr;         // per (SNP split, column) max |U|, |W|
    unsigned *qmax = nullptr;         // per (sample segment, column) max |Q|
    float *Tpart = nullptr, *Apart = nullptr;
    int nseg = 1, st_tiles = 8, nsplit = 1, sa_tiles = 4;
    int ns = 3;                       // int8 slices of the FP64 operand
    int za_waves = 8, zt_waves = 8;   // waves per int8 product workgroup
    int64_t seg_bytes = 0, rps = 0;
    int nseg64 = 1;                   // FP64 Z Q sample segments
    int64_t seg64 = 0;
    double *Tseg = nullptr;           // their partials (nseg64 > 1)
    // streaming from host memory or the .bed file (matrix above device_cap)
    bool streamed = false;
    int hmode = DPCA_HOST_PINNED;     // where the streamed matrix lives
    uint8_t *hpacked = nullptr;       // host copy, Npad x ldp, device layout
    uint8_t *bounce[2] = {nullptr, nullptr};  // pinned staging (pageable / file)
    uint8_t *draw[2] = {nullptr, nullptr};    // file mode: raw .bed rows on the device
    hipEvent_t ev_bounce[2] = {nullptr, nullptr};  // last copy out of bounce[b]
    int fd = -1;                      // file mode: the .bed file
    uint8_t *dbuf[2] = {nullptr, nullptr};
    int64_t chunk_rows = 0;           // rows per streamed block (multiple of 1024)
    uint8_t *resident = nullptr;      // streamed: leading blocks kept in HBM
    int64_t res_blocks = 0;           // how many (filled by the first pass)
    bool res_filled = false;
    int64_t block_rows = 0;           // rows the per-block buffers are sized for
    hipStream_t cst = nullptr;        // copy stream
    hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    bool stats_done = false;
    double t_h2d = 0, bytes_h2d = 0, t_stream_pass = 0;
    int *flags = nullptr;             // [0] shifted used, [1] QR failure, [2] Jacobi
    double *cinfo = nullptr;          // dependent columns of the last mode-0 factor
    bool robust = false;              // orthonormalise completes rank-deficient bases
    int qr_repairs = 0;               // bases completed in the last run
    IterState *its = nullptr;
    std::vector<double> crit_hist;
    // timers
    std::vector<Timer> pending;
    std::vector<hipEvent_t> free_ev;
    double t[T_NKIND] = {};
    double t_generate_rhs = 0;
    int passes = 0;
    // iterations after the first replay captured graphs (resident, non-robust
    // contexts)
    std::vector<IterGraph> graphs;
    bool capturing = false, graphs_off = false;
    int graph_its = 0;                // iterations replayed in the last run
    double iter_share[T_NKIND] = {};
};

static hipEvent_t get_event(dpca_ctx *c)
{
    if (!c->free_ev.empty()) {
        hipEvent_t e = c->free_ev.back();
        c->free_ev.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct Stage {
    dpca_ctx *c;
    hipStream_t s;
    Timer tm;
    Stage(dpca_ctx *ctx, int kind, hipStream_t on = nullptr) : c(ctx), s(on ? on : ctx->st)
    {
        tm.kind = kind;
        tm.a = tm.b = nullptr;
        if (c->capturing) return;  // a graph being captured is timed as a whole
        tm.a = get_event(c);
        tm.b = get_event(c);
        if (tm.a) (void)hipEventRecord(tm.a, s);
    }
    ~Stage()
    {
        if (c->capturing) return;
        if (tm.b) (void)hipEventRecord(tm.b, s);
        c->pending.push_back(tm);
    }
};

// Call after a stream synchronisation: fold the recorded stage times in.
static void collect_timers(dpca_ctx *c)
{
    for (auto &tm : c->pending) {
        float ms = 0;
        if (tm.a && tm.b && hipEventElapsedTime(&ms, tm.a, tm.b) == hipSuccess) {
            if (tm.kind == T_GRAPH) {
                for (int i = 0; i < T_NKIND; i++) c->t[i] += ms * 1e-3 * c->iter_share[i];
            } else {
                c->t[tm.kind] += ms * 1e-3;
            }
        }
        if (tm.a) c->free_ev.push_back(tm.a);
        if (tm.b) c->free_ev.push_back(tm.b);
    }
    c->pending.clear();
}

// When set, dalloc only adds up sizes (footprint estimate of a context).
static thread_local size_t *g_dry = nullptr;
// Contexts opened by this thread are serving jobs (see dpca::set_job_contexts).
static thread_local bool g_job_ctx = false;

namespace dpca {
void set_job_contexts(bool on) { g_job_ctx = on; }
}  // namespace dpca

template <class P>
static int dalloc(P **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (g_dry) {
        *g_dry += bytes;
        return 0;
    }
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return -3;
    }
    return 0;
}

static void free_ctx(dpca_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->dev);
    void *ptrs[] = {c->packed, c->zval, c->mu, c->dsc, c->ssum, c->sctr, c->Q, c->ARHS,
                    c->RHS2, c->RHS2old, c->tmp, c->T, c->part, c->gpart, c->G, c->Rinv,
                    c->w, c->X, c->dots, c->flags, c->cinfo, c->its, c->UW8, c->Q8, c->qmax, c->smax, c->Tpart, c->Apart, c->Tseg,
                    c->zt_img};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->resident) (void)hipFree(c->resident);
    for (int b = 0; b < 2; b++) {
        if (c->dbuf[b]) (void)hipFree(c->dbuf[b]);
        if (c->draw[b]) (void)hipFree(c->draw[b]);
        if (c->bounce[b]) (void)hipHostFree(c->bounce[b]);
        if (c->ev_copied[b]) (void)hipEventDestroy(c->ev_copied[b]);
        if (c->ev_free[b]) (void)hipEventDestroy(c->ev_free[b]);
        if (c->ev_bounce[b]) (void)hipEventDestroy(c->ev_bounce[b]);
    }
    if (c->hpacked) {
        if (c->hmode == DPCA_HOST_PINNED)
            (void)hipHostFree(c->hpacked);
        else
            free(c->hpacked);
    }
    if (c->fd >= 0) close(c->fd);
    if (c->cst) (void)hipStreamDestroy(c->cst);
    for (auto &tm : c->pending) {
        if (tm.a) (void)hipEventDestroy(tm.a);
        if (tm.b) (void)hipEventDestroy(tm.b);
    }
    for (auto e : c->free_ev) (void)hipEventDestroy(e);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

// Host RAM in KB: MemTotal of /proc/meminfo as GetRamInKB reads it
// (src/utilities.cpp:60-80), as a 64-bit value (the reference's int
// overflows above 2 TB); the page count where /proc is unreadable.
static int64_t ram_kb()
{
    if (FILE *f = fopen("/proc/meminfo", "r")) {
        char line[256];
        long long kb = -1;
        while (fgets(line, sizeof line, f))
            if (sscanf(line, "MemTotal: %lld kB", &kb) == 1) break;
        fclose(f);
        if (kb > 0) return kb;
    }
    return (int64_t)sysconf(_SC_PHYS_PAGES) * (int64_t)sysconf(_SC_PAGE_SIZE) / 1024;
}

// bsize and NSV/NRHS clamps of src/distpca.cpp:218-352 for a single rank.
static int resolve_params(dpca_ctx *c, const dpca_params *in)
{
    c->p = *in;
    dpca_params &p = c->p;
    const int M = c->M;
    const int64_t N = c->N;
    if (M <= 0 || N <= 0) {
        set_error("M and/or N were either zero or negative. Aborting...");
        return -1;
    }
    if (N > 0x7fffffffLL) {
        set_error("N = %lld exceeds the reference's int range", (long long)N);
        return -1;
    }
    if (p.bsize <= 0) {
        const double ram_gb = (double)ram_kb() / 1e6;  // KB / 1e6 (src/distpca.cpp:215)
        const double mem = (p.memory == 0.0) ? ram_gb / 10.0 : p.memory;
        const double membuff = 3.0 * N * 8.0 + (double)M * p.nsv * 8.0 + 2.0 * N * p.nrhs +
                               3.0 * M * 8.0 + 2048.0 * 100000.0;
        const double avail = mem * 1e9 - membuff;
        const double blk = avail > 0.0 ? avail / (8.0 * M) : (membuff + M * 8.0) / 1e7;
        p.bsize = (int)blk;
        if (p.bsize <= 0) p.bsize = p.nsv;
        if (p.bsize >= N) p.bsize = (int)N;
    } else {
        if (p.bsize > N) p.bsize = (int)N;
        if (p.bsize < 1) p.bsize = p.nsv;
    }
    if (p.nsv <= 0) {
        set_error("NSV was either zero or negative. Aborting...");
        return -1;
    }
    const int mn = (int)std::min<int64_t>(M, N);
    if (p.nsv > mn) p.nsv = mn;
    if (p.nrhs <= 0) p.nrhs = std::min(2 * p.nsv, mn);
    if (p.nrhs > mn) p.nrhs = mn;
    if (p.nrhs < p.nsv) p.nrhs = std::min(2 * p.nsv, mn);
    if (p.nrhs > 64) {
        set_error("nrhs = %d: this build supports at most 64 right-hand sides", p.nrhs);
        return -1;
    }
    if (p.crit < 0 || p.crit > 2) {
        set_error("crit must be 0, 1 or 2");
        return -1;
    }
    if (p.power < 1) p.power = 1;  // the reference always runs at least one pass pair
    c->k = p.nrhs;
    c->kp = (int)round_up(c->k, 16);
    c->nsv = p.nsv;
    c->bsize = p.bsize;
    c->incore = (p.bsize == N);
    c->np = (M + 3) / 4;
    c->ldp = round_up(c->np, kLdpAlign);
    c->Mpad = 4 * c->ldp;
    c->Npad = round_up(N, kRowAlign);
    return 0;
}

// Workgroup slots of the device (resident workgroups per CU x CUs), 0 if
// unknown (no device).
static int64_t device_slots(int blocks_per_cu)
{
    int dev = 0, ncu = 0;
    if (blocks_per_cu <= 0 || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return (int64_t)blocks_per_cu * ncu;
}

// SNP split count for a product kernel with `tiles` workgroups per split over
// B rows: n in [lo, hi] minimising rounds x (rows per split + 256 rows of
// per-workgroup start-up and write-out), rounds = ceil(tiles n / slots) of
// resident workgroups; `dflt` unless that saves > 3% (each split adds a
// partial to reduce) or when the slots are unknown.
static int64_t fit_splits(int64_t tiles, int64_t B, int64_t align, int64_t dflt, int64_t lo,
                          int64_t hi, int64_t slots)
{
    if (slots <= 0) return dflt;
    auto cost = [&](int64_t n) {
        return (tiles * n + slots - 1) / slots * (round_up((B + n - 1) / n, align) + 256);
    };
    int64_t best = dflt;
    for (int64_t n = std::max<int64_t>(1, lo); n <= hi; n++)
        if (cost(n) < cost(best)) best = n;
    return cost(best) < 0.97 * cost(dflt) ? best : dflt;
}

static int alloc_work(dpca_ctx *c)
{
    const size_t mk = (size_t)c->Mpad * c->kp * sizeof(double);
    const int64_t B = c->block_rows;  // rows of the largest block a pass works on
    int r = 0;
    r |= dalloc(&c->zval, (size_t)c->Npad * 4 * sizeof(double));
    r |= dalloc(&c->mu, (size_t)c->Npad * sizeof(double));
    r |= dalloc(&c->dsc, (size_t)c->Npad * sizeof(double));
    r |= dalloc(&c->ssum, (size_t)c->Npad * sizeof(int64_t));
    r |= dalloc(&c->sctr, (size_t)c->Npad * sizeof(int64_t));
    r |= dalloc(&c->Q, mk);
    r |= dalloc(&c->ARHS, mk);
    r |= dalloc(&c->RHS2, mk);
    r |= dalloc(&c->RHS2old, mk);
    r |= dalloc(&c->tmp, mk);
    r |= dalloc(&c->T, (size_t)c->Npad * c->kp * sizeof(double));
    // FP64 Z^T T: aim for >= ~2048 workgroups, each split >= 256 rows, then
    // fit the count (half to twice that) to the resident workgroup slots
    const int64_t tiles = c->ldp / 64;
    int64_t sp = (2048 + tiles - 1) / tiles;
    sp = std::max<int64_t>(1, std::min<int64_t>(sp, B / 256));
    if (!g_job_ctx)
        sp = fit_splits(tiles, B, 64, sp, sp / 2, std::min<int64_t>(2 * sp, std::max<int64_t>(1, B / 256)),
                        device_slots(zt_f64_blocks_per_cu(c->k, c->kp)));
    c->splits = (int)sp;
    {
        // FP64 Z Q (the FP64 form's passes, and the loadings of a model in
        // either form): 128 or 256 SNP rows per workgroup; short blocks
        // (streamed) also split the samples so that there are >= ~2048
        // workgroups
        const int64_t rg = std::max<int64_t>(1, B / za_f64_rows_per_wg(c->k, c->kp));
        const int64_t ns = std::max<int64_t>(1, std::min<int64_t>((2048 + rg - 1) / rg, c->ldp / 64));
        c->seg64 = round_up((c->ldp + ns - 1) / ns, 64);
        c->nseg64 = (int)((c->ldp + c->seg64 - 1) / c->seg64);
        if (c->nseg64 > 1) r |= dalloc(&c->Tseg, (size_t)c->nseg64 * B * c->kp * sizeof(double));
    }
    if (c->p.precision == DPCA_FP64) {
        r |= dalloc(&c->part, (size_t)c->splits * mk);
    } else {
        // T = Z Q: sample segments of <= 65536 samples (the int32 bound of
        // the centred rows; fewer segments, fewer FP32 partials to reduce),
        // 2 or 1 SNP tiles per wave by column tiles.
        // (short blocks: enough segments for >= ~1024 workgroups)
        const int64_t nsegs = std::min<int64_t>(
            c->ldp / 128, std::max<int64_t>((c->ldp + 16383) / 16384,
                                            (1024 + B / 256 - 1) / std::max<int64_t>(1, B / 256)));
        c->seg_bytes = round_up((c->ldp + nsegs - 1) / nsegs, 128);
        c->nseg = (int)((c->ldp + c->seg_bytes - 1) / c->seg_bytes);
        // serving jobs keep the two-tile forms (fewer resident waves, room
        // for the projections beside them)
        const int nsl = c->ns;
        // tuning knobs: DPCA_I8_WAVES (8, or 4: one wave per SIMD with larger
        // tiles), DPCA_ZA_TILES (1, 2, 4) and DPCA_ZT_TILES (2, 4, 8) where
        // the accumulators fit (anything else keeps the computed choice)
        if (const char *e = getenv("DPCA_I8_WAVES"))
            if (!strcmp(e, "4")) c->za_waves = c->zt_waves = 4;
        const bool two = g_job_ctx && za_i8_valid(c->k, nsl, 2, c->za_waves);
        c->st_tiles = two ? 2 : za_i8_tiles(c->k, nsl, c->za_waves);
        if (const char *e = getenv("DPCA_ZA_TILES"))
            if (za_i8_valid(c->k, nsl, atoi(e), c->za_waves)) c->st_tiles = atoi(e);
        // Z^T T: SNP splits of <= 32768 rows (fewer partials), >= ~1024
        // workgroups
        c->sa_tiles = (two && zt_i8_valid(c->k, nsl, 4, c->ldp, c->zt_waves))
                          ? 4
                          : zt_i8_tiles(c->k, nsl, c->ldp, c->zt_waves);
        if (const char *e = getenv("DPCA_ZT_TILES"))
            if (zt_i8_valid(c->k, nsl, atoi(e), c->ldp, c->zt_waves)) c->sa_tiles = atoi(e);
        const int64_t tl = c->ldp / zt_i8_tile_bytes(c->sa_tiles, c->zt_waves);
        int64_t ns = std::max<int64_t>((1024 + tl - 1) / tl, (B + 32767) / 32768);
        ns = std::min<int64_t>(ns, B / 128);
        // up to 4x more splits where they fill the last round of resident
        // workgroups better
        if (!g_job_ctx)
            ns = fit_splits(tl, B, 128, ns, ns, std::min<int64_t>(4 * ns, B / 128),
                            device_slots(zt_i8_blocks_per_cu(c->k, nsl, c->sa_tiles, c->zt_waves)));
        c->rps = round_up((B + ns - 1) / ns, 128);
        c->nsplit = (int)((B + c->rps - 1) / c->rps);
        const int vt = col_tiles(c->k, nsl).vt();
        r |= dalloc(&c->Q8, (size_t)c->Mpad / 16 * vt * 16 * sizeof(uint4));
        r |= dalloc(&c->qmax, (size_t)c->nseg * c->kp * sizeof(unsigned));
        r |= dalloc(&c->UW8, (size_t)c->Npad / 8 * vt * 16 * sizeof(uint4));
        r |= dalloc(&c->smax, (size_t)c->nsplit * c->kp * sizeof(unsigned));
        r |= dalloc(&c->Tpart, (size_t)c->nseg * B * c->kp * sizeof(float));
        r |= dalloc(&c->Apart, (size_t)c->nsplit * c->Mpad * c->kp * sizeof(float));
    }
    r |= dalloc(&c->gpart, (size_t)c->ngparts * 64 * 64 * sizeof(double));
    r |= dalloc(&c->G, 64 * 64 * sizeof(double));
    r |= dalloc(&c->Rinv, 64 * 64 * sizeof(double));
    r |= dalloc(&c->w, 64 * sizeof(double));
    r |= dalloc(&c->X, 64 * 64 * sizeof(double));
    r |= dalloc(&c->dots, 64 * sizeof(double));
    r |= dalloc(&c->flags, 16 * sizeof(int));
    r |= dalloc(&c->cinfo, (kMaxKp + 1) * sizeof(double));
    r |= dalloc(&c->its, sizeof(IterState));
    if (r) return -3;
    if (g_dry) return 0;
    DPCA_CHECK(hipMemsetAsync(c->Q, 0, mk, c->st));
    DPCA_CHECK(hipMemsetAsync(c->ARHS, 0, mk, c->st));
    DPCA_CHECK(hipMemsetAsync(c->RHS2, 0, mk, c->st));
    DPCA_CHECK(hipMemsetAsync(c->RHS2old, 0, mk, c->st));
    DPCA_CHECK(hipMemsetAsync(c->tmp, 0, mk, c->st));
    return 0;
}

// Per-SNP statistics of rows [r0, r0+rows) held at `blk` (first touch of
// those SNPs, as src/methods.cpp:819-836).
static int block_stats(dpca_ctx *c, const uint8_t *blk, int64_t r0, int64_t rows)
{
    Stage s(c, T_LOAD);
    DPCA_CHECK(launch_snp_stats(c->st, blk, rows, c->ldp, 1.0 / sqrt((double)c->N),
                                c->ssum + r0, c->sctr + r0, c->zval + 4 * r0, c->mu + r0,
                                c->dsc + r0));
    return 0;
}

// Z^T T reads a tile-order image of the matrix (a second copy of the same
// size, launch_zt_image) in mixed, HBM-resident contexts that are not
// serving jobs, with 8-wave Z^T T workgroups, unless p.zt_image is 0; a
// device_cap has to hold both copies.
static bool want_zt_image(const dpca_ctx *c, int64_t bytes)
{
    return c->p.zt_image && c->p.precision == DPCA_MIXED && !c->streamed && !g_job_ctx &&
           c->zt_waves == 8 && (c->p.device_cap <= 0 || 2 * bytes <= c->p.device_cap);
}

// Once the matrix is resident: per-SNP statistics and, when allocated, the
// Z^T T image.
static int compute_stats(dpca_ctx *c)
{
    if (c->streamed) return 0;  // computed on the first streamed pass
    int r = block_stats(c, c->packed, 0, c->Npad);
    if (!r) c->stats_done = true;
    if (!r && c->zt_img) {
        Stage s(c, T_LOAD);
        DPCA_CHECK(launch_zt_image(c->st, c->packed, c->Npad, c->ldp, c->sa_tiles, c->zt_waves,
                                   c->zt_img));
    }
    return r;
}

static int new_ctx(const dpca_params *p, int M, int64_t N, dpca_ctx **out, bool from_file = false)
{
    *out = nullptr;
    if (!p) {
        set_error("params is NULL");
        return -1;
    }
    dpca_ctx *c = new dpca_ctx();
    c->dev = p->device;
    c->M = M;
    c->N = N;
    // parameter errors are reported before any device call
    int r = resolve_params(c, p);
    if (r) {
        delete c;
        return r;
    }
    if (c->p.slices == 0) c->p.slices = 3;
    c->ns = c->p.slices;
    if (c->p.slices != 2 && c->p.slices != 3) {
        set_error("slices must be 2 or 3 (0 for the default)");
        delete c;
        return -1;
    }
    if (c->p.precision != DPCA_FP64 && c->p.precision != DPCA_MIXED) {
        set_error("unknown precision %d", c->p.precision);
        delete c;
        return -1;
    }
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) {
        set_error("hipSetDevice(%d): %s", c->dev, hipGetErrorString(e));
        delete c;
        return -2;
    }
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        delete c;
        return -2;
    }
    const int64_t bytes = c->Npad * c->ldp;
    c->streamed = (c->p.device_cap > 0 && bytes > c->p.device_cap);
    c->block_rows = c->Npad;
    if (c->streamed) {
        // blocks of ~512 MB (or a quarter of the cap), whole multiples of 1024
        // rows; two of them in HBM at a time
        int64_t target = std::min<int64_t>(512ll << 20, c->p.device_cap / 4);
        c->chunk_rows = std::max<int64_t>(1024, target / c->ldp / 1024 * 1024);
        c->chunk_rows = std::min(c->chunk_rows, c->Npad);
        c->block_rows = c->chunk_rows;
        if (hipStreamCreateWithFlags(&c->cst, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreate failed");
            free_ctx(c);
            return -2;
        }
        c->hmode = c->p.host_mode;
        if (c->hmode < DPCA_HOST_PINNED || c->hmode > DPCA_HOST_FILE ||
            (c->hmode == DPCA_HOST_FILE && !from_file)) {
            set_error("host_mode %d is not available here (file mode needs dpca_open)", c->hmode);
            free_ctx(c);
            return -1;
        }
        for (int b = 0; b < 2; b++) {
            if ((r = dalloc(&c->dbuf[b], (size_t)(c->chunk_rows * c->ldp)))) {
                free_ctx(c);
                return r;
            }
            if (hipEventCreateWithFlags(&c->ev_copied[b], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_free[b], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_bounce[b], hipEventDisableTiming) != hipSuccess) {
                set_error("hipEventCreate failed");
                free_ctx(c);
                return -2;
            }
            (void)hipEventRecord(c->ev_free[b], c->st);
            (void)hipEventRecord(c->ev_bounce[b], c->st);
        }
        // Pinned host memory is limited by the driver (typically half the
        // RAM); above that the matrix stays pageable and each block is staged
        // through a pinned buffer.
        if (c->hmode == DPCA_HOST_PINNED &&
            hipHostMalloc((void **)&c->hpacked, (size_t)bytes) != hipSuccess) {
            c->hpacked = nullptr;
            (void)hipGetLastError();
            c->hmode = DPCA_HOST_PAGEABLE;
        }
        if (c->hmode == DPCA_HOST_PAGEABLE && !(c->hpacked = (uint8_t *)malloc((size_t)bytes))) {
            set_error("malloc(%lld bytes) for the streamed matrix failed", (long long)bytes);
            free_ctx(c);
            return -3;
        }
        if (c->hmode != DPCA_HOST_PINNED) {
            const int64_t row_bytes = (c->hmode == DPCA_HOST_FILE) ? c->np : c->ldp;
            for (int b = 0; b < 2; b++) {
                if (hipHostMalloc((void **)&c->bounce[b], (size_t)(c->chunk_rows * row_bytes)) !=
                    hipSuccess) {
                    set_error("hipHostMalloc of a %lld-byte staging buffer failed",
                              (long long)(c->chunk_rows * row_bytes));
                    free_ctx(c);
                    return -3;
                }
                if (c->hmode == DPCA_HOST_FILE &&
                    (r = dalloc(&c->draw[b], (size_t)(c->chunk_rows * c->np)))) {
                    free_ctx(c);
                    return r;
                }
            }
        }
    }
    if ((r = alloc_work(c))) {
        free_ctx(c);
        return r;
    }
    if (getenv("DPCA_SHOW_PLAN"))  // launch shapes, for tuning
        fprintf(stderr,
                "dpca plan: block rows %lld; fp64 splits %d, Z Q segments %d, Z Q tiles %d; "
                "int8 slices %d, Z Q tiles %d x %d waves segments %d, Z^T T tiles %d x %d waves "
                "splits %d (%lld rows)\n",
                (long long)c->block_rows, c->splits, c->nseg64,
                za_f64_rows_per_wg(c->k, c->kp) / 64, c->ns, c->st_tiles, c->za_waves, c->nseg,
                c->sa_tiles, c->zt_waves, c->nsplit, (long long)c->rps);
    if (!c->streamed && (r = dalloc(&c->packed, (size_t)bytes))) {
        free_ctx(c);
        return r;
    }
    if (want_zt_image(c, bytes)) {
        // only where the copy leaves 1 GB of the device free; without the
        // memory the context simply runs the staging kernel
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > (size_t)bytes + (1ull << 30) &&
            hipMalloc((void **)&c->zt_img, (size_t)bytes) != hipSuccess) {
            c->zt_img = nullptr;
            (void)hipGetLastError();
        }
    }
    if (c->streamed) {
        // The rest of the cap keeps the leading blocks in HBM once the first
        // pass has copied them, so later passes stream only the others.
        // Allocated last, so it never takes memory the work buffers need;
        // without the memory the context just streams every block.
        const int64_t cb = c->chunk_rows * c->ldp;
        const int64_t nblk = (c->Npad + c->chunk_rows - 1) / c->chunk_rows;
        int64_t rb = std::max<int64_t>(
            0, std::min<int64_t>((c->p.device_cap - 2 * cb) / cb, nblk - 1));
        if (rb > 0 && hipMalloc((void **)&c->resident, (size_t)(rb * cb)) != hipSuccess) {
            (void)hipGetLastError();
            c->resident = nullptr;
            rb = 0;
        }
        c->res_blocks = rb;
    }
    *out = c;
    return 0;
}

// .bed rows (np bytes) -> device layout rows (ldp bytes, padding codes 01)
// on the host, for the streamed matrix (same bytes as k_repack).
static void host_repack_rows(const uint8_t *raw, int64_t n, int64_t np, int64_t ldp, int M,
                             uint8_t *dst)
{
    const int valid = M - 4 * (int)(np - 1);
    const uint8_t keep = (uint8_t)((1u << (2 * valid)) - 1u);
    const int T = host_threads();
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([=] {
            for (int64_t r = t; r < n; r += T) {
                uint8_t *d = dst + r * ldp;
                memcpy(d, raw + r * np, (size_t)np);
                if (valid < 4) d[np - 1] = (uint8_t)((d[np - 1] & keep) | (kPadByte & ~keep));
                memset(d + np, kPadByte, (size_t)(ldp - np));
            }
        });
    for (auto &x : th) x.join();
}

// Host copy split over threads (pageable -> pinned staging and back).
static void par_copy(uint8_t *dst, const uint8_t *src, size_t n)
{
    const int T = host_threads();
    if (n < (16u << 20) || T == 1) {
        memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + T - 1) / T + 4095) & ~(size_t)4095;
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++) {
        const size_t a = (size_t)t * per;
        if (a >= n) break;
        const size_t len = std::min(per, n - a);
        th.emplace_back([=] { memcpy(dst + a, src + a, len); });
    }
    for (auto &x : th) x.join();
}

// Reads n bytes at offset off of fd into dst, split over threads.
static int par_pread(int fd, uint8_t *dst, size_t n, off_t off)
{
    const int T = host_threads();
    const size_t per = std::max<size_t>(((n + T - 1) / T + 4095) & ~(size_t)4095, 1 << 20);
    std::vector<std::thread> th;
    std::vector<int> bad(T, 0);
    for (int t = 0; t < T; t++) {
        const size_t a = (size_t)t * per;
        if (a >= n) break;
        const size_t len = std::min(per, n - a);
        th.emplace_back([=, &bad] {
            size_t done = 0;
            while (done < len) {
                const ssize_t got = pread(fd, dst + a + done, len - done, off + (off_t)(a + done));
                if (got <= 0) {
                    if (got < 0 && errno == EINTR) continue;
                    bad[t] = 1;
                    return;
                }
                done += (size_t)got;
            }
        });
    }
    for (auto &x : th) x.join();
    for (int b : bad)
        if (b) return -1;
    return 0;
}

// Enqueues on stream s the copy of streamed rows [r0, r0 + rows) (a block of
// at most chunk_rows) into device buffer dst in the device layout.  Pinned
// host copy: one H2D copy.  Pageable copy: the rows are first copied into
// pinned staging buffer b; .bed file: the raw rows are read into staging
// buffer b, copied to the device and repacked there (k_repack, padding rows
// filled).  The host waits until the previous copy out of staging buffer b
// has finished before it refills it.
static int fetch_block(dpca_ctx *c, int b, int64_t r0, int64_t rows, uint8_t *dst, hipStream_t s)
{
    if (c->hmode == DPCA_HOST_PINNED) {
        DPCA_CHECK(hipMemcpyAsync(dst, c->hpacked + r0 * c->ldp, (size_t)(rows * c->ldp),
                                  hipMemcpyHostToDevice, s));
        c->bytes_h2d += (double)(rows * c->ldp);
        return 0;
    }
    DPCA_CHECK(hipEventSynchronize(c->ev_bounce[b]));
    if (c->hmode == DPCA_HOST_PAGEABLE) {
        par_copy(c->bounce[b], c->hpacked + r0 * c->ldp, (size_t)(rows * c->ldp));
        DPCA_CHECK(hipMemcpyAsync(dst, c->bounce[b], (size_t)(rows * c->ldp),
                                  hipMemcpyHostToDevice, s));
        c->bytes_h2d += (double)(rows * c->ldp);
    } else {
        const int64_t real = std::max<int64_t>(0, std::min(rows, c->N - r0));
        if (real > 0) {
            if (par_pread(c->fd, c->bounce[b], (size_t)(real * c->np), (off_t)(3 + r0 * c->np))) {
                set_error("short read of SNPs [%lld, %lld) from the .bed file", (long long)r0,
                          (long long)(r0 + real));
                return -1;
            }
            DPCA_CHECK(hipMemcpyAsync(c->draw[b], c->bounce[b], (size_t)(real * c->np),
                                      hipMemcpyHostToDevice, s));
            DPCA_CHECK(launch_repack(s, c->draw[b], real, c->np, dst, c->ldp, c->M));
            c->bytes_h2d += (double)(real * c->np);
        }
        if (rows > real)
            DPCA_CHECK(launch_fill_u8(s, dst + real * c->ldp, (rows - real) * c->ldp, kPadByte));
    }
    DPCA_CHECK(hipEventRecord(c->ev_bounce[b], s));
    return 0;
}

// Streamed matrix: per-SNP statistics without a pass (one copy sweep).
static int stream_stats_sweep(dpca_ctx *c)
{
    for (int64_t r0 = 0; r0 < c->Npad; r0 += c->chunk_rows) {
        const int64_t rows = std::min(c->chunk_rows, c->Npad - r0);
        int r = fetch_block(c, 0, r0, rows, c->dbuf[0], c->st);
        if (!r) r = block_stats(c, c->dbuf[0], r0, rows);
        if (r) return r;
        DPCA_CHECK(hipStreamSynchronize(c->st));
    }
    c->stats_done = true;
    return 0;
}

// Device-layout block rows [r0, r0 + rows) (from the device) into the host
// copy of a streamed matrix, through staging buffer 0 when it is pageable.
static int park_block(dpca_ctx *c, const uint8_t *src, int64_t r0, int64_t rows)
{
    const size_t n = (size_t)(rows * c->ldp);
    if (c->hmode == DPCA_HOST_PINNED) {
        DPCA_CHECK(hipMemcpyAsync(c->hpacked + r0 * c->ldp, src, n, hipMemcpyDeviceToHost, c->st));
        DPCA_CHECK(hipStreamSynchronize(c->st));
        return 0;
    }
    DPCA_CHECK(hipMemcpyAsync(c->bounce[0], src, n, hipMemcpyDeviceToHost, c->st));
    DPCA_CHECK(hipStreamSynchronize(c->st));
    par_copy(c->hpacked + r0 * c->ldp, c->bounce[0], n);
    return 0;
}

// Raw .bed rows from host memory -> padded device layout, in chunks.
static int upload_rows(dpca_ctx *c, const uint8_t *host, int64_t rows)
{
    if (c->streamed) {
        Stage s(c, T_LOAD);
        host_repack_rows(host, rows, c->np, c->ldp, c->M, c->hpacked);
        memset(c->hpacked + rows * c->ldp, kPadByte, (size_t)((c->Npad - rows) * c->ldp));
        return 0;
    }
    Stage s(c, T_LOAD);
    const int64_t chunk_rows = std::max<int64_t>(1, (256ll << 20) / c->np);
    uint8_t *raw = nullptr;
    const int64_t nrows = std::min(chunk_rows, rows);
    if (dalloc(&raw, (size_t)nrows * c->np)) return -3;
    int rc = 0;
    for (int64_t r0 = 0; r0 < rows && !rc; r0 += chunk_rows) {
        const int64_t n = std::min(chunk_rows, rows - r0);
        hipError_t e = hipMemcpyAsync(raw, host + r0 * c->np, (size_t)(n * c->np),
                                      hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess)
            e = launch_repack(c->st, raw, n, c->np, c->packed + r0 * c->ldp, c->ldp, c->M);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) {
            set_error("upload: %s", hipGetErrorString(e));
            rc = -2;
        }
    }
    (void)hipFree(raw);
    if (rc) return rc;
    DPCA_CHECK(launch_fill_u8(c->st, c->packed + rows * c->ldp, (c->Npad - rows) * c->ldp,
                              kPadByte));
    return 0;
}

// --------------------------------------------------------------- one pass --
// A pass ARHS = Z^T (Z Qin) (src/methods.cpp:786-936) is a loop over blocks
// of SNP rows: the whole resident matrix as one block, or ~512 MB blocks
// streamed from pinned host memory.  Each block adds its Z_b^T (Z_b Qin)
// into `out` (FP64, block order); `half` stops after T = Z_b Qin.

// Mixed form: Q -> int8 slices once per pass.
static int pass_prologue(dpca_ctx *c, const double *Qin)
{
    if (c->p.precision != DPCA_MIXED) return 0;
    Stage s(c, T_MM);
    DPCA_CHECK(launch_q8(c->st, Qin, c->Mpad, c->k, c->ns, c->kp, c->seg_bytes, c->qmax, c->Q8));
    return 0;
}

static int pass_block(dpca_ctx *c, const uint8_t *blk, int64_t r0, int64_t rows,
                      const double *Qin, double *out, bool accumulate, bool half)
{
    double *T = c->T + r0 * c->kp;
    if (c->p.precision == DPCA_FP64) {
        {
            Stage s(c, T_MM_A);
            DPCA_CHECK(launch_za_f64(c->st, blk, rows, c->ldp, c->zval + 4 * r0, Qin, c->k, c->kp, T,
                                     c->seg64, c->nseg64, c->Tseg));
        }
        if (half) return 0;
        {
            Stage s(c, T_MM_AT);
            DPCA_CHECK(launch_zt_f64(c->st, blk, rows, c->ldp, c->zval + 4 * r0, T, c->k, c->kp,
                                     c->splits, c->part));
        }
        Stage s(c, T_MM);
        DPCA_CHECK(launch_reduce_parts(c->st, c->part, c->splits, c->Mpad * c->kp, out,
                                       accumulate ? 1 : 0));
        return 0;
    }
    {
        Stage s(c, T_MM_A);
        DPCA_CHECK(launch_za_i8(c->st, blk, rows, c->ldp, c->Q8, c->qmax, c->mu + r0,
                                c->dsc + r0, c->k, c->ns, c->st_tiles, c->za_waves, c->seg_bytes,
                                c->nseg,
                                c->Tpart));
    }
    {
        Stage s(c, T_MM);
        // the U / W maxima per SNP split come with T (not for a half pass)
        DPCA_CHECK(launch_tprep(c->st, c->Tpart, c->nseg, rows, c->kp, T, c->mu + r0,
                                c->dsc + r0, c->rps, half ? nullptr : c->smax));
    }
    if (half) return 0;
    const int nsplit = (int)((rows + c->rps - 1) / c->rps);
    uint4 *UW8 = c->UW8 + (r0 / 8) * col_tiles(c->k, c->ns).vt() * 16;
    {
        Stage s(c, T_MM);
        DPCA_CHECK(launch_uw8(c->st, T, rows, c->k, c->ns, c->kp, c->mu + r0, c->dsc + r0, c->rps,
                              c->smax, UW8, true));
    }
    {
        Stage s(c, T_MM_AT);
        // the image covers the resident matrix (rows from r0 at the same offset)
        const uint8_t *img =
            (c->zt_img && blk == c->packed + r0 * c->ldp) ? c->zt_img + r0 * c->ldp : nullptr;
        DPCA_CHECK(launch_zt_i8(c->st, blk, img, rows, c->ldp, UW8, c->smax, c->k, c->ns,
                                c->sa_tiles, c->zt_waves, c->rps, nsplit, c->Apart));
    }
    Stage s(c, T_MM);
    DPCA_CHECK(launch_reduce_f32(c->st, c->Apart, nsplit, c->Mpad * c->kp, out,
                                 accumulate ? 1 : 0));
    return 0;
}

// Streamed pass: block j is copied on the copy stream into buffer j&1 while
// block j-1 is multiplied on the compute stream; events order the reuse of
// each buffer.  With a pageable host copy or the .bed file as the source, the
// host fills staging buffer j&1 while the copy engine and the compute stream
// work on the two blocks before it.  The first streamed pass also computes the per-SNP statistics
// of each block before its products (first touch, src/methods.cpp:819-836).
static int pass_streamed(dpca_ctx *c, const double *Qin, double *out, bool half)
{
    const double t0 = now_s();
    const int64_t nblk = (c->Npad + c->chunk_rows - 1) / c->chunk_rows;
    for (int64_t j = 0; j < nblk; j++) {
        const int b = (int)(j & 1);
        const int64_t r0 = j * c->chunk_rows;
        const int64_t rows = std::min(c->chunk_rows, c->Npad - r0);
        int r;
        if (j < c->res_blocks) {
            // a resident block: copied once, by the first pass
            uint8_t *dst = c->resident + r0 * c->ldp;
            if (!c->res_filled) {
                {
                    Stage s(c, T_H2D, c->cst);
                    if ((r = fetch_block(c, b, r0, rows, dst, c->cst))) return r;
                }
                DPCA_CHECK(hipEventRecord(c->ev_copied[b], c->cst));
                DPCA_CHECK(hipStreamWaitEvent(c->st, c->ev_copied[b], 0));
            }
            if (!c->stats_done && (r = block_stats(c, dst, r0, rows))) return r;
            if ((r = pass_block(c, dst, r0, rows, Qin, out, j > 0, half))) return r;
            continue;
        }
        DPCA_CHECK(hipStreamWaitEvent(c->cst, c->ev_free[b], 0));
        {
            Stage s(c, T_H2D, c->cst);
            if ((r = fetch_block(c, b, r0, rows, c->dbuf[b], c->cst))) return r;
        }
        DPCA_CHECK(hipEventRecord(c->ev_copied[b], c->cst));
        DPCA_CHECK(hipStreamWaitEvent(c->st, c->ev_copied[b], 0));
        if (!c->stats_done && (r = block_stats(c, c->dbuf[b], r0, rows))) return r;
        if ((r = pass_block(c, c->dbuf[b], r0, rows, Qin, out, j > 0, half))) return r;
        DPCA_CHECK(hipEventRecord(c->ev_free[b], c->st));
    }
    c->stats_done = true;
    c->res_filled = true;
    c->t_stream_pass += now_s() - t0;  // host-side issue time only
    return 0;
}

static int do_pass(dpca_ctx *c, const double *Qin, double *out)
{
    int r = pass_prologue(c, Qin);
    if (r) return r;
    if (c->streamed)
        r = pass_streamed(c, Qin, out, false);
    else
        r = pass_block(c, c->packed, 0, c->Npad, Qin, out, false, false);
    if (!r) c->passes++;
    return r;
}

// Z Q only (the in-core variant's half pass), into c->T.
static int half_pass(dpca_ctx *c, const double *Qin)
{
    int r = pass_prologue(c, Qin);
    if (r) return r;
    if (c->streamed) return pass_streamed(c, Qin, nullptr, true);
    return pass_block(c, c->packed, 0, c->Npad, Qin, nullptr, false, true);
}

// CholQR2 (+ shifted fallback) of A (Mpad x kp) into c->Q; uses c->tmp.
// In robust mode a basis that spans fewer than k dimensions (the shifted
// fallback still meets a non-positive pivot) has its dependent columns
// replaced in place by k_replace_cols and is factored again, so the result
// is an orthonormal completion of its span, as dgeqrf + dorgqr would give;
// this costs one host synchronisation per call, so it is only switched on for
// a context whose plain run met such a basis.
static int orthonormalise(dpca_ctx *c, double *A)
{
    Stage s(c, T_GS);
    const double u = ldexp(1.0, -53);
    const double shift = 11.0 * ((double)c->M * c->k + (double)c->k * (c->k + 1)) * u;
    for (int attempt = 0;; attempt++) {
        const double *src = A;
        double *dst[3] = {c->tmp, c->Q, c->tmp};
        for (int mode = 0; mode < 3; mode++) {
            DPCA_CHECK(launch_gram(c->st, src, src, c->Mpad, c->kp, c->gpart, c->ngparts, c->G));
            DPCA_CHECK(launch_chol_inv(c->st, c->G, c->k, c->kp, shift, mode, c->Rinv, c->flags,
                                       c->robust ? c->cinfo : nullptr));
            DPCA_CHECK(launch_apply(c->st, src, c->Rinv, c->Mpad, c->kp, dst[mode]));
            src = dst[mode];
        }
        if (!c->robust) break;
        int failed = 0;
        DPCA_CHECK(hipMemcpyAsync(&failed, c->flags + 1, sizeof(int), hipMemcpyDeviceToHost,
                                  c->st));
        DPCA_CHECK(hipStreamSynchronize(c->st));
        if (!failed) break;
        if (attempt == 1) {
            set_error("CholeskyQR failed after completing a rank-deficient basis");
            return -4;
        }
        DPCA_CHECK(launch_replace_cols(c->st, A, c->M, c->Mpad, c->k, c->kp, c->cinfo));
        DPCA_CHECK(hipMemsetAsync(c->flags + 1, 0, sizeof(int), c->st));
        c->qr_repairs++;
    }
    std::swap(c->Q, c->tmp);  // the third pass landed in tmp
    return 0;
}

// Right-hand sides cross the host boundary densely (M x k) and are padded /
// compacted on the device; c->tmp is free at these points.
static int copy_in_rhs(dpca_ctx *c, const double *host, double *dev)
{
    DPCA_CHECK(hipMemcpyAsync(c->tmp, host, (size_t)c->M * c->k * sizeof(double),
                              hipMemcpyHostToDevice, c->st));
    DPCA_CHECK(launch_pad_rhs(c->st, c->tmp, c->M, c->k, c->Mpad, c->kp, dev));
    return 0;
}

static int copy_out_rhs(dpca_ctx *c, const double *dev, double *host)
{
    DPCA_CHECK(launch_compact_rhs(c->st, dev, c->M, c->k, c->kp, c->tmp));
    DPCA_CHECK(hipMemcpyAsync(host, c->tmp, (size_t)c->M * c->k * sizeof(double),
                              hipMemcpyDeviceToHost, c->st));
    DPCA_CHECK(hipStreamSynchronize(c->st));
    return 0;
}

// Returned by run_once when a plain run met a rank-deficient basis; run()
// then repeats it in robust mode.
constexpr int kRerunRobust = 1;

// One outer iteration up to the stopping test (its state in c->its), as in
// SubspaceIteration_MPI (src/methods.cpp:52-192) for the in-core variant and
// BlockSubspaceIter_MPI_OOC_double_buffering (src/methods.cpp:938-1063)
// otherwise.  No host synchronisation: every iteration after the first
// (ii > 0) enqueues the same work, so it can be captured as a graph.
static int iter_body(dpca_ctx *c, int ii)
{
    const dpca_params &p = c->p;
    int r;
    if (c->incore) {
        // `power` full passes with QR between them, then QR, the half pass
        // Z Q and the Rayleigh-Ritz matrix (Z Q)^T (Z Q).
        if ((r = do_pass(c, c->RHS2, c->ARHS))) return r;
        for (int kk = 0; kk < p.power - 1; kk++) {
            if ((r = orthonormalise(c, c->ARHS))) return r;
            if ((r = do_pass(c, c->Q, c->ARHS))) return r;
        }
        if ((r = orthonormalise(c, c->ARHS))) return r;
        if ((r = half_pass(c, c->Q))) return r;
        {
            Stage s(c, T_RR);
            DPCA_CHECK(launch_gram(c->st, c->T, c->T, c->Npad, c->kp, c->gpart, c->ngparts,
                                   c->G));
        }
    } else {
        if (ii > 0 && !p.strict_passes) {
            // A A^T (Q X) = (A A^T Q) X: the previous iteration's last
            // pass, rotated by the Ritz vectors, replaces the first pass.
            Stage s(c, T_MM);
            DPCA_CHECK(launch_apply(c->st, c->ARHS, c->X, c->Mpad, c->kp, c->tmp));
            std::swap(c->ARHS, c->tmp);
        } else if ((r = do_pass(c, c->RHS2, c->ARHS))) {
            return r;
        }
        for (int kk = 0; kk < p.power - 1; kk++) {
            if ((r = orthonormalise(c, c->ARHS))) return r;
            if ((r = do_pass(c, c->Q, c->ARHS))) return r;
        }
        if ((r = orthonormalise(c, c->ARHS))) return r;
        if ((r = do_pass(c, c->Q, c->ARHS))) return r;
        Stage s(c, T_RR);
        DPCA_CHECK(launch_gram(c->st, c->Q, c->ARHS, c->Mpad, c->kp, c->gpart, c->ngparts,
                               c->G));
    }
    Stage s(c, T_RR);
    DPCA_CHECK(launch_syev(c->st, c->G, c->k, c->kp, c->w, c->X, c->flags + 2));
    DPCA_CHECK(launch_apply(c->st, c->Q, c->X, c->Mpad, c->kp, c->RHS2));
    if (p.crit == 2 && ii > 0)
        DPCA_CHECK(launch_col_dots(c->st, c->RHS2old, c->RHS2, c->Mpad, c->kp, c->nsv, c->gpart,
                                   c->ngparts, c->dots));
    DPCA_CHECK(launch_crit(c->st, c->w, c->nsv, p.crit, p.tol, ii, c->dots, c->its));
    return 0;
}

// Graphs replay iterations of resident contexts; streamed passes are driven
// from the host and the robust orthonormalisation synchronises, so those run
// eagerly, as does everything with DPCA_NO_GRAPH set.
static bool use_graphs(const dpca_ctx *c)
{
    return !c->streamed && !c->robust && !c->graphs_off && !getenv("DPCA_NO_GRAPH");
}

namespace dpca {
void ctx_eager_iterations(dpca_ctx *c)
{
    if (c) c->graphs_off = true;
}
}  // namespace dpca

// Iteration ii > 0 from the graph for the current (Q, tmp, ARHS) state,
// captured on first use.  The launches, their order and their arguments are
// those iter_body enqueues eagerly, so the results are the same bit for bit.
static int iter_replay(dpca_ctx *c, int ii)
{
    IterGraph *g = nullptr;
    for (auto &x : c->graphs)
        if (x.in[0] == c->Q && x.in[1] == c->tmp && x.in[2] == c->ARHS) g = &x;
    if (!g) {
        IterGraph ng{};
        ng.in[0] = c->Q;
        ng.in[1] = c->tmp;
        ng.in[2] = c->ARHS;
        const int passes0 = c->passes;
        hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
        int r = 0;
        hipGraph_t graph = nullptr;
        if (e == hipSuccess) {
            c->capturing = true;
            r = iter_body(c, ii);
            c->capturing = false;
            e = hipStreamEndCapture(c->st, &graph);
        }
        if (e == hipSuccess && !r) e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        ng.out[0] = c->Q;
        ng.out[1] = c->tmp;
        ng.out[2] = c->ARHS;
        ng.passes = c->passes - passes0;
        // nothing ran yet: back to the state the graph starts from
        c->passes = passes0;
        c->Q = ng.in[0];
        c->tmp = ng.in[1];
        c->ARHS = ng.in[2];
        if (e != hipSuccess || r) {
            // capture not available here: this and later iterations run eagerly
            (void)hipGetLastError();
            c->graphs_off = true;
            return iter_body(c, ii);
        }
        c->graphs.push_back(ng);
        g = &c->graphs.back();
    }
    {
        Stage s(c, T_GRAPH);
        DPCA_CHECK(hipGraphLaunch(g->exec, c->st));
    }
    c->Q = g->out[0];
    c->tmp = g->out[1];
    c->ARHS = g->out[2];
    c->passes += g->passes;
    c->graph_its++;
    return 0;
}

static int run_once(dpca_ctx *c, double *rhs, double *sigma, dpca_report *rep, double t0,
                    bool fresh)
{
    const dpca_params &p = c->p;
    if (fresh) {
        for (int i = 0; i < T_NKIND; i++) c->t[i] = 0;
        c->passes = 0;
        c->bytes_h2d = 0;
        c->qr_repairs = 0;
    }
    c->crit_hist.clear();
    c->graph_its = 0;
    std::vector<double> om;
    double *rhs_in = rhs;
    if (!rhs_in) {
        const double tg = now_s();
        om.resize((size_t)c->M * c->k);
        omega_generate(c->M, c->k, om.data());
        c->t_generate_rhs = now_s() - tg;
        rhs_in = om.data();
    }
    int r;
    if ((r = copy_in_rhs(c, rhs_in, c->RHS2))) return r;
    if (p.crit == 2)
        DPCA_CHECK(hipMemcpyAsync(c->RHS2old, c->RHS2, (size_t)c->Mpad * c->kp * 8,
                                  hipMemcpyDeviceToDevice, c->st));
    DPCA_CHECK(hipMemsetAsync(c->flags, 0, 16 * sizeof(int), c->st));
    DPCA_CHECK(hipMemsetAsync(c->its, 0, sizeof(IterState), c->st));
    IterState hs{};
    int ii = 0;
    for (ii = 0; ii < p.miter; ii++) {
        double t_before[T_NKIND];
        for (int i = 0; i < T_NKIND; i++) t_before[i] = c->t[i];
        if (ii > 0 && use_graphs(c))
            r = iter_replay(c, ii);
        else
            r = iter_body(c, ii);
        if (r) return r;
        DPCA_CHECK(hipMemcpyAsync(&hs, c->its, sizeof(IterState), hipMemcpyDeviceToHost, c->st));
        DPCA_CHECK(hipStreamSynchronize(c->st));
        collect_timers(c);
        if (ii == 0) {
            double sum = 0;
            for (int i = 0; i < T_NKIND; i++) sum += c->t[i] - t_before[i];
            for (int i = 0; i < T_NKIND; i++)
                c->iter_share[i] = sum > 0 ? (c->t[i] - t_before[i]) / sum : (i == T_MM ? 1.0 : 0.0);
        }
        if (ii > 0) {
            c->crit_hist.push_back(hs.crit);
            if (p.verbose > 1) {
                if (p.crit == 0)
                    printf("Iteration %d: rel. error: %02.13f\n", ii, hs.crit);
                else if (p.crit == 1)
                    printf("Iteration %d: %d converged\n", ii, hs.conv);
                else
                    printf("Iteration %d: MEV error: %02.13f\n", ii, hs.crit);
                fflush(stdout);
            }
        }
        if (hs.stop) break;
        if (p.crit == 2)
            DPCA_CHECK(hipMemcpyAsync(c->RHS2old, c->RHS2, (size_t)c->Mpad * c->kp * 8,
                                      hipMemcpyDeviceToDevice, c->st));
    }
    const int total_its = (ii < p.miter) ? ii + 1 : ii;
    int flags[16];
    DPCA_CHECK(hipMemcpyAsync(flags, c->flags, sizeof flags, hipMemcpyDeviceToHost, c->st));
    DPCA_CHECK(hipStreamSynchronize(c->st));
    if (flags[1] && !c->robust) return kRerunRobust;  // rhs is still the caller's input
    if (rhs && (r = copy_out_rhs(c, c->RHS2, rhs))) return r;
    DPCA_CHECK(hipStreamSynchronize(c->st));
    collect_timers(c);
    if (flags[1]) {
        set_error("CholeskyQR failed even with the shifted fallback (rank-deficient basis)");
        return -4;
    }
    if (flags[2]) {
        set_error("Jacobi eigensolver did not converge");
        return -4;
    }
    if (sigma)
        for (int j = 0; j < c->nsv; j++) sigma[j] = hs.sigma[j];
    if (rep) {
        memset(rep, 0, sizeof *rep);
        rep->M = c->M;
        rep->N = c->N;
        rep->nsv = c->nsv;
        rep->nrhs = c->k;
        rep->power = p.power;
        rep->bsize = c->bsize;
        rep->crit = p.crit;
        rep->total_its = total_its;
        rep->passes = c->passes;
        rep->incore = c->incore;
        rep->qr_shifted = flags[0];
        rep->t_generate_rhs = c->t_generate_rhs;
        rep->t_load = c->t[T_LOAD];
        rep->t_mm_a = c->t[T_MM_A];
        rep->t_mm_at = c->t[T_MM_AT];
        rep->t_mm = c->t[T_MM_A] + c->t[T_MM_AT] + c->t[T_MM];
        rep->t_gs = c->t[T_GS];
        rep->t_rr = c->t[T_RR];
        rep->t_other = c->t[T_OTHER];
        rep->wall = now_s() - t0;
        rep->last_crit = hs.crit;
        rep->streamed = c->streamed;
        rep->t_h2d = c->t[T_H2D];
        rep->h2d_bytes = c->bytes_h2d;
        rep->host_mode = c->streamed ? c->hmode : -1;
        rep->qr_repaired = c->qr_repairs;
        rep->resident_bytes = c->streamed ? (double)(c->res_blocks * c->chunk_rows * c->ldp) : 0.0;
        rep->graph_its = c->graph_its;
        rep->zt_image = c->zt_img != nullptr;
    }
    return 0;
}

static int run(dpca_ctx *c, double *rhs, double *sigma, dpca_report *rep)
{
    const double t0 = now_s();
    int r = run_once(c, rhs, sigma, rep, t0, true);
    if (r == kRerunRobust) {
        // the time of the plain attempt stays in the stage timers
        c->robust = true;
        r = run_once(c, rhs, sigma, rep, t0, false);
    }
    return r;
}

// ---------------------------------------------------------------- PLINK IO --
// Lines of a text file the way src/distpca.cpp:137-147 counts them: every
// newline-terminated line; an unterminated last line is dropped.
static int read_lines(const std::string &path, std::vector<std::string> &out)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        set_error("Error reading file '%s': %s", path.c_str(), strerror(errno));
        return -1;
    }
    std::string cur;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) {
        for (size_t i = 0; i < n; i++) {
            if (buf[i] == '\n') {
                out.push_back(cur);
                cur.clear();
            } else {
                cur.push_back(buf[i]);
            }
        }
    }
    fclose(f);
    return 0;
}

static std::string token(const std::string &line, int idx)
{
    int t = 0;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i])) i++;
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j])) j++;
        if (j > i) {
            if (t == idx) return line.substr(i, j - i);
            t++;
        }
        i = j;
    }
    return std::string();
}

// ------------------------------------------------------------------ C ABI --
extern "C" {

const char *dpca_last_error(void) { return g_err; }

int dpca_params_init(dpca_params *p)
{
    if (!p) return -1;
    memset(p, 0, sizeof *p);
    p->nsv = 10;
    p->nrhs = 20;
    p->power = 1;
    p->crit = 1;
    p->tol = 1e-3;
    p->bsize = 0;
    p->miter = 100;
    p->verbose = 1;
    p->fwrite = 0;
    p->fullsvd = 0;
    p->memory = 0;
    snprintf(p->prefix, sizeof p->prefix, "noprefix");
    p->precision = DPCA_FP64;
    p->device = 0;
    p->device_cap = 0;
    p->strict_passes = 1;
    p->host_mode = DPCA_HOST_PINNED;
    p->slices = 0;
    p->zt_image = 1;
    return 0;
}

int dpca_omega(int M, int k, double *out)
{
    if (M <= 0 || k <= 0 || !out) {
        set_error("dpca_omega: bad arguments");
        return -1;
    }
    return omega_generate(M, k, out);
}

int dpca_open(const char *bfile, const dpca_params *p, dpca_ctx **out)
{
    if (!bfile || !out) {
        set_error("dpca_open: NULL argument");
        return -1;
    }
    const std::string base(bfile);
    const std::string bed = base + ".bed";
    FILE *fb = fopen(bed.c_str(), "rb");
    if (!fb) {
        set_error("[Data::read_bed] Error reading file %s", bed.c_str());
        return -1;
    }
    std::vector<std::string> fam, bim;
    if (read_lines(base + ".fam", fam) || read_lines(base + ".bim", bim)) {
        fclose(fb);
        return -1;
    }
    const int M = (int)fam.size();
    const int64_t N = (int64_t)bim.size();
    const int64_t np = (M + 3) / 4;
    struct stat sb;
    if (fstat(fileno(fb), &sb) != 0 || (int64_t)sb.st_size < 3 + N * np) {
        set_error("%s holds %lld bytes; %d samples x %lld SNPs need %lld", bed.c_str(),
                  (long long)sb.st_size, M, (long long)N, (long long)(3 + N * np));
        fclose(fb);
        return -1;
    }
    dpca_ctx *c = nullptr;
    int r = new_ctx(p, M, N, &c, true);
    if (r) {
        fclose(fb);
        return r;
    }
    c->iids.reserve(M);
    for (auto &l : fam) c->iids.push_back(token(l, 1));
    fseek(fb, 3, SEEK_SET);  // 3-byte magic, not validated (src/distpca.cpp:105)
    if (c->streamed && c->hmode == DPCA_HOST_FILE) {
        // every pass reads the file again (src/methods.cpp:795-811)
        c->fd = dup(fileno(fb));
        fclose(fb);
        if (c->fd < 0) {
            set_error("dup of the .bed descriptor failed");
            free_ctx(c);
            return -1;
        }
        *out = c;
        return 0;
    }
    if (c->streamed) {
        // the matrix stays in pinned host memory in the device row layout
        const int64_t chunk = std::max<int64_t>(1, (256ll << 20) / np);
        std::vector<uint8_t> buf((size_t)(std::min(chunk, N) * np));
        for (int64_t r0 = 0; r0 < N && !r; r0 += chunk) {
            const int64_t n = std::min(chunk, N - r0);
            if (fread(buf.data(), 1, (size_t)(n * np), fb) != (size_t)(n * np)) {
                set_error("short read on %s", bed.c_str());
                r = -1;
                break;
            }
            host_repack_rows(buf.data(), n, np, c->ldp, M, c->hpacked + r0 * c->ldp);
        }
        fclose(fb);
        if (!r) memset(c->hpacked + N * c->ldp, kPadByte, (size_t)((c->Npad - N) * c->ldp));
        if (r) {
            free_ctx(c);
            return r;
        }
        *out = c;
        return 0;
    }
    // read in pinned chunks and stage
    const int64_t chunk_rows = std::max<int64_t>(1, (256ll << 20) / np);
    uint8_t *pin = nullptr;
    if (hipHostMalloc((void **)&pin, (size_t)std::min(chunk_rows, N) * np) != hipSuccess) {
        set_error("hipHostMalloc failed");
        fclose(fb);
        free_ctx(c);
        return -3;
    }
    uint8_t *raw = nullptr;
    if (dalloc(&raw, (size_t)std::min(chunk_rows, N) * np)) {
        (void)hipHostFree(pin);
        fclose(fb);
        free_ctx(c);
        return -3;
    }
    {
        Stage s(c, T_LOAD);
        for (int64_t r0 = 0; r0 < N && !r; r0 += chunk_rows) {
            const int64_t n = std::min(chunk_rows, N - r0);
            if (fread(pin, 1, (size_t)(n * np), fb) != (size_t)(n * np)) {
                set_error("short read on %s", bed.c_str());
                r = -1;
                break;
            }
            hipError_t e = hipMemcpyAsync(raw, pin, (size_t)(n * np), hipMemcpyHostToDevice, c->st);
            if (e == hipSuccess)
                e = launch_repack(c->st, raw, n, np, c->packed + r0 * c->ldp, c->ldp, M);
            if (e == hipSuccess) e = hipStreamSynchronize(c->st);
            if (e != hipSuccess) {
                set_error("bed upload: %s", hipGetErrorString(e));
                r = -2;
            }
        }
    }
    fclose(fb);
    (void)hipHostFree(pin);
    (void)hipFree(raw);
    if (!r && launch_fill_u8(c->st, c->packed + N * c->ldp, (c->Npad - N) * c->ldp, kPadByte) !=
                  hipSuccess) {
        set_error("fill failed");
        r = -2;
    }
    if (!r) r = compute_stats(c);
    if (!r && hipStreamSynchronize(c->st) != hipSuccess) r = -2;
    if (r) {
        free_ctx(c);
        return r;
    }
    *out = c;
    return 0;
}

int dpca_open_packed(const uint8_t *packed, int M, long long N, const dpca_params *p,
                     dpca_ctx **out)
{
    if (!packed || !out) {
        set_error("dpca_open_packed: NULL argument");
        return -1;
    }
    dpca_ctx *c = nullptr;
    int r = new_ctx(p, M, N, &c);
    if (r) return r;
    for (int i = 0; i < M; i++) c->iids.push_back(std::to_string(i));
    if ((r = upload_rows(c, packed, N)) || (r = compute_stats(c))) {
        free_ctx(c);
        return r;
    }
    if (hipStreamSynchronize(c->st) != hipSuccess) {
        set_error("stream sync failed");
        free_ctx(c);
        return -2;
    }
    *out = c;
    return 0;
}

static int prep_gen(const dpca_gen_spec *g, GenSpec *spec, GenTables *tab)
{
    if (!g) return -1;
    int r = gen_prepare(g->M, g->N, g->npop, g->fst, g->admix_frac, g->missing, g->seed,
                        g->pmin > 0 ? g->pmin : 0.05, g->pmax > 0 ? g->pmax : 0.5, spec, tab);
    if (r) set_error("invalid generator spec");
    return r;
}

int dpca_open_synthetic(const dpca_gen_spec *g, const dpca_params *p, dpca_ctx **out)
{
    GenSpec spec;
    GenTables tab;
    if (prep_gen(g, &spec, &tab)) return -1;
    dpca_ctx *c = nullptr;
    int r = new_ctx(p, g->M, g->N, &c);
    if (r) {
        gen_free(&tab);
        return r;
    }
    for (int i = 0; i < g->M; i++) c->iids.push_back(std::to_string(i));
    uint32_t *dthr = nullptr;
    uint16_t *dwts = nullptr;
    const size_t nthr = (size_t)g->N * g->npop, nw = (size_t)(g->M - spec.m_pure) * g->npop;
    if (dalloc(&dthr, nthr * 4) || dalloc(&dwts, std::max<size_t>(nw, 1) * 2)) {
        gen_free(&tab);
        free_ctx(c);
        return -3;
    }
    hipError_t e = hipMemcpy(dthr, tab.thr, nthr * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && nw) e = hipMemcpy(dwts, tab.wts, nw * 2, hipMemcpyHostToDevice);
    {
        Stage s(c, T_LOAD);
        if (c->streamed) {
            // generate block by block in HBM and park it in pinned host memory
            for (int64_t r0 = 0; r0 < c->Npad && e == hipSuccess; r0 += c->chunk_rows) {
                const int64_t rows = std::min(c->chunk_rows, c->Npad - r0);
                e = launch_generate(c->st, spec, dthr, dwts, r0, rows, c->dbuf[0], c->ldp);
                if (e == hipSuccess && park_block(c, c->dbuf[0], r0, rows)) e = hipErrorUnknown;
            }
        } else if (e == hipSuccess) {
            e = launch_generate(c->st, spec, dthr, dwts, 0, c->Npad, c->packed, c->ldp);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void)hipFree(dthr);
    (void)hipFree(dwts);
    gen_free(&tab);
    if (e != hipSuccess) {
        set_error("generate: %s", hipGetErrorString(e));
        free_ctx(c);
        return -2;
    }
    if ((r = compute_stats(c)) || hipStreamSynchronize(c->st) != hipSuccess) {
        free_ctx(c);
        return r ? r : -2;
    }
    collect_timers(c);
    *out = c;
    return 0;
}

int dpca_write_synthetic(const char *prefix, const dpca_gen_spec *g)
{
    GenSpec spec;
    GenTables tab;
    if (!prefix || prep_gen(g, &spec, &tab)) return -1;
    const std::string base(prefix);
    FILE *f = fopen((base + ".fam").c_str(), "w");
    if (!f) {
        set_error("cannot write %s.fam", prefix);
        gen_free(&tab);
        return -1;
    }
    for (int i = 0; i < g->M; i++) {
        const int pop = i < spec.m_pure ? gen_pop(spec, i) : -1;
        if (pop >= 0)
            fprintf(f, "POP%d IND%d 0 0 0 -9\n", pop, i);
        else
            fprintf(f, "ADMIX IND%d 0 0 0 -9\n", i);
    }
    fclose(f);
    f = fopen((base + ".bim").c_str(), "w");
    if (!f) {
        set_error("cannot write %s.bim", prefix);
        gen_free(&tab);
        return -1;
    }
    for (int64_t s = 0; s < g->N; s++)
        fprintf(f, "1\tsnp%lld\t0\t%lld\tA\tG\n", (long long)s, (long long)(s + 1));
    fclose(f);
    f = fopen((base + ".bed").c_str(), "wb");
    if (!f) {
        set_error("cannot write %s.bed", prefix);
        gen_free(&tab);
        return -1;
    }
    const uint8_t magic[3] = {0x6c, 0x1b, 0x01};
    fwrite(magic, 1, 3, f);
    const int64_t np = (g->M + 3) / 4;
    const int64_t chunk = std::max<int64_t>(1, (64ll << 20) / np);
    std::vector<uint8_t> buf((size_t)(std::min(chunk, (int64_t)g->N) * np));
    for (int64_t s0 = 0; s0 < g->N; s0 += chunk) {
        const int64_t n = std::min(chunk, (int64_t)g->N - s0);
        gen_rows_host(spec, tab, s0, n, buf.data(), host_threads());
        fwrite(buf.data(), 1, (size_t)(n * np), f);
    }
    fclose(f);
    gen_free(&tab);
    return 0;
}

int dpca_resolve(int M, long long N, const dpca_params *in, dpca_params *out)
{
    if (!in || !out) {
        set_error("params is NULL");
        return -1;
    }
    dpca_ctx c;  // sizes and parameters only: no device state is touched
    c.M = M;
    c.N = N;
    const int r = resolve_params(&c, in);
    if (r) return r;
    *out = c.p;
    return 0;
}

int dpca_get_dims(dpca_ctx *c, int *M, long long *N, int *nsv, int *nrhs, int *bsize)
{
    if (!c) return -1;
    if (M) *M = c->M;
    if (N) *N = c->N;
    if (nsv) *nsv = c->nsv;
    if (nrhs) *nrhs = c->k;
    if (bsize) *bsize = c->bsize;
    return 0;
}

int dpca_snp_stats(dpca_ctx *c, long long *sum, long long *ctr, double *zval)
{
    if (!c) return -1;
    (void)hipSetDevice(c->dev);
    if (!c->stats_done) {
        int r = stream_stats_sweep(c);
        if (r) return r;
    }
    if (sum) DPCA_CHECK(hipMemcpy(sum, c->ssum, (size_t)c->N * 8, hipMemcpyDeviceToHost));
    if (ctr) DPCA_CHECK(hipMemcpy(ctr, c->sctr, (size_t)c->N * 8, hipMemcpyDeviceToHost));
    if (zval) DPCA_CHECK(hipMemcpy(zval, c->zval, (size_t)c->N * 32, hipMemcpyDeviceToHost));
    return 0;
}

int dpca_decode(dpca_ctx *c, long long snp0, long long rows, uint8_t *out)
{
    if (!c || !out || snp0 < 0 || rows < 0 || snp0 + rows > c->N) {
        set_error("dpca_decode: bad arguments");
        return -1;
    }
    (void)hipSetDevice(c->dev);
    uint8_t *d = nullptr, *src = nullptr;
    if (dalloc(&d, (size_t)rows * c->M)) return -3;
    hipError_t e = hipSuccess;
    const uint8_t *rowsrc = c->packed ? c->packed + snp0 * c->ldp : nullptr;
    if (c->streamed) {
        if (dalloc(&src, (size_t)(rows * c->ldp))) {
            (void)hipFree(d);
            return -3;
        }
        if (c->hmode != DPCA_HOST_FILE) {
            e = hipMemcpyAsync(src, c->hpacked + snp0 * c->ldp, (size_t)(rows * c->ldp),
                               hipMemcpyHostToDevice, c->st);
        } else {
            std::vector<uint8_t> raw((size_t)(rows * c->np));
            uint8_t *draw = nullptr;
            if (par_pread(c->fd, raw.data(), raw.size(), (off_t)(3 + snp0 * c->np)) ||
                dalloc(&draw, raw.size())) {
                set_error("dpca_decode: reading the .bed file failed");
                (void)hipFree(d);
                (void)hipFree(src);
                return -1;
            }
            e = hipMemcpy(draw, raw.data(), raw.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = launch_repack(c->st, draw, rows, c->np, src, c->ldp, c->M);
            if (e == hipSuccess) e = hipStreamSynchronize(c->st);
            (void)hipFree(draw);
        }
        rowsrc = src;
    }
    if (e == hipSuccess) e = launch_decode_codes(c->st, rowsrc, rows, c->ldp, c->M, d);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, (size_t)rows * c->M, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void)hipFree(d);
    if (src) (void)hipFree(src);
    if (e != hipSuccess) {
        set_error("dpca_decode: %s", hipGetErrorString(e));
        return -2;
    }
    return 0;
}

int dpca_copy_packed(dpca_ctx *c, long long snp0, long long rows, uint8_t *out)
{
    if (!c || !out || snp0 < 0 || rows < 0 || snp0 + rows > c->N) {
        set_error("dpca_copy_packed: bad arguments");
        return -1;
    }
    (void)hipSetDevice(c->dev);
    if (c->streamed && c->hmode == DPCA_HOST_FILE) {
        if (par_pread(c->fd, out, (size_t)(rows * c->np), (off_t)(3 + snp0 * c->np))) {
            set_error("dpca_copy_packed: reading the .bed file failed");
            return -1;
        }
    } else if (c->streamed) {
        for (long long r = 0; r < rows; r++)
            memcpy(out + r * c->np, c->hpacked + (snp0 + r) * c->ldp, (size_t)c->np);
    } else {
        DPCA_CHECK(hipMemcpy2D(out, c->np, c->packed + snp0 * c->ldp, c->ldp, c->np, rows,
                               hipMemcpyDeviceToHost));
    }
    // the device copy holds missing codes past sample M; the file holds 00
    const int valid = c->M - 4 * (int)(c->np - 1);
    if (valid < 4) {
        const uint8_t keep = (uint8_t)((1u << (2 * valid)) - 1u);
        for (long long r = 0; r < rows; r++) out[r * c->np + c->np - 1] &= keep;
    }
    return 0;
}

int dpca_pass(dpca_ctx *c, const double *Q, double *AAtQ)
{
    if (!c || !Q || !AAtQ) {
        set_error("dpca_pass: NULL argument");
        return -1;
    }
    (void)hipSetDevice(c->dev);
    int r;
    if ((r = copy_in_rhs(c, Q, c->RHS2))) return r;
    if ((r = do_pass(c, c->RHS2, c->ARHS))) return r;
    if ((r = copy_out_rhs(c, c->ARHS, AAtQ))) return r;
    collect_timers(c);
    return 0;
}

int dpca_run(dpca_ctx *c, double *rhs_inout, double *sigma, dpca_report *report)
{
    if (!c) {
        set_error("dpca_run: NULL context");
        return -1;
    }
    (void)hipSetDevice(c->dev);
    return run(c, rhs_inout, sigma, report);
}

int dpca_crit_history(dpca_ctx *c, double *vals, int cap)
{
    if (!c || cap < 0 || (cap > 0 && !vals)) {
        set_error("dpca_crit_history: bad arguments");
        return -1;
    }
    int n = std::min<int>(cap, (int)c->crit_hist.size());
    for (int i = 0; i < n; i++) vals[i] = c->crit_hist[i];
    return n;
}

int dpca_write_outputs(dpca_ctx *c, const char *prefix, const double *rhs, const double *sigma)
{
    if (!c || !prefix || !rhs || !sigma) {
        set_error("dpca_write_outputs: NULL argument");
        return -1;
    }
    const std::string base(prefix);
    FILE *f = fopen((base + "_singularValues.txt").c_str(), "w");
    if (!f) {
        set_error("fopen singularValues: %s", strerror(errno));
        return -1;
    }
    for (int j = 0; j < c->nsv; j++) fprintf(f, "%2.13lf\n", sigma[j]);
    fclose(f);
    f = fopen((base + "_leftSingularVectors.txt").c_str(), "w");
    if (!f) {
        set_error("fopen leftSingularVectors: %s", strerror(errno));
        return -1;
    }
    fprintf(f, "FID");
    for (int j = 0; j < c->nsv; j++) fprintf(f, "\tPC%d", j);
    fprintf(f, "\n");
    for (int i = 0; i < c->M; i++) {
        fprintf(f, "%8s", c->iids[i].c_str());
        for (int j = 0; j < c->nsv; j++) fprintf(f, "\t%2.13f", rhs[(size_t)i * c->k + j]);
        fprintf(f, "\n");
    }
    fclose(f);
    return 0;
}

int dpca_orthonormalize(const double *A, long long m, int k, double *Q, int *shifted, int device)
{
    if (!A || !Q || m <= 0 || k <= 0 || k > 64 || k > m) {
        set_error("dpca_orthonormalize: bad arguments");
        return -1;
    }
    DPCA_CHECK(hipSetDevice(device));
    dpca_ctx c;
    c.M = (int)std::min<long long>(m, 0x7fffffff);
    c.k = k;
    c.kp = (int)round_up(k, 16);
    c.Mpad = m;
    DPCA_CHECK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking));
    const size_t bytes = (size_t)m * c.kp * 8;
    int r = dalloc(&c.tmp, bytes) | dalloc(&c.Q, bytes) | dalloc(&c.ARHS, bytes) |
            dalloc(&c.gpart, (size_t)c.ngparts * 64 * 64 * 8) | dalloc(&c.G, 64 * 64 * 8) |
            dalloc(&c.Rinv, 64 * 64 * 8) | dalloc(&c.flags, 64) |
            dalloc(&c.cinfo, (kMaxKp + 1) * 8);
    c.robust = true;  // complete a rank-deficient A, as dgeqrf + dorgqr do
    int fl[16] = {0};
    if (!r) {
        hipError_t e = hipMemsetAsync(c.ARHS, 0, bytes, c.st);
        if (e == hipSuccess) e = hipMemsetAsync(c.flags, 0, 64, c.st);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(c.ARHS, c.kp * 8, A, k * 8, k * 8, m, hipMemcpyHostToDevice, c.st);
        if (e != hipSuccess) r = -2;
    }
    if (!r) r = orthonormalise(&c, c.ARHS);
    if (!r) {
        hipError_t e = hipMemcpy2DAsync(Q, k * 8, c.Q, c.kp * 8, k * 8, m, hipMemcpyDeviceToHost, c.st);
        if (e == hipSuccess) e = hipMemcpyAsync(fl, c.flags, sizeof fl, hipMemcpyDeviceToHost, c.st);
        if (e == hipSuccess) e = hipStreamSynchronize(c.st);
        if (e != hipSuccess) {
            set_error("dpca_orthonormalize: %s", hipGetErrorString(e));
            r = -2;
        }
    }
    collect_timers(&c);
    for (void *p : {(void *)c.tmp, (void *)c.Q, (void *)c.ARHS, (void *)c.gpart, (void *)c.G,
                    (void *)c.Rinv, (void *)c.flags, (void *)c.cinfo})
        if (p) (void)hipFree(p);
    c.tmp = c.Q = c.ARHS = c.gpart = c.G = c.Rinv = c.cinfo = nullptr;
    c.flags = nullptr;
    for (auto e : c.free_ev) (void)hipEventDestroy(e);
    c.free_ev.clear();
    (void)hipStreamDestroy(c.st);
    c.st = nullptr;
    if (shifted) *shifted = fl[0];
    if (!r && fl[1]) {
        set_error("CholeskyQR failed even with the shifted fallback");
        r = -4;
    }
    return r;
}

int dpca_syev(const double *B, int k, double *w, double *X, int device)
{
    if (!B || !w || !X || k <= 0 || k > 64) {
        set_error("dpca_syev: bad arguments");
        return -1;
    }
    DPCA_CHECK(hipSetDevice(device));
    const int kp = (int)round_up(k, 16);
    double *dB = nullptr, *dw = nullptr, *dX = nullptr;
    int *ds = nullptr;
    if (dalloc(&dB, kp * kp * 8) | dalloc(&dw, 64 * 8) | dalloc(&dX, kp * kp * 8) | dalloc(&ds, 64))
        return -3;
    std::vector<double> hb((size_t)kp * kp, 0.0), hx((size_t)kp * kp);
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) hb[(size_t)i * kp + j] = B[(size_t)i * k + j];
    int st = 0;
    hipError_t e = hipMemcpy(dB, hb.data(), hb.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ds, 0, 64);
    if (e == hipSuccess) e = launch_syev(nullptr, dB, k, kp, dw, dX, ds);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(w, dw, k * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hx.data(), dX, hx.size() * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&st, ds, sizeof st, hipMemcpyDeviceToHost);
    (void)hipFree(dB);
    (void)hipFree(dw);
    (void)hipFree(dX);
    (void)hipFree(ds);
    if (e != hipSuccess) {
        set_error("dpca_syev: %s", hipGetErrorString(e));
        return -2;
    }
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) X[(size_t)i * k + j] = hx[(size_t)i * kp + j];
    if (st) {
        set_error("Jacobi eigensolver did not converge");
        return -4;
    }
    return 0;
}

int dpca_bench_pass(dpca_ctx *c, const double *Q, int npass, double *ms)
{
    if (!c || !Q || !ms || npass < 1) {
        set_error("dpca_bench_pass: bad arguments");
        return -1;
    }
    DPCA_CHECK(hipSetDevice(c->dev));
    int r;
    if ((r = copy_in_rhs(c, Q, c->RHS2))) return r;
    DPCA_CHECK(hipStreamSynchronize(c->st));
    collect_timers(c);
    for (int i = 0; i < T_NKIND; i++) c->t[i] = 0;
    const double b0 = c->bytes_h2d;
    hipEvent_t e0 = get_event(c), e1 = get_event(c);
    const double h0 = now_s();
    DPCA_CHECK(hipEventRecord(e0, c->st));
    for (int i = 0; i < npass; i++)
        if ((r = do_pass(c, c->RHS2, c->ARHS))) return r;
    DPCA_CHECK(hipEventRecord(e1, c->st));
    DPCA_CHECK(hipStreamSynchronize(c->st));
    const double hwall = now_s() - h0;
    float tot = 0;
    DPCA_CHECK(hipEventElapsedTime(&tot, e0, e1));
    c->free_ev.push_back(e0);
    c->free_ev.push_back(e1);
    // per-kernel times: the stage events bracket the kernels of each stage
    double ka = 0, kb = 0, kc = 0, kd = 0;
    for (auto &tm : c->pending) {
        float t = 0;
        if (tm.a && tm.b && hipEventElapsedTime(&t, tm.a, tm.b) == hipSuccess) {
            if (tm.kind == T_MM_A) ka += t;
            if (tm.kind == T_MM_AT) kb += t;
            if (tm.kind == T_H2D) kc += t;
            if (tm.kind == T_MM) kd += t;
        }
    }
    collect_timers(c);
    ms[0] = tot / npass;
    ms[1] = ka / npass;
    ms[2] = kb / npass;
    ms[3] = ms[0] - ms[1] - ms[2];
    ms[4] = hwall * 1e3 / npass;
    ms[5] = kc / npass;
    ms[6] = kd / npass;
    ms[7] = c->streamed ? c->hmode : -1;
    ms[8] = (c->bytes_h2d - b0) / npass;
    return 0;
}

// Largest singular value of the k x k row-major matrix C: square root of the
// largest eigenvalue of C^T C by cyclic Jacobi (k <= 64).
static double largest_singular_value(const std::vector<double> &C, int k)
{
    std::vector<double> A((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
            double t = 0;
            for (int r = 0; r < k; r++) t += C[(size_t)r * k + i] * C[(size_t)r * k + j];
            A[(size_t)i * k + j] = t;
        }
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < k; i++)
            for (int j = 0; j < k; j++)
                (i == j ? diag : off) += A[(size_t)i * k + j] * A[(size_t)i * k + j];
        if (off <= 1e-30 * diag) break;
        for (int p = 0; p < k - 1; p++)
            for (int q = p + 1; q < k; q++) {
                const double apq = A[(size_t)p * k + q];
                if (apq == 0) continue;
                const double tau = (A[(size_t)q * k + q] - A[(size_t)p * k + p]) / (2 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                const double cs = 1 / sqrt(1 + t * t), sn = t * cs;
                for (int r = 0; r < k; r++) {  // A <- J^T A J
                    const double x = A[(size_t)r * k + p], y = A[(size_t)r * k + q];
                    A[(size_t)r * k + p] = cs * x - sn * y;
                    A[(size_t)r * k + q] = sn * x + cs * y;
                }
                for (int r = 0; r < k; r++) {
                    const double x = A[(size_t)p * k + r], y = A[(size_t)q * k + r];
                    A[(size_t)p * k + r] = cs * x - sn * y;
                    A[(size_t)q * k + r] = sn * x + cs * y;
                }
            }
    }
    double m = 0;
    for (int i = 0; i < k; i++) m = std::max(m, A[(size_t)i * k + i]);
    return sqrt(std::max(m, 0.0));
}

int dpca_full_svd(dpca_ctx *c, const double *rhs, const double *sigma, double *true_sigma,
                  double *true_U, double *sigma_relerr, double *vec_relerr, double *cos_values,
                  dpca_fullsvd_stats *st)
{
    if (!c || !rhs || !sigma || !true_sigma || !true_U || !sigma_relerr || !vec_relerr ||
        !cos_values || !st) {
        set_error("dpca_full_svd: NULL argument");
        return -1;
    }
    DPCA_CHECK(hipSetDevice(c->dev));
    const int M = c->M, k = c->nsv;
    const int64_t N = c->N;
    int r;
    if (!c->stats_done && (r = stream_stats_sweep(c))) return r;
    const double t0 = now_s();
    double *B = nullptr;
    if (dalloc(&B, (size_t)N * M * sizeof(double))) {
        set_error("dpca_full_svd: the dense %d x %lld FP64 matrix (%.1f GB) does not fit on the device",
                  M, (long long)N, (double)N * M * 8 / 1e9);
        return -3;
    }
    hipError_t e = hipSuccess;
    r = 0;
    if (!c->streamed) {
        e = launch_decode_f64(c->st, c->packed, N, c->ldp, M, c->zval, B);
    } else {
        for (int64_t r0 = 0; r0 < N && !r && e == hipSuccess; r0 += c->chunk_rows) {
            const int64_t rows = std::min(c->chunk_rows, c->Npad - r0);
            if ((r = fetch_block(c, 0, r0, rows, c->dbuf[0], c->st))) break;
            e = launch_decode_f64(c->st, c->dbuf[0], std::min(rows, N - r0), c->ldp, M,
                                  c->zval + 4 * r0, B + r0 * M);
        }
    }
    int info = 0;
    if (!r && e == hipSuccess) r = dense_top_svd(c->st, B, M, N, k, true_sigma, true_U, &info);
    (void)hipFree(B);
    if (e != hipSuccess) {
        set_error("dpca_full_svd: %s", hipGetErrorString(e));
        return -2;
    }
    if (r) return r;
    st->seconds = now_s() - t0;
    st->info = info;
    // comparisons of src/distpca.cpp:510-570 and src/utilities.cpp:234-270
    for (int j = 0; j < k; j++) {
        const double den = true_sigma[j] ? true_sigma[j] : 1.0;
        sigma_relerr[j] = fabs(true_sigma[j] - sigma[j]) / den;
    }
    std::vector<double> C((size_t)k * k, 0.0);
    for (int i = 0; i < M; i++)
        for (int a = 0; a < k; a++) {
            const double uh = rhs[(size_t)i * c->k + a], ut = true_U[(size_t)i * k + a];
            vec_relerr[(size_t)i * k + a] = fabs(fabs(ut) - fabs(uh)) / (fabs(ut) + 1e-12);
            for (int b = 0; b < k; b++) C[(size_t)a * k + b] += uh * true_U[(size_t)i * k + b];
        }
    double frob = 0;
    for (int a = 0; a < k; a++)
        for (int b = 0; b < k; b++) {
            const double x = C[(size_t)a * k + b];
            frob += (a == b) ? (fabs(x) - 1.0) * (fabs(x) - 1.0) : x * x;
        }
    st->frob_norm_angle = sqrt(frob);
    for (int j = 0; j < k; j++) {
        cos_values[j] = C[(size_t)j * k + j];
        C[(size_t)j * k + j] = fabs(C[(size_t)j * k + j]) - 1.0;
    }
    st->cos_error = largest_singular_value(C, k);
    return 0;
}

int dpca_write_fullsvd(const char *prefix, int M, int nsv, const double *true_U,
                       const double *true_sigma, const double *vec_relerr,
                       const double *cos_values)
{
    if (!prefix || !true_U || !true_sigma || !vec_relerr || !cos_values || M < 1 || nsv < 1) {
        set_error("dpca_write_fullsvd: bad arguments");
        return -1;
    }
    const std::string p(prefix);
    FILE *fvec = fopen((p + "_realLeftSingularVectors.txt").c_str(), "w");
    FILE *fval = fopen((p + "_realSingularValues.txt").c_str(), "w");
    FILE *ferr = fopen((p + "_singvecs_accuracy.txt").c_str(), "w");
    FILE *fcos = fopen((p + "_cosineValues.txt").c_str(), "w");
    int rc = 0;
    if (!fvec || !fval || !ferr || !fcos) {
        set_error("File error");
        rc = -1;
    } else {
        for (int i = 0; i < M; i++) {
            for (int j = 0; j < nsv; j++) fprintf(fvec, "%2.13f ", true_U[(size_t)i * nsv + j]);
            fprintf(fvec, "\n");
        }
        for (int j = 0; j < nsv; j++) fprintf(fval, "%2.13f\n", true_sigma[j]);
        for (int64_t e = 0; e < (int64_t)M * nsv; e++) fprintf(ferr, "%2.13lf\n", vec_relerr[e]);
        for (int j = 0; j < nsv; j++) fprintf(fcos, "%2.13lf\n", cos_values[j]);
    }
    FILE *fs[] = {fvec, fval, ferr, fcos};
    for (FILE *f : fs)
        if (f) fclose(f);
    return rc;
}

void dpca_close(dpca_ctx *c) { free_ctx(c); }

int dpca_footprint(int M, long long N, const dpca_params *p, long long *bytes)
{
    if (!p || !bytes) {
        set_error("dpca_footprint: NULL argument");
        return -1;
    }
    dpca_ctx c;
    c.M = M;
    c.N = N;
    int r = resolve_params(&c, p);
    if (r) return r;
    size_t total = 0;
    const int64_t packed = c.Npad * c.ldp;
    c.streamed = (c.p.device_cap > 0 && packed > c.p.device_cap);
    c.block_rows = c.Npad;
    if (c.streamed) {
        int64_t target = std::min<int64_t>(512ll << 20, c.p.device_cap / 4);
        c.chunk_rows = std::min(c.Npad, std::max<int64_t>(1024, target / c.ldp / 1024 * 1024));
        c.block_rows = c.chunk_rows;
        total += 2 * (size_t)(c.chunk_rows * c.ldp);
        const int64_t cb = c.chunk_rows * c.ldp, nblk = (c.Npad + c.chunk_rows - 1) / c.chunk_rows;
        total += (size_t)(std::max<int64_t>(0, std::min<int64_t>((c.p.device_cap - 2 * cb) / cb, nblk - 1)) * cb);
        if (c.p.host_mode == DPCA_HOST_FILE) total += 2 * (size_t)(c.chunk_rows * c.np);
    } else {
        total += (size_t)packed;
    }
    g_dry = &total;
    r = alloc_work(&c);
    g_dry = nullptr;
    if (!r && want_zt_image(&c, packed)) total += (size_t)packed;
    *bytes = (long long)total;
    return r;
}

int dpca_model_from_run(dpca_ctx *c, const double *rhs, const double *sigma, int nsv,
                        dpca_model **out)
{
    if (!c || !rhs || !sigma || !out || nsv < 1 || nsv > c->k) {
        set_error("dpca_model_from_run: bad arguments");
        return -1;
    }
    DPCA_CHECK(hipSetDevice(c->dev));
    for (int j = 0; j < nsv; j++)
        if (!(sigma[j] > 0)) {
            set_error("dpca_model_from_run: sigma[%d] is not positive", j);
            return -1;
        }
    dpca_model *m = new dpca_model();
    m->dev = c->dev;
    m->precision = c->p.precision;
    m->ns = c->ns;
    m->N = c->N;
    m->Npad = c->Npad;
    m->nsv = nsv;
    m->k = c->k;
    m->kp = c->kp;
    int r = 0;
    const size_t n1 = (size_t)c->Npad * sizeof(double);
    r |= dalloc(&m->zval, 4 * n1);
    r |= dalloc(&m->mu, n1);
    r |= dalloc(&m->dsc, n1);
    r |= dalloc(&m->W, n1 * c->kp);
    const int64_t prps = proj_rps_mx(c->Npad);
    const size_t uw8_bytes = (size_t)c->Npad / 8 * col_tiles(c->k, c->ns).vt() * 16 * sizeof(uint4);
    const size_t smax_bytes = (size_t)((c->Npad + prps - 1) / prps) * c->kp * sizeof(unsigned);
    if (m->precision == DPCA_MIXED) {
        r |= dalloc(&m->UW8, uw8_bytes);
        r |= dalloc(&m->smax, smax_bytes);
    }
    m->bytes = 6 * n1 + n1 * c->kp + (m->UW8 ? uw8_bytes + smax_bytes : 0);
    double *inv = nullptr;
    r |= dalloc(&inv, 64 * sizeof(double));
    if (r) {
        dpca_model_free(m);
        return -3;
    }
    // U = the first nsv Ritz vectors, padded; T = Z U in FP64 over the matrix
    std::vector<double> u((size_t)c->M * c->k, 0.0), is2(64, 0.0);
    for (int i = 0; i < c->M; i++)
        for (int j = 0; j < nsv; j++) u[(size_t)i * c->k + j] = rhs[(size_t)i * c->k + j];
    for (int j = 0; j < nsv; j++) is2[j] = 1.0 / (sigma[j] * sigma[j]);
    if (!c->stats_done && (r = stream_stats_sweep(c))) {
        dpca_model_free(m);
        (void)hipFree(inv);
        return r;
    }
    hipError_t e = hipSuccess;
    if ((r = copy_in_rhs(c, u.data(), c->RHS2)) == 0) {
        if (c->streamed) {
            const int prec = c->p.precision;
            c->p.precision = DPCA_FP64;  // loadings in FP64 whatever the form
            r = half_pass(c, c->RHS2);
            c->p.precision = prec;
        } else {
            e = launch_za_f64(c->st, c->packed, c->Npad, c->ldp, c->zval, c->RHS2, c->k, c->kp, c->T,
                              c->seg64, c->nseg64, c->Tseg);
        }
    }
    if (!r && e == hipSuccess) e = hipMemcpyAsync(inv, is2.data(), 64 * sizeof(double), hipMemcpyHostToDevice, c->st);
    if (!r && e == hipSuccess) e = launch_loadings(c->st, c->T, c->Npad, c->kp, nsv, inv, m->W);
    if (!r && e == hipSuccess) e = hipMemcpyAsync(m->zval, c->zval, 4 * n1, hipMemcpyDeviceToDevice, c->st);
    if (!r && e == hipSuccess) e = hipMemcpyAsync(m->mu, c->mu, n1, hipMemcpyDeviceToDevice, c->st);
    if (!r && e == hipSuccess) e = hipMemcpyAsync(m->dsc, c->dsc, n1, hipMemcpyDeviceToDevice, c->st);
    if (!r && e == hipSuccess && m->UW8)
        e = launch_uw8(c->st, m->W, c->Npad, c->k, c->ns, c->kp, m->mu, m->dsc, prps, m->smax, m->UW8,
                       false);
    if (!r && e == hipSuccess) e = hipStreamSynchronize(c->st);
    collect_timers(c);
    (void)hipFree(inv);
    if (r || e != hipSuccess) {
        if (e != hipSuccess) set_error("dpca_model_from_run: %s", hipGetErrorString(e));
        dpca_model_free(m);
        return r ? r : -2;
    }
    *out = m;
    return 0;
}

}  // extern "C"
