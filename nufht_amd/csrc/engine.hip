// Device engine of the MI355X NUFHT: builds a "program" (flat work lists for
// every stage over all boxes of one or many transforms), runs it on a HIP
// stream, and implements the device-side entry points of nufht_amd.h plus
// the host-pointer nufht_batch of the drop-in ABI.
//
// Stage order per program (one launch per stage, all transforms at once):
//   zero gs -> local projection -> local evaluation -> direct tiles
//   -> asymptotic spread -> rocFFT (one plan per grid-length class)
//   -> asymptotic interpolation
// which covers the reference's execute_boxes_batch (src/fast_hankel_nufht.c:
// 923-987) and nufht_batch dispatch (:989-1107) for every transform.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "program.h"

using namespace nufht_dev;

// ---------------------------------------------------------------------------
// errors
static thread_local char g_err[1024] = "";

extern "C" void nufht_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *nufht_amd_last_error(void) { return g_err; }
extern "C" const char *nufht_amd_version(void) { return "nufht-amd 0.1 gfx950"; }

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            nufht_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                            __FILE__, __LINE__);                                       \
            return -510;                                                               \
        }                                                                              \
    } while (0)

#define RC_TRY(expr)                   \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != 0) return rc_;      \
    } while (0)

// Phase times of program builds on stderr when NUFHT_AMD_TRACE is set.
static const bool g_trace = getenv("NUFHT_AMD_TRACE") != nullptr;
struct PhaseClock {
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char *what)
    {
        if (!g_trace) return;
        auto n = std::chrono::steady_clock::now();
        fprintf(stderr, "[nufht_amd] %-24s %8.2f ms\n", what,
                std::chrono::duration<double, std::milli>(n - t).count());
        t = n;
    }
};

// ---------------------------------------------------------------------------
// per-stage timing
static const char *kStageNames[ST_COUNT] = {"zero", "permute", "direct", "loc_project",
                                            "loc_eval", "asy_grid_zero", "asy_spread",
                                            "asy_spread_mfma", "asy_fft", "asy_interp",
                                            "asy_interp_mfma", "grad_combine", "asy_fused"};
struct PendingEv { hipEvent_t a, b; int stage; };
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<PendingEv> g_prof_pending;
static double g_prof_ms[ST_COUNT];
static long long g_prof_n[ST_COUNT];

bool nufht_dev::prof_enabled() { return g_prof_on; }

void nufht_dev::prof_push(hipEvent_t a, hipEvent_t b, int stage)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_pending.push_back({a, b, stage});
}

extern "C" int nufht_amd_profile(int enable)
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_on = enable != 0;
    return 0;
}

static void prof_drain()
{
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto &p : g_prof_pending) {
        hipEventSynchronize(p.b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        g_prof_ms[p.stage] += ms;
        g_prof_n[p.stage] += 1;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_prof_pending.clear();
}

extern "C" int nufht_amd_profile_read(double *ms, long long *launches, int max_stages,
                                      const char **names)
{
    prof_drain();
    std::lock_guard<std::mutex> g(g_prof_mu);
    int n = max_stages < ST_COUNT ? max_stages : ST_COUNT;
    for (int i = 0; i < n; ++i) {
        if (ms) ms[i] = g_prof_ms[i];
        if (launches) launches[i] = g_prof_n[i];
        if (names) names[i] = kStageNames[i];
    }
    return ST_COUNT;
}

extern "C" void nufht_amd_profile_reset(void)
{
    prof_drain();
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (int i = 0; i < ST_COUNT; ++i) { g_prof_ms[i] = 0.0; g_prof_n[i] = 0; }
}

// ---------------------------------------------------------------------------
// ES kernel parameters and the table of its Fourier transform
struct KernelTable {
    int kw;
    double beta, kmax;
    std::vector<double> coef;   // kNFT Chebyshev coefficients
};

static void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &w)
{
    x.resize(q);
    w.resize(q);
    for (int i = 0; i < q; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (q + 0.5)), p0 = 1, p1 = z, dp = 1;
        for (int it = 0; it < 100; ++it) {
            p0 = 1.0;
            p1 = z;
            for (int k = 2; k <= q; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = q * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        p0 = 1.0;
        p1 = z;
        for (int k = 2; k <= q; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = q * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

static void es_params(double tol, int mixed, int *kw, double *beta)
{
    // the FP32 grids of mixed mode keep the FP64 width: their rounding
    // (~1e-7 relative) then stays below the tolerances they serve (>= 1e-6)
    int w = (int)std::ceil(std::log10(1.0 / tol)) + 2;
    w = std::max(4, std::min(16, w));
    if (mixed) w = std::min(w, 9);
    *kw = w;
    *beta = 2.30 * w;
}

static const KernelTable &kernel_table(int kw, double beta)
{
    static std::mutex mu;
    static std::vector<KernelTable *> cache;
    std::lock_guard<std::mutex> g(mu);
    for (auto *t : cache)
        if (t->kw == kw && t->beta == beta) return *t;
    KernelTable *t = new KernelTable();
    t->kw = kw;
    t->beta = beta;
    t->kmax = kw * M_PI / 2.0;
    std::vector<double> gx, gw;
    gauss_legendre(200, gx, gw);
    const int Q = 128;
    std::vector<double> F(Q);
    for (int q = 0; q < Q; ++q) {
        double y = std::cos(M_PI * (q + 0.5) / Q);
        double k = t->kmax * std::sqrt(std::max(0.0, (y + 1.0) / 2.0));
        double s = 0.0;
        for (size_t a = 0; a < gx.size(); ++a) {
            double u = gx[a], tt = 1.0 - u * u;
            double es = tt > 0 ? std::exp(beta * (std::sqrt(tt) - 1.0)) : 0.0;
            s += gw[a] * es * std::cos(k * u);
        }
        F[q] = s;
    }
    t->coef.assign(kNFT + kTapMax * kEsN, 0.0);
    for (int n = 0; n < kNFT; ++n) {
        double s = 0.0;
        for (int q = 0; q < Q; ++q) s += F[q] * std::cos(n * M_PI * (q + 0.5) / Q);
        t->coef[n] = s * (n == 0 ? 1.0 : 2.0) / Q;
    }
    // kernel values at the taps as Chebyshev series in the cell offset
    for (int k = 0; k < kTapMax; ++k) {
        double fv[kEsN];
        for (int j = 0; j < kEsN; ++j) {
            double y = std::cos(M_PI * (j + 0.5) / kEsN), f = 0.5 * (y + 1.0);
            double u = (2.0 / kw) * (k - 0.5 * kw + f), tt = 1.0 - u * u;
            fv[j] = tt > 0 ? std::exp(beta * (std::sqrt(tt) - 1.0)) : 0.0;
        }
        for (int n = 0; n < kEsN; ++n) {
            double s = 0.0;
            for (int j = 0; j < kEsN; ++j) s += fv[j] * std::cos(n * M_PI * (j + 0.5) / kEsN);
            t->coef[kNFT + k * kEsN + n] = s * (n == 0 ? 1.0 : 2.0) / kEsN;
        }
    }
    cache.push_back(t);
    return *t;
}

// largest argument covered by a direct-sum Bessel table (unit intervals)
static const int kJTabMax = 8192;

// sources per grid cell from which a box uses the split-K spread kernel
static const double kDenseCell = 4.0;
// (NUFHT_AMD_DENSE_CELL overrides the threshold)
static double g_dense_cell = [] {
    const char *e = getenv("NUFHT_AMD_DENSE_CELL");
    return e ? atof(e) : kDenseCell;
}();
double nufht_dev::knobs_default_dense_cell() { return g_dense_cell; }
// (NUFHT_AMD_FUSED_MAX overrides the fused kernel's box size limit)
static int g_fused_max = [] {
    const char *e = getenv("NUFHT_AMD_FUSED_MAX");
    return e ? std::max(0, atoi(e)) : kFusedMax;
}();
int nufht_dev::knobs_default_fused_max() { return g_fused_max; }


// matrix-core spread: source density below which items stage their
// chunks' deconvolution factors (AsyItem::pad; measured on the uneven batch:
// spread 8.85 -> 8.18 ms per step at 16 to 64, the other workloads unchanged
// below 64 and slower when every item stages them)
constexpr double kDecDensity = 32.0;

// Two-pass factorisation Nf = N1 N2 for the FFT kernels (Nf = 2^a, 3 2^a or
// 9 2^a above 4096; the factor 3 or 9 goes into N1), sequences per block and
// radices.
static void fft_radices(int N, FftPass &p)
{
    p.nstage = 0;
    const int f3 = N % 9 == 0 ? 9 : N % 3 == 0 ? 3 : 1;
    N /= f3;
    int k = 0;
    while ((1 << k) < N) ++k;
    for (int q = 0; q < k / 3; ++q) p.radix[p.nstage++] = 8;
    if (k % 3 == 2) p.radix[p.nstage++] = 4;
    if (k % 3 == 1) p.radix[p.nstage++] = 2;
    if (f3 > 1) p.radix[p.nstage++] = f3;
}
static int fft_pick_m(int N, int other, int maxe)
{
    int m = 16;
    while (m > 1 && (m * N > maxe || other % m)) m >>= 1;
    return m;
}
// maxe: elements per block (fft_max_elems of the grid precision)
static bool fft_factor(int nf, int maxe, FftPass &p1, FftPass &p2)
{
    if (nf <= 4096) return false;
    const int f3 = nf % 9 == 0 ? 9 : nf % 3 == 0 ? 3 : 1;
    int q = nf / f3, a = 0;
    while ((1 << a) < q) ++a;
    if ((1 << a) != q) return false;
    // N2 = 2^j; the factorisation with the widest contiguous runs (min M)
    int bestN2 = 0, bestScore = -1;
    for (int j = 4; j <= 10; ++j) {
        int N2 = 1 << j;
        if (nf % N2) continue;
        int N1 = nf / N2;
        if (N1 < 16 || N1 > 1536) continue;
        int m1 = fft_pick_m(N1, N2, maxe), m2 = fft_pick_m(N2, N1, maxe);
        int score = 64 * std::min(m1, m2) + std::max(m1, m2);
        if (score > bestScore) { bestScore = score; bestN2 = N2; }
    }
    if (!bestN2) return false;
    int N2 = bestN2, N1 = nf / N2;
    for (FftPass *p : {&p1, &p2}) {
        p->N1 = N1;
        p->N2 = N2;
        p->nf = nf;
    }
    p1.N = N1;
    p1.M = fft_pick_m(N1, N2, maxe);
    p1.tiles = N2 / p1.M;
    p2.N = N2;
    p2.M = fft_pick_m(N2, N1, maxe);
    p2.tiles = N1 / p2.M;
    for (FftPass *p : {&p1, &p2}) {
        p->lgM = 0;
        while ((1 << p->lgM) < p->M) ++p->lgM;
        p->lgN = 0;
        while ((1 << p->lgN) < p->N) ++p->lgN;
    }
    fft_radices(N1, p1);
    fft_radices(N2, p2);
    return p1.M >= 4 && p2.M >= 4;
}
static void fft_table(std::vector<double> &tw, int n, int count, int *off)
{
    *off = (int)(tw.size() / 2);
    for (int m = 0; m < count; ++m) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        tw.push_back(std::cos(a));
        tw.push_back(std::sin(a));
    }
}

// ---------------------------------------------------------------------------
// per-plan device state
struct DevPlan {
    double *d_r = nullptr, *d_w = nullptr;
    int *d_perm_r = nullptr, *d_perm_w = nullptr;
};



// Identity of a plan's device state: order, options, sizes, device and
// 128 bits of hash of the caller's r and k arrays.
struct StateKey {
    double nu, tol, z_split;
    int max_levels, min_dim_prod, K_asy, K_loc, m, n, device;
    unsigned long long hr[2], hw[2];
};

struct EngineState {
    std::mutex mu;
    StateKey key;
    DevPlan dev;
    bool dev_ready = false;
    Knobs knobs;            // the plan's choices
    Program *prog = nullptr;
    int prog_B = 0;
    int prog_prec = -1;
    Knobs prog_knobs;       // the choices prog was built with
    // staging for host-pointer calls and permutations
    double *d_in = nullptr, *d_out = nullptr;
    size_t in_cap = 0, out_cap = 0;
    double *d_tmp[6] = {nullptr};
    size_t tmp_cap[6] = {0};
    // the last call's end on its stream: a call on another stream waits for
    // it before rewriting the program's descriptors or the staging buffers
    hipEvent_t done_ev = nullptr;
    hipStream_t done_st = nullptr;
};

// Orders a call on `st` after the previous call on this state (a no-op on
// the same stream).
static int order_after_last(EngineState *E, hipStream_t st)
{
    if (E->done_ev && E->done_st != st) HIP_TRY(hipStreamWaitEvent(st, E->done_ev, 0));
    return 0;
}

static int mark_done(EngineState *E, hipStream_t st)
{
    if (!E->done_ev) HIP_TRY(hipEventCreateWithFlags(&E->done_ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(E->done_ev, st));
    E->done_st = st;
    return 0;
}

static int ensure_buf(double **p, size_t *cap, size_t need)
{
    if (*cap >= need && *p) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc((void **)p, std::max<size_t>(need, 1) * sizeof(double)));
    *cap = need;
    return 0;
}

template <typename T>
static int upload(T **dptr, const std::vector<T> &v)
{
    if (*dptr) hipFree(*dptr);
    *dptr = nullptr;
    if (v.empty()) return 0;
    HIP_TRY(hipMalloc((void **)dptr, v.size() * sizeof(T)));
    HIP_TRY(hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static void free_program(Program *P)
{
    if (!P) return;
    hipFree(P->d_tfs); hipFree(P->d_zero); hipFree(P->d_dir); hipFree(P->d_dir_rng);
    hipFree(P->d_lbox);
    hipFree(P->d_lproj); hipFree(P->d_leval); hipFree(P->d_leval1); hipFree(P->d_lev_boxes);
    hipFree(P->d_leval1h); hipFree(P->d_lfold);
    hipFree(P->d_coef);
    hipFree(P->d_abox);
    hipFree(P->d_aspread); hipFree(P->d_ainterp); hipFree(P->d_grid); hipFree(P->d_ft);
    hipFree(P->d_grid2); hipFree(P->d_fftw); hipFree(P->d_fft_chinfo);
    hipFree(P->d_azero); hipFree(P->d_arange); hipFree(P->d_adense); hipFree(P->d_adrange);
    hipFree(P->d_funits); hipFree(P->d_fpass); hipFree(P->d_fqueue);
    hipFree(P->d_aimf);
    hipFree(P->d_aimt);
    hipFree(P->d_cheb);
    hipFree(P->d_jtab);
    if (P->side) {
        hipStreamSynchronize(P->side);
        hipStreamDestroy(P->side);
    }
    if (P->ev_fork) hipEventDestroy(P->ev_fork);
    if (P->ev_join) hipEventDestroy(P->ev_join);
    for (auto &c : P->fft) {
        if (c.plan) rocfft_plan_destroy(c.plan);
        if (c.info) rocfft_execution_info_destroy(c.info);
        if (c.work) hipFree(c.work);
    }
    delete P;
}

// Two independent 64-bit mixes (splitmix64 finaliser per word, chained).
static void hash_doubles(const double *x, int n, unsigned long long out[2])
{
    unsigned long long a = 0x243F6A8885A308D3ull ^ (unsigned long long)n,
                       b = 0x13198A2E03707344ull + (unsigned long long)n;
    for (int i = 0; i < n; ++i) {
        unsigned long long v;
        memcpy(&v, x + i, sizeof v);
        unsigned long long z = v + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        a = (a ^ z) * 0x100000001B3ull;
        a ^= a >> 29;
        b = (b + (z ^ (v >> 17))) * 0xD6E8FEB86659FD93ull;
        b ^= b >> 32;
    }
    out[0] = a;
    out[1] = b;
}

static StateKey state_key(const NufhtPlan *plan)
{
    const NufhtImplHost *h = (const NufhtImplHost *)plan->impl;
    StateKey k;
    memset(&k, 0, sizeof k);
    k.nu = plan->nu;
    k.tol = h->opt.tol;
    k.z_split = h->opt.z_split;
    k.max_levels = h->opt.max_levels;
    k.min_dim_prod = h->opt.min_dim_prod;
    k.K_asy = h->opt.K_asy;
    k.K_loc = h->opt.K_loc;
    k.m = plan->m;
    k.n = plan->n;
    hipGetDevice(&k.device);
    hash_doubles(plan->rs, plan->m, k.hr);
    hash_doubles(plan->ws, plan->n, k.hw);
    return k;
}

// Released device states kept for reuse by a later plan with the same key
// (the reference's Python entry points plan every call); most recent first,
// at most NUFHT_AMD_PLAN_CACHE entries (default 2, 0 disables).
static std::mutex g_cache_mu;
static std::vector<EngineState *> *g_cache = new std::vector<EngineState *>();
static int g_cache_cap = [] {
    const char *e = getenv("NUFHT_AMD_PLAN_CACHE");
    return e ? std::max(0, atoi(e)) : 2;
}();
static long long g_cache_hits = 0;

static EngineState *engine_of(const NufhtPlan *plan)
{
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    if (!h->dev) {
        StateKey k = state_key(plan);
        EngineState *E = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            for (size_t q = 0; q < g_cache->size(); ++q)
                if (!memcmp(&(*g_cache)[q]->key, &k, sizeof k)) {
                    E = (*g_cache)[q];
                    g_cache->erase(g_cache->begin() + q);
                    g_cache_hits++;
                    break;
                }
        }
        if (!E) {
            E = new EngineState();
            E->key = k;
        }
        h->dev = E;
    }
    return (EngineState *)h->dev;
}

static int ensure_dev(const NufhtPlan *plan, EngineState *E)
{
    if (E->dev_ready) return 0;
    PhaseClock clk;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    HIP_TRY(hipMalloc((void **)&E->dev.d_r, sizeof(double) * plan->m));
    HIP_TRY(hipMalloc((void **)&E->dev.d_w, sizeof(double) * plan->n));
    HIP_TRY(hipMemcpy(E->dev.d_r, h->r, sizeof(double) * plan->m, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(E->dev.d_w, h->w, sizeof(double) * plan->n, hipMemcpyHostToDevice));
    if (h->perm_r) {
        HIP_TRY(hipMalloc((void **)&E->dev.d_perm_r, sizeof(int) * plan->m));
        HIP_TRY(hipMemcpy(E->dev.d_perm_r, h->perm_r, sizeof(int) * plan->m, hipMemcpyHostToDevice));
    }
    if (h->perm_w) {
        HIP_TRY(hipMalloc((void **)&E->dev.d_perm_w, sizeof(int) * plan->n));
        HIP_TRY(hipMemcpy(E->dev.d_perm_w, h->perm_w, sizeof(int) * plan->n, hipMemcpyHostToDevice));
    }
    E->dev_ready = true;
    clk.mark("nodes to device");
    return 0;
}

static void engine_destroy(EngineState *E)
{
    free_program(E->prog);
    hipFree(E->dev.d_r); hipFree(E->dev.d_w); hipFree(E->dev.d_perm_r); hipFree(E->dev.d_perm_w);
    hipFree(E->d_in); hipFree(E->d_out);
    for (int i = 0; i < 6; ++i) hipFree(E->d_tmp[i]);
    if (E->done_ev) hipEventDestroy(E->done_ev);
    delete E;
}

extern "C" void nufht_engine_release(NufhtImplHost *h)
{
    EngineState *E = (EngineState *)h->dev;
    if (!E) return;
    h->dev = nullptr;
    std::vector<EngineState *> drop;
    if (E->prog && g_cache_cap > 0) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_cache->insert(g_cache->begin(), E);
        while ((int)g_cache->size() > g_cache_cap) {
            drop.push_back(g_cache->back());
            g_cache->pop_back();
        }
    } else {
        drop.push_back(E);
    }
    for (EngineState *d : drop) engine_destroy(d);
}

extern "C" int nufht_amd_plan_cache(int capacity, long long *hits)
{
    std::vector<EngineState *> drop;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        if (hits) *hits = g_cache_hits;
        if (capacity >= 0) {
            g_cache_cap = capacity;
            while ((int)g_cache->size() > g_cache_cap) {
                drop.push_back(g_cache->back());
                g_cache->pop_back();
            }
        }
    }
    for (EngineState *d : drop) engine_destroy(d);
    return 0;
}

extern "C" void nufht_engine_scratch_release(NufhtScratch *s) { (void)s; }

static int effective_prec(const NufhtPlan *plan)
{
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    if (h->precision == NUFHT_PREC_MIXED) return 2;
    if (h->precision == NUFHT_PREC_AUTO && plan->tol >= 1e-6) return 2;
    return 0;
}

// ---------------------------------------------------------------------------
// program construction
struct TfInput { const NufhtPlan *plan; int B; };

static void add_chunks_2d(std::vector<LocItem> &v, int box, int count, int chunk, int B, int cols)
{
    for (int c0 = 0; c0 < B; c0 += cols)
        for (int s = 0; s < count; s += chunk) v.push_back({box, s, std::min(chunk, count - s), c0});
}

// Local-expansion evaluation items.  Where the atomic adds saved are many,
// B (boxes per covered target - 1) >= 40 (the clustered workload: 32 x
// 1.6-1.8, where this form measured 1.12 -> 0.75 ms per step; the headline
// workload has 16 x 1.1 and measured 0.19 -> 0.34 ms), per
// segment of target rows covered by the same boxes, chunks of loc_tpb
// targets x loc_bmax columns that loop over all of those boxes (k_loc_eval:
// one atomic add per output element per item).  Otherwise per box, chunks of
// loc_tpb targets x all columns (k_loc_eval1: fewer registers; segments cut
// at every box boundary would leave many small items).
static void add_loc_eval_items(Program *P)
{
    std::map<int, std::vector<int>> bytf;
    for (size_t b = 0; b < P->lbox.size(); ++b) bytf[P->lbox[b].tf].push_back((int)b);
    for (auto &kv : bytf) {
        const int tf = kv.first, B = P->tfs[tf].B;
        const std::vector<int> &bl = kv.second;
        std::vector<int> cuts;
        long long cover = 0;
        for (int b : bl) {
            cuts.push_back(P->lbox[b].i0);
            cuts.push_back(P->lbox[b].i0 + P->lbox[b].ni);
            cover += P->lbox[b].ni;
        }
        std::sort(cuts.begin(), cuts.end());
        cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
        // targets covered at least once
        long long covered = 0;
        for (size_t q = 0; q + 1 < cuts.size(); ++q)
            for (int b : bl)
                if (P->lbox[b].i0 <= cuts[q] && P->lbox[b].i0 + P->lbox[b].ni >= cuts[q + 1]) {
                    covered += cuts[q + 1] - cuts[q];
                    break;
                }
        // atomics saved per target ~ B (boxes per covered target - 1)
        if (covered == 0 || (double)B * ((double)cover / (double)covered - 1.0) < 40.0) {
            // (half-integer orders: their own launch, 256 targets per item, no LDS)
            const bool half = P->tfs[tf].half != 0;
            const int tpb = half ? 256 : P->loc_tpb;
            for (int b : bl)
                for (int s = 0; s < P->lbox[b].ni; s += tpb)
                    (half ? P->leval1h : P->leval1).push_back({b, s, std::min(tpb, P->lbox[b].ni - s), 0});
            continue;
        }
        for (size_t q = 0; q + 1 < cuts.size(); ++q) {
            const int a = cuts[q], e = cuts[q + 1];
            const int roff = (int)P->lev_boxes.size();
            for (int b : bl)
                if (P->lbox[b].i0 <= a && P->lbox[b].i0 + P->lbox[b].ni >= e)
                    P->lev_boxes.push_back(b);
            const int rcnt = (int)P->lev_boxes.size() - roff;
            if (!rcnt) continue;
            for (int c0 = 0; c0 < B; c0 += P->loc_bmax)
                for (int i = a; i < e; i += P->loc_tpb)
                    P->leval.push_back({tf, i, std::min(P->loc_tpb, e - i), roff, rcnt, c0});
        }
    }
}

// Direct-sum items.  A target row is covered by several direct boxes (about
// 4 at the headline workload), each with few sources per row; one item per
// (row chunk, source slice, column chunk) of a row segment whose covering
// boxes are the same sums the sources of all of them, so that each output
// element receives one atomic add per slice instead of one per box.
constexpr int kDirSlice = 8192;      // sources per item (slices of long strips)

static void add_dir_items(Program *P, int tf, int B, const std::vector<Box> &dirb)
{
    std::vector<int> cuts;
    for (auto &b : dirb)
        if (b.i1 >= b.i0 && b.j1 >= b.j0) {
            cuts.push_back(b.i0);
            cuts.push_back(b.i1 + 1);
        }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    for (size_t q = 0; q + 1 < cuts.size(); ++q) {
        const int a = cuts[q], e = cuts[q + 1], ni = e - a;
        // the segment's source ranges (boxes tile the plane: disjoint in j)
        std::vector<int2> rg;
        for (auto &b : dirb)
            if (b.i1 >= b.i0 && b.j1 >= b.j0 && b.i0 <= a && b.i1 >= e - 1)
                rg.push_back(make_int2(b.j0, b.j1 - b.j0 + 1));
        if (rg.empty()) continue;
        std::sort(rg.begin(), rg.end(), [](const int2 &x, const int2 &y) { return x.x < y.x; });
        long long S = 0;
        for (auto &r : rg) S += r.y;
        // lanes per row: rows with few sources get one lane each (no
        // reduction); blocks of few rows cover all of them at once
        int tpr;
        if (ni >= 64) {
            tpr = S <= 32 ? 1 : (S <= 64 ? 2 : 4);
        } else {
            int p2 = 1;
            while (p2 < ni) p2 <<= 1;
            tpr = std::min(64, std::max(4, 256 / p2));
        }
        const int rows = 256 / tpr;
        // source slices of at most kDirSlice sources, as lists of ranges
        std::vector<std::pair<int, int>> slices;   // (offset in dir_rng, count)
        {
            size_t k = 0;
            int off = 0;
            while (k < rg.size()) {
                int roff = (int)P->dir_rng.size(), taken = 0;
                while (k < rg.size() && taken < kDirSlice) {
                    int n = std::min(rg[k].y - off, kDirSlice - taken);
                    P->dir_rng.push_back(make_int2(rg[k].x + off, n));
                    taken += n;
                    off += n;
                    if (off == rg[k].y) { ++k; off = 0; }
                }
                slices.push_back({roff, (int)P->dir_rng.size() - roff});
            }
        }
        for (int c0 = 0; c0 < B; c0 += kDirBM)
            for (int i = 0; i < ni; i += rows)
                for (auto &sl : slices)
                    P->dir.push_back({tf, a + i, std::min(rows, ni - i), sl.first, sl.second, tpr,
                                      c0, 0});
    }
}

// Interpolation items, per asymptotic box in order.  The box's targets are
// cut into 16-target tiles; a tile whose kernel supports fit a 16-position
// window (one position of slack; a 24-position form measured slower than
// the VALU kernel on the tiles it would add) goes to the matrix-core kernel,
// in runs of up to kIrTiles consecutive such tiles whose windows lie within
// kIrPos positions of the run's first window start; the other tiles form
// VALU items of up to 256/tb targets and tb columns.
static void add_interp_items(Program *P, const std::vector<TfInput> &in)
{
    for (size_t bi = 0; bi < P->abox.size(); ++bi) {
        if (P->afused[bi]) continue;
        const int box = (int)bi;
        const AsyBoxD &A = P->abox[bi];
        const TfDesc &T = P->tfs[A.tf];
        const NufhtImplHost *h = (const NufhtImplHost *)in[A.tf].plan->impl;
        int tb = 1;
        while (tb < T.B && tb < 16) tb <<= 1;
        const int tp = 256 / tb;
        auto m0_of = [&](int s) { return (int)std::ceil((h->w[A.i0 + s] - A.D - A.a2) / A.delta); };
        int run0 = -1;   // start of the pending VALU run of tiles
        auto flush_valu = [&](int upto) {
            for (int c0 = 0; run0 >= 0 && c0 < T.B; c0 += tb)
                for (int s = run0; s < upto; s += tp) {
                    const int n = std::min(tp, upto - s);
                    P->ainterp.push_back({box, s, n, c0, tb, 0});
                    P->stats.interp_tc_valu += (long long)n * A.E * std::min(tb, T.B - c0);
                }
            run0 = -1;
        };
        InterpRun cur = {box, 0, 0, 0, 0, 0};
        auto flush_run = [&]() {
            if (cur.count) {
                P->aimf.push_back(cur);
                P->stats.interp_tc_mfma += (long long)cur.count * A.E * T.B;
            }
            cur.count = 0;
        };
        const bool tiles = P->knobs.interp_mfma == 2;
        for (int s = 0; s < A.ni; s += 16) {
            const int cnt = std::min(16, A.ni - s);
            const int m0f = m0_of(s), m0l = m0_of(s + cnt - 1);
            // window positions the tile's taps need, with one of slack: at
            // most 16 for the one-tile items, 4 kIrSteps for the runs
            const int need = m0l - m0f + T.kw + 2;
            if (!(P->knobs.interp_mfma && need <= (tiles ? 16 : 4 * kIrSteps))) {
                flush_run();
                if (run0 < 0) run0 = s;
                continue;
            }
            flush_valu(s);
            // the tile's window: [mw, mw + 16) for the one-tile items; the run
            // kernel starts it at the device's first tap, which may differ from
            // the host's m0f by one either way, and reads 4 K steps (16
            // positions) or, for a wider tile, kIrSteps, so runs stage
            // [mw, mw + 18) or [mw, mw + 4 kIrSteps + 2)
            const int mw = m0f - 1;
            const int wdt = need <= 16 ? 18 : 4 * kIrSteps + 2;
            if (tiles) {
                // one item per tile, every column chunk
                P->aimt.push_back({box, s, cnt, 0, tb, mw});
                P->stats.interp_tc_mfma += (long long)cnt * A.E * T.B;
                continue;
            }
            if (cur.count && (cur.count == 16 * kIrTiles || mw + wdt - cur.mw0 > kIrPos ||
                              mw < cur.mw0))
                flush_run();
            if (!cur.count) {
                cur.start = s;
                cur.mw0 = mw;
                cur.npos = wdt;
            }
            cur.count += cnt;
            cur.npos = std::max(cur.npos, mw + wdt - cur.mw0);
        }
        flush_run();
        flush_valu(A.ni);
    }
}

// Local expansions the local kernels run: K_loc + 1 terms in at most
// kLocMaxNlt row tiles of 16, and a per-target scratch (loc_slots words) that
// 64 targets per block keep within 64 KiB of LDS.
constexpr int kLocSlotsMax = 128;
static bool loc_fits(int order, int half, int K)
{
    return K >= 0 && (K + 16) / 16 <= kLocMaxNlt && loc_slots(order, half, K) <= kLocSlotsMax;
}

// Whether asymptotic box A runs in the fused kernel: a uniform grid of at
// most kFusedNfMax points, at most fused_max sources and targets (one block
// takes each of its units), a kernel width the tap staging holds, and phases
// D x and w C below 2^20, which the kernel reduces without the library's
// full-range path.
static bool fused_fits(const Program *P, const AsyBoxD &A, const TfDesc &T)
{
    if (P->knobs.fused_max <= 0 || A.Nf > kFusedNfMax || A.nj > P->knobs.fused_max ||
        A.ni > P->knobs.fused_max || T.kw + 1 > kTapMax)
        return false;
    const double X = (A.L + 1) * A.h, S = M_PI / A.h;          // bounds of |x| and |s|
    return std::fabs(A.D) * X < 1048576.0 && (std::fabs(A.D) + S) * std::fabs(A.C) < 1048576.0;
}

// LDS layout of a fused unit's grid (Nf points x M channels): padded rows,
// or compact ones where the padding would pass the grid budget.
static void fused_layout(FftPass &p, size_t esz)
{
    p.lds_stride = p.M == 1 ? 1 : p.M + 1;
    p.lds_shift = 3;
    if (fft_lds_elems(p.N, p.lds_stride, p.lds_shift, p.M) * esz > (size_t)kFusedGridBytes) {
        p.lds_stride = p.M;
        p.lds_shift = 2;
    }
}

// Units of fused box `box` (column chunk tb): as many channels per unit as
// the block's lanes (2L + 1 positions x column pairs x exponent groups <=
// kFusedThreads) and the LDS grid budget allow.
static void add_fused_units(Program *P, int box, int tb, bool mixed,
                            std::map<std::pair<int, int>, int> &fpass_of)
{
    const AsyBoxD &A = P->abox[box];
    const TfDesc &T = P->tfs[A.tf];
    const size_t esz = mixed ? sizeof(float2) : sizeof(double2);
    const int lch = fused_lane_ch(mixed), Pn = 2 * A.L + 1;
    auto ntl_of = [](int t) { return t >= 2 ? t / 2 : 1; };
    int tbu = tb, ne = 0, M = 1;
    for (; tbu >= 1; tbu >>= 1) {
        if (Pn * ntl_of(tbu) > kFusedThreads) continue;
        const int eb = tbu >= 2 ? lch / 2 : lch;
        ne = std::min(A.E, (kFusedThreads / (Pn * ntl_of(tbu))) * eb);
        for (; ne >= 1; --ne) {
            M = 1;
            while (M < ne * tbu) M <<= 1;
            FftPass p;
            memset(&p, 0, sizeof(p));
            p.N = A.Nf;
            p.M = M;
            fused_layout(p, esz);
            if (fft_lds_elems(p.N, p.lds_stride, p.lds_shift, p.M) * esz <= (size_t)kFusedGridBytes) break;
        }
        if (ne >= 1) break;
    }
    auto key = std::make_pair(A.Nf, M);
    if (!fpass_of.count(key)) {
        FftPass p;
        memset(&p, 0, sizeof(p));
        p.N = A.Nf;
        p.nf = A.Nf;
        p.M = M;
        while ((1 << p.lgM) < M) ++p.lgM;
        fft_radices(A.Nf, p);
        fused_layout(p, esz);
        fpass_of[key] = (int)P->fpass.size();
        P->fpass.push_back(p);
        P->fused_grid_elems = std::max(P->fused_grid_elems,
                                       (int)fft_lds_elems(p.N, p.lds_stride, p.lds_shift, p.M));
    }
    const int fp = fpass_of[key];
    for (int c0 = 0; c0 < T.B; c0 += tbu)
        for (int e0 = 0; e0 < A.E; e0 += ne)
            P->funits.push_back({box, c0, e0, std::min(ne, A.E - e0), tbu, fp});
}

static int build_program(Program *P, const std::vector<TfInput> &in)
{
    PhaseClock clk;
    static std::once_flag fft_once;
    std::call_once(fft_once, []() { rocfft_setup(); });
    clk.mark("rocfft setup");
    P->prec = 0;
    for (auto &x : in) P->prec = std::max(P->prec, effective_prec(x.plan) == 2 ? 2 : 0);
    // a program mixes transforms only at one grid precision: FP64 wins
    for (auto &x : in)
        if (effective_prec(x.plan) == 0) P->prec = 0;
    bool mixed = P->prec == 2;
    memset(&P->stats, 0, sizeof(P->stats));
    std::vector<double> ft_all, cheb_all;
    std::map<std::pair<int, double>, int> ft_off;
    int kloc_max = 0, slots_max = 0, e_max = 0;
    struct AsyPending { AsyBoxD b; int tb; };
    std::vector<AsyPending> apend;
    for (size_t k = 0; k < in.size(); ++k) {
        const NufhtPlan *plan = in[k].plan;
        NufhtImplHost *h = (NufhtImplHost *)plan->impl;
        EngineState *E = engine_of(plan);
        int B = in[k].B;
        TfDesc T;
        memset(&T, 0, sizeof(T));
        T.r = E->dev.d_r;
        T.w = E->dev.d_w;
        T.m = plan->m;
        T.n = plan->n;
        T.B = B;
        double anu = std::fabs(plan->nu);
        long twice = std::lround(2.0 * anu);
        T.half = (int)(twice & 1);
        T.order = (int)(twice / 2);
        T.K_asy = h->exec_K_asy;
        T.K_loc = h->exec_K_loc;
        bool nu_int = !T.half;
        T.sign = (plan->nu < 0 && nu_int && (std::llround(plan->nu) & 1)) ? -1.0 : 1.0;
        for (int q = 0; q < kMaxA; ++q) T.a[q] = q < 64 ? plan->asy_coef[q] : 0.0;
        es_params(plan->tol, mixed, &T.kw, &T.beta);
        const KernelTable &kt = kernel_table(T.kw, T.beta);
        auto key = std::make_pair(T.kw, T.beta);
        if (!ft_off.count(key)) {
            ft_off[key] = (int)ft_all.size();
            ft_all.insert(ft_all.end(), kt.coef.begin(), kt.coef.end());
        }
        T.ft_off = ft_off[key];
        T.kmax = kt.kmax;
        int tf = (int)k;
        // zero items
        for (int t = 0; t < B; ++t)
            for (int s = 0; s < plan->n; s += 8192)
                P->zero.push_back({tf, t, s, std::min(8192, plan->n - s)});
        // direct boxes
        std::vector<Box> dirb, locb, asyb;
        if (h->use_direct) {
            dirb.push_back({0, plan->n - 1, 0, plan->m - 1});
        } else {
            const Boxes &bx = h->exec_boxes;
            for (size_t q = 0; q < bx.dir.length; ++q) dirb.push_back(bx.dir.data[q]);
            if (T.K_loc >= 0)
                for (size_t q = 0; q < bx.loc.length; ++q) locb.push_back(bx.loc.data[q]);
            for (size_t q = 0; q < bx.asy.length; ++q) asyb.push_back(bx.asy.data[q]);
        }
        // local boxes whose expansion exceeds what the local kernels hold (K_loc
        // up to 16 kLocMaxNlt - 1; per-target scratch of at most kLocSlotsMax
        // words, which high orders pass: order 60 at tol 1e-8 has K_loc = 207)
        // are summed directly instead
        if (!locb.empty() && !loc_fits(T.order, T.half, T.K_loc)) {
            dirb.insert(dirb.end(), locb.begin(), locb.end());
            locb.clear();
        }
        // direct-sum Bessel table over [0, max w r of the direct boxes]
        double xdir = 0.0;
        for (auto &b : dirb)
            if (b.i1 >= b.i0 && b.j1 >= b.j0) xdir = std::max(xdir, h->w[b.i1] * h->r[b.j1]);
        T.jtab_n = xdir > 0.0 ? (int)std::min(std::ceil(xdir) + 1.0, (double)kJTabMax) : 0;
        T.jtab_off = P->jtab_total;
        P->jtab_total += (long long)T.jtab_n * (kJDeg + 1);
        for (auto &b : dirb) {
            int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
            if (ni <= 0 || nj <= 0) continue;
            P->stats.pairs_dir += (long long)ni * nj;
            P->stats.n_dir++;
            P->stats.model_f64 += 2.0 * B * (double)ni * (double)nj;
            P->stats.model_dir += 2.0 * B * (double)ni * (double)nj;
        }
        // compulsory traffic: the nodes, the coefficients and the outputs
        P->stats.model_bytes += 8.0 * ((double)plan->m + plan->n) * (1.0 + B);
        add_dir_items(P, tf, B, dirb);
        if (!locb.empty()) {
            kloc_max = std::max(kloc_max, T.K_loc);
            slots_max = std::max(slots_max, loc_lds_slots(T.order, T.half, T.K_loc));
            if (T.half) {
                int K = T.K_loc, N = 2 * K + 2, p = T.order & 1;
                T.cheb_off = (int)cheb_all.size();
                for (int l = 0; l <= K; ++l)
                    for (int q = 0; q <= K; ++q) {
                        double th = M_PI * (q + 0.5) / N;
                        double v;
                        if (l == 0 && p == 0) v = 2.0 / N;
                        else v = (4.0 / N) * std::cos((2 * l + p) * th) / 2.0;
                        cheb_all.push_back(v);
                    }
                // the nodes x_q = cos(pi (q + 1/2) / N) after the matrix
                for (int q = 0; q <= K; ++q) cheb_all.push_back(std::cos(M_PI * (q + 0.5) / N));
            }
        }
        for (auto &b : locb) {
            int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
            if (ni <= 0 || nj <= 0) continue;
            LocBoxD L;
            L.tf = tf; L.i0 = b.i0; L.ni = ni; L.j0 = b.j0; L.nj = nj;
            L.coef_off = (int)P->ncoef;
            L.r_end = h->r[b.j1];
            P->ncoef += (long long)(T.K_loc + 1) * B;
            P->lbox.push_back(L);
            P->stats.n_loc++;
            P->stats.loc_src += nj;
            P->stats.loc_tgt += ni;
            P->stats.model_f64 += 2.0 * (T.K_loc + 1) * B * (double)(ni + nj);
            P->stats.model_loc += 2.0 * (T.K_loc + 1) * B * (double)(ni + nj);
        }
        int E2 = 2 * (T.K_asy + 1);
        int tb = 1;
        while (tb < B && tb < 16) tb <<= 1;
        for (auto &b : asyb) {
            int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
            if (ni <= 0 || nj <= 0) continue;
            double r0 = h->r[b.j0], r1 = h->r[b.j1], w0 = h->w[b.i0], w1 = h->w[b.i1];
            AsyBoxD A;
            memset(&A, 0, sizeof(A));
            A.tf = tf; A.i0 = b.i0; A.ni = ni; A.j0 = b.j0; A.nj = nj; A.E = E2;
            A.C = 0.5 * (r0 + r1);
            double X = 0.5 * (r1 - r0);
            A.D = 0.5 * (w0 + w1);
            double S = 0.5 * (w1 - w0);
            double Smin = 1e-12 * (std::fabs(A.D) + 1.0);
            if (S < Smin) S = Smin;
            const double sigma = 2.0;
            A.h = M_PI / (sigma * S);
            RC_TRY(nufht_asy_grid(X, A.h, T.kw, &A.L, &A.Nf));
            A.delta = 2.0 * M_PI / ((double)A.Nf * A.h);
            A.a1 = 0.5 * T.kw * A.h;
            A.a2 = 0.5 * T.kw * A.delta;
            A.rho = r0;
            e_max = std::max(e_max, E2);
            apend.push_back({A, tb});
            P->stats.n_asy++;
            P->stats.asy_src += nj;
            P->stats.asy_tgt += ni;
            P->stats.asy_channels += (long long)E2 * B;
            const double nfb = 4.0 * X * S / M_PI + T.kw + 1.0;
            P->stats.asy_fine += (double)E2 * B * nfb;
            P->stats.asy_fine_nlogn += (double)E2 * B * nfb * std::log2(nfb);
            const double C = (double)E2 * B;
            const double fn = 2.0 * T.kw * C * nj + 4.0 * T.kw * C * ni + 2.5 * C * nfb * std::log2(nfb);
            (mixed ? P->stats.model_f32 : P->stats.model_f64) += fn;
            P->stats.model_bytes += (mixed ? 8.0 : 16.0) * C * nfb;
        }
        P->tfs.push_back(T);
    }
    // small boxes whose sources and targets fit the fused kernel's units run
    // there (no HBM grid, no FFT class)
    std::vector<char> fz(apend.size(), 0);
    for (size_t q = 0; q < apend.size(); ++q) fz[q] = fused_fits(P, apend[q].b, P->tfs[apend[q].b.tf]);
    // the other asymptotic boxes grouped by FFT length; classes laid out back to back
    std::map<int, std::vector<int>> byNf;
    for (size_t q = 0; q < apend.size(); ++q)
        if (!fz[q]) byNf[apend[q].b.Nf].push_back((int)q);
    long long off = 0;
    std::map<int, bool> own_nf;
    for (auto &kv : byNf) {
        FftClass c;
        c.nf = kv.first;
        c.off = off;
        const int maxe = mixed ? fft_max_elems<float>() : fft_max_elems<double>();
        c.own = P->knobs.own_fft && fft_factor(c.nf, maxe, c.p1, c.p2);
        if (P->knobs.own_fft && !c.own && c.nf <= 4096 && c.nf <= maxe) {
            // one pass: M whole channels of Nf points per block
            c.own = c.small = true;
            c.p1.N = c.p1.nf = c.nf;
            c.p1.M = 1;
            while (c.p1.M < 16 && 2 * c.p1.M * c.nf <= maxe) c.p1.M <<= 1;
            c.p1.lgM = 0;
            while ((1 << c.p1.lgM) < c.p1.M) ++c.p1.lgM;
            fft_radices(c.nf, c.p1);
        }
        own_nf[c.nf] = c.own;
        c.p1.chinfo_off = c.p2.chinfo_off = (int)P->fft_chinfo.size();
        for (int q : kv.second) {
            AsyBoxD &A = apend[q].b;
            const TfDesc &T = P->tfs[A.tf];
            A.goff = off;
            long long ch = (long long)A.E * T.B;
            off += ch * A.Nf;
            c.batch += ch;
            if (c.own) {
                // the transformed grid is read only at frequencies [mlo, mhi]
                // (mod Nf) by the interpolation of the box's targets
                const NufhtImplHost *h = (const NufhtImplHost *)in[A.tf].plan->impl;
                double smin = h->w[A.i0] - A.D, smax = h->w[A.i0 + A.ni - 1] - A.D;
                int mlo = (int)std::ceil((smin - A.a2) / A.delta) - 4;
                // (up to the last target's first tap + 4 kIrSteps + 2: an
                // interpolation run's window may reach that far past the
                // supports)
                int mhi = (int)std::ceil((smax - A.a2) / A.delta) + std::max(T.kw + 4, 4 * kIrSteps + 2);
                if (mhi - mlo + 1 >= A.Nf) { mlo = 0; mhi = A.Nf - 1; }
                int4 info = make_int4(A.L, mlo, mhi - mlo, 0);
                for (long long k = 0; k < ch; ++k) P->fft_chinfo.push_back(info);
                P->stats.fft_bytes += (double)ch * (mixed ? 8.0 : 16.0) *
                                      ((2.0 * A.L + 1.0) + 2.0 * A.Nf + (mhi - mlo + 1.0));
            } else {
                P->stats.fft_bytes += (double)ch * (mixed ? 8.0 : 16.0) * 2.0 * A.Nf;
            }
        }
        c.p1.nch = c.p2.nch = (int)c.batch;
        if (g_trace)
            fprintf(stderr, "[nufht_amd] fft class Nf %d channels %lld %s N1 %d M1 %d N2 %d M2 %d\n",
                    c.nf, c.batch, !c.own ? "rocfft" : c.small ? "one-pass" : "two-pass",
                    c.small ? c.nf : c.p1.N1, c.p1.M, c.small ? 0 : c.p2.N2, c.small ? 0 : c.p2.M);
        P->fft.push_back(c);
    }
    P->grid_elems = off;
    // sources per matrix-core spread item: the knob, or (0) 512 where every
    // dense box has kw <= 8 (tol >= 1e-6; the clustered workload's spread
    // 3.08 -> 2.73 ms per step in mixed precision, 4.06 -> 3.87 in FP64,
    // against 1024) and 1024 otherwise (the headline: 3.02, against 3.21 at
    // 512); the autotuner tries others
    int dense_split = P->knobs.dense_split;
    if (dense_split == 0) {
        int kwd = 0;
        for (auto &ap : apend) {
            const AsyBoxD &A = ap.b;
            const TfDesc &T = P->tfs[A.tf];
            if (T.kw <= 13 && (double)A.nj >= P->knobs.dense_cell * (2.0 * A.L + 1.0))
                kwd = std::max(kwd, T.kw);
        }
        dense_split = kwd <= 8 ? 512 : 1024;
    }
    std::map<std::pair<int, int>, int> fpass_of;
    for (size_t q = 0; q < apend.size(); ++q) {
        const AsyPending &ap = apend[q];
        int box = (int)P->abox.size();
        const AsyBoxD &A = ap.b;
        const TfDesc &T = P->tfs[A.tf];
        const NufhtImplHost *h = (const NufhtImplHost *)in[A.tf].plan->impl;
        P->abox.push_back(A);
        P->afused.push_back(fz[q]);
        if (fz[q]) {
            add_fused_units(P, box, ap.tb, mixed, fpass_of);
            P->stats.fused_boxes++;
            P->stats.fused_sc += (long long)A.nj * A.E * T.B;
            P->stats.fused_tc += (long long)A.ni * A.E * T.B;
            P->stats.asy_grid += A.Nf;
            P->stats.asy_grid_points += (long long)A.E * T.B * A.Nf;
            continue;
        }
        int tb = ap.tb;
        // spread positions per VALU item: 256 lanes over column pairs x exponent groups
        int sp = std::max(1, 256 / spread_valu_lanes(A.E, tb, mixed));
        const double *rb = h->r + A.j0, *re = h->r + A.j0 + A.nj;
        // source-dense boxes (many sources per grid cell) use the split-K kernel
        // (the matrix-core spread slides a 16-position window: kw <= 13 keeps
        // the kw + 1 taps of a source inside it with room to move)
        bool dense = T.kw <= 13 && (double)A.nj >= P->knobs.dense_cell * (2.0 * A.L + 1.0);
        // sources whose kernel support reaches grid positions [s, s + np)
        auto src_range = [&](int s, int np) {
            int lmin = INT32_MAX, lmax = INT32_MIN;
            for (int p = s; p < s + np; ++p) {
                int l = p <= A.L ? p : p - A.Nf;
                if (l < -A.L || l > A.L) continue;
                lmin = std::min(lmin, l);
                lmax = std::max(lmax, l);
            }
            // a source at x has taps at l = lf .. lf + kw, lf = ceil((x - a1) / h):
            // it reaches [lmin, lmax] iff (lmin - kw - 1) h + a1 < x <= lmax h + a1;
            // a quarter cell of margin on each side absorbs rounding (a source the
            // device places one tap differently has a kernel value ~e^-beta there)
            int2 jr = make_int2(0, 0);
            if (lmin <= lmax) {
                double xlo = A.C + (lmin - T.kw - 1.25) * A.h + A.a1, xhi = A.C + (lmax + 0.25) * A.h + A.a1;
                jr.x = A.j0 + (int)(std::lower_bound(rb, re, xlo) - rb);
                jr.y = A.j0 + (int)(std::upper_bound(rb, re, xhi) - rb);
            }
            return jr;
        };
        for (int c0 = 0; c0 < T.B; c0 += tb) {
            long long chans = (long long)A.E * std::min(tb, T.B - c0);
            (dense ? P->stats.spread_sc_mfma : P->stats.spread_sc_valu) += (long long)A.nj * chans;
            // position ranges to spread: the whole grid, or for the two-pass
            // FFT (which reads only l in [-L, L]) just [0, L] and [Nf - L, Nf)
            std::vector<std::pair<int, int>> pr;
            if (own_nf[A.Nf]) {
                pr.push_back({0, A.L + 1});
                pr.push_back({A.Nf - A.L, A.Nf});
            } else {
                pr.push_back({0, A.Nf});
            }
            if (dense) {
                // the box's grid rows are zeroed, then pieces of dense_split
                // consecutive sources each write the positions their kernels
                // reach: plain stores where no other piece reaches (host and
                // device first taps may differ by one: a margin of one
                // position), atomic adds elsewhere
                for (auto &rg : pr)
                    for (int s = rg.first; s < rg.second; s += 256)
                        P->azero.push_back({box, s, std::min(256, rg.second - s), c0, tb, 0});
                auto lf_of = [&](int j) {
                    return (int)std::ceil((h->r[j] - A.C - A.a1) / A.h);
                };
                const int split = dense_split;
                // boxes with fewer than kDecDensity sources per grid position
                // move their window often: their items stage the
                // deconvolution factors of each chunk for the flushes
                const int stage_dec = (double)A.nj < kDecDensity * (2.0 * A.L + 1.0) ? 1 : 0;
                const int npiece = (A.nj + split - 1) / split;
                for (int pk = 0; pk < npiece; ++pk) {
                    int j0 = A.j0 + pk * split, j1 = std::min(A.j0 + A.nj, j0 + split);
                    SpreadRange sr;
                    sr.piece = make_int2(j0, j1);
                    sr.wr = make_int2(std::max(-A.L, lf_of(j0) - 1),
                                      std::min(A.L, lf_of(j1 - 1) + T.kw + 1));
                    sr.ex = make_int2(pk > 0 ? lf_of(j0 - 1) + T.kw + 2 : sr.wr.x,
                                      pk + 1 < npiece ? lf_of(j1) - 2 : sr.wr.y);
                    sr.pad = make_int2(0, 0);
                    P->adense.push_back({box, 0, 0, c0, tb, stage_dec});
                    P->adrange.push_back(sr);
                }
            } else {
                for (auto &rg : pr)
                for (int s = rg.first; s < rg.second; s += sp) {
                    int np = std::min(sp, rg.second - s);
                    int2 jr = src_range(s, np);
                    if (jr.y - jr.x <= P->knobs.spread_split) {
                        P->aspread.push_back({box, s, np, c0, tb, 0});
                        P->arange.push_back(jr);
                    } else {
                        P->azero.push_back({box, s, np, c0, tb, 0});
                        for (int j = jr.x; j < jr.y; j += P->knobs.spread_split) {
                            P->aspread.push_back({box, s, np, c0, tb, 1});
                            P->arange.push_back(make_int2(j, std::min(jr.y, j + P->knobs.spread_split)));
                        }
                    }
                }
            }
        }
        P->stats.asy_grid += A.Nf;
        P->stats.asy_grid_points += (long long)A.E * T.B * A.Nf;
    }
    add_interp_items(P, in);
    // fused units in decreasing cost (the queue hands out large units first)
    if (!P->funits.empty()) {
        std::vector<double> cost(P->funits.size());
        std::vector<size_t> order(cost.size());
        for (size_t q = 0; q < cost.size(); ++q) {
            const FusedUnit &u = P->funits[q];
            const AsyBoxD &A = P->abox[u.box];
            const int kw = P->tfs[A.tf].kw;
            const double ch = (double)u.ne * u.tb;
            cost[q] = ch * ((kw + 1.0) * A.nj + 1.6 * kw * A.ni + 5.0 * A.Nf * std::log2((double)A.Nf)) +
                      400.0 * (A.nj + A.ni);
            order[q] = q;
        }
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
        std::vector<FusedUnit> u2;
        u2.reserve(order.size());
        for (size_t q : order) u2.push_back(P->funits[q]);
        P->funits.swap(u2);
    }
    P->stats.fft_classes = (long long)P->fft.size();
    P->emax = e_max <= 6 ? 6 : e_max <= 10 ? 10 : e_max <= 14 ? 14 : 22;
    // matrix-core spread items grouped (stably: box order kept within a
    // group) by products per consumer wave, one launch per group
    {
        const int kwmax = (2 * P->emax + 3) / 4;
        std::vector<int> kw(P->adense.size());
        for (size_t q = 0; q < kw.size(); ++q)
            kw[q] = spread_mfma_kw(P->abox[P->adense[q].box].E, P->adense[q].tb);
        std::vector<size_t> order(kw.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return kw[x] < kw[y]; });
        std::vector<AsyItem> it2;
        std::vector<SpreadRange> rg2;
        it2.reserve(order.size());
        rg2.reserve(order.size());
        P->adense_off.assign(kwmax + 2, 0);
        for (size_t q : order) {
            it2.push_back(P->adense[q]);
            rg2.push_back(P->adrange[q]);
            ++P->adense_off[kw[q] + 1];
        }
        for (int k = 1; k <= kwmax + 1; ++k) P->adense_off[k] += P->adense_off[k - 1];
        P->adense.swap(it2);
        P->adrange.swap(rg2);
    }
    // split (producer / consumer) schedule of the matrix-core spread unless
    // the knob asks for the unified one: it measured faster on every workload
    // (the clustered one in mixed precision: 3.90 -> 3.53 ms per step)
    P->spread_split = P->knobs.spread_sched != 0;
    // local-expansion launch geometry: projection items of loc_ch sources x
    // 16 columns with ceil((K+1)/16) row tiles; evaluation items of loc_tpb
    // targets x all columns, the per-target scratch within 48 KiB of LDS
    if (!P->lbox.empty()) {
        P->loc_nlt = (kloc_max + 16) / 16;
        if (P->loc_nlt > kLocMaxNlt) {
            // (loc_fits sends such boxes to the direct sums: not reached)
            nufht_set_error("local expansion with K_loc = %d exceeds the supported %d", kloc_max,
                            16 * kLocMaxNlt - 1);
            return -540;
        }
        P->loc_tpb = 256;
        while (P->loc_tpb > 64 && (size_t)slots_max * P->loc_tpb * sizeof(double) > 48 * 1024)
            P->loc_tpb >>= 1;
        P->loc_lds_e = (size_t)slots_max * P->loc_tpb * sizeof(double);
        if (P->loc_lds_e > 160 * 1024) {
            nufht_set_error("local expansion scratch of %d words per target exceeds LDS", slots_max);
            return -540;
        }
        if (P->loc_lds_e > 64 * 1024) {
            for (const void *f : {(const void *)k_loc_eval1<double>, (const void *)k_loc_eval1<float>})
                HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)P->loc_lds_e));
            for (const void *f : {(const void *)k_loc_eval<16, double>, (const void *)k_loc_eval<32, double>,
                                  (const void *)k_loc_eval<16, float>, (const void *)k_loc_eval<32, float>})
                HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)P->loc_lds_e));
        }
        for (size_t b = 0; b < P->lbox.size(); ++b) {
            const TfDesc &T = P->tfs[P->lbox[b].tf];
            add_chunks_2d(P->lproj, (int)b, P->lbox[b].nj, P->loc_ch, T.B, kLocBM);
            if (T.B > 16) P->loc_bmax = 32;
            if (T.half) P->lfold.push_back((int)b);
        }
        add_loc_eval_items(P);
        P->stats.loc_eval_multi = (long long)P->leval.size();
    }
    if (!P->tfs.empty()) {
        P->stats.kernel_width = P->tfs[0].kw;
        P->stats.K_asy = P->tfs[0].K_asy;
        P->stats.K_loc = P->tfs[0].K_loc;
        P->stats.z_split = ((NufhtImplHost *)in[0].plan->impl)->exec_z;
    }
    P->stats.precision = P->prec;
    clk.mark("work lists (host)");
    // uploads
    RC_TRY(upload(&P->d_zero, P->zero));
    {
        // direct items grouped (stably) by column class: <= 4, <= 8, <= 16 columns
        auto cls = [&](const DirItem &d) {
            const int nb = std::min(kDirBM, P->tfs[d.tf].B - d.col0);
            return nb <= 4 ? 0 : nb <= 8 ? 1 : 2;
        };
        std::vector<DirItem> d2;
        d2.reserve(P->dir.size());
        for (int c = 0; c < 3; ++c) {
            P->dir_off[c] = (int)d2.size();
            for (auto &d : P->dir)
                if (cls(d) == c) d2.push_back(d);
        }
        P->dir_off[3] = (int)d2.size();
        P->dir.swap(d2);
    }
    RC_TRY(upload(&P->d_dir, P->dir));
    RC_TRY(upload(&P->d_dir_rng, P->dir_rng));
    RC_TRY(upload(&P->d_lbox, P->lbox));
    RC_TRY(upload(&P->d_lproj, P->lproj));
    RC_TRY(upload(&P->d_leval, P->leval));
    RC_TRY(upload(&P->d_leval1, P->leval1));
    RC_TRY(upload(&P->d_leval1h, P->leval1h));
    RC_TRY(upload(&P->d_lfold, P->lfold));
    RC_TRY(upload(&P->d_lev_boxes, P->lev_boxes));
    RC_TRY(upload(&P->d_abox, P->abox));
    RC_TRY(upload(&P->d_aspread, P->aspread));
    RC_TRY(upload(&P->d_ainterp, P->ainterp));
    RC_TRY(upload(&P->d_azero, P->azero));
    RC_TRY(upload(&P->d_arange, P->arange));
    RC_TRY(upload(&P->d_adense, P->adense));
    if (g_trace && !P->aimf.empty()) {
        long long hist[kIrTiles + 1] = {0};
        for (auto &r : P->aimf) ++hist[(r.count + 15) / 16];
        fprintf(stderr, "[nufht_amd] interpolation runs %zu, by tiles:", P->aimf.size());
        for (int t = 1; t <= kIrTiles; ++t) fprintf(stderr, " %lld", hist[t]);
        fprintf(stderr, "\n");
    }
    RC_TRY(upload(&P->d_aimf, P->aimf));
    RC_TRY(upload(&P->d_aimt, P->aimt));
    RC_TRY(upload(&P->d_adrange, P->adrange));
    RC_TRY(upload(&P->d_funits, P->funits));
    if (!P->funits.empty()) {
        int dev = 0, ncu = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_TRY(hipMalloc((void **)&P->d_fqueue, sizeof(int)));
        P->fused_blocks = (int)std::min<size_t>(P->funits.size(), (size_t)std::max(1, 2 * ncu));
    }
    for (auto &it : P->ainterp) P->interp_tmax = std::max(P->interp_tmax, 256 / it.tb);
    RC_TRY(upload(&P->d_ft, ft_all));
    if (cheb_all.empty()) cheb_all.push_back(0.0);
    RC_TRY(upload(&P->d_cheb, cheb_all));
    if (P->ncoef) HIP_TRY(hipMalloc((void **)&P->d_coef, P->ncoef * sizeof(double)));
    HIP_TRY(hipMalloc((void **)&P->d_jtab, std::max<long long>(1, P->jtab_total) * sizeof(double)));
    for (auto &T : P->tfs)
        if (T.jtab_n > 0)
            hipLaunchKernelGGL(k_jtab_build, dim3(T.jtab_n), dim3(kJDeg + 1), 0, nullptr, T.order,
                               T.half, T.jtab_n, P->d_jtab + T.jtab_off);
    HIP_TRY(hipDeviceSynchronize());
    clk.mark("uploads + J tables");
    HIP_TRY(hipMalloc((void **)&P->d_tfs, std::max<size_t>(1, P->tfs.size()) * sizeof(TfDesc)));
    size_t esz = mixed ? sizeof(float2) : sizeof(double2);
    if (P->grid_elems) {
        HIP_TRY(hipMalloc(&P->d_grid, P->grid_elems * esz));
        HIP_TRY(hipMalloc(&P->d_grid2, P->grid_elems * esz));
    }
    clk.mark("grid allocation");
    // two-pass FFT classes: twiddle tables (in the grid's precision)
    std::vector<double> tw;
    for (auto &c : P->fft) {
        if (!c.own) continue;
        fft_table(tw, c.p1.N, c.p1.N, &c.p1.w_off);
        if (c.small) {
            FftPass &fp = c.p1;
            fp.lds_stride = fp.M == 1 ? 1 : fp.M + 1;
            fp.lds_shift = 3;
            if (fft_lds_elems(fp.N, fp.lds_stride, fp.lds_shift, fp.M) * esz > 80 * 1024) {
                fp.lds_stride = fp.M;
                fp.lds_shift = 2;
            }
            c.lds1 = fft_lds_elems(fp.N, fp.lds_stride, fp.lds_shift, fp.M) * esz;
            continue;
        }
        fft_table(tw, c.p2.N, c.p2.N, &c.p2.w_off);
        // w^a = hi[a >> 8] lo[a & 255], w = e^{2 pi i / Nf}
        c.p1.hi_off = (int)(tw.size() / 2);
        for (int j = 0; j < (c.nf + 255) / 256; ++j) {
            double a = 2.0 * M_PI * (256.0 * j) / (double)c.nf;
            tw.push_back(std::cos(a));
            tw.push_back(std::sin(a));
        }
        fft_table(tw, c.nf, 256, &c.p1.lo_off);
        c.p2.hi_off = c.p1.hi_off;
        c.p2.lo_off = c.p1.lo_off;
        // LDS layout of each pass: padded rows unless that keeps a second
        // block off the CU (above 80 KiB), then the compact layout
        for (FftPass *fp : {&c.p1, &c.p2}) {
            fp->lds_stride = fp->M + 1;
            fp->lds_shift = 3;
            if (fft_lds_elems(fp->N, fp->lds_stride, fp->lds_shift, fp->M) * esz > 80 * 1024) {
                fp->lds_stride = fp->M;
                fp->lds_shift = 2;
            }
        }
        c.lds1 = fft_lds_elems(c.p1.N, c.p1.lds_stride, c.p1.lds_shift, c.p1.M) * esz;
        c.lds2 = fft_lds_elems(c.p2.N, c.p2.lds_stride, c.p2.lds_shift, c.p2.M) * esz;
    }
    {
        std::map<int, int> w_of;
        for (auto &p : P->fpass) {
            if (!w_of.count(p.N)) {
                int off;
                fft_table(tw, p.N, p.N, &off);
                w_of[p.N] = off;
            }
            p.w_off = w_of[p.N];
        }
        RC_TRY(upload(&P->d_fpass, P->fpass));
    }
    if (!P->fft_chinfo.empty()) RC_TRY(upload(&P->d_fft_chinfo, P->fft_chinfo));
    if (!tw.empty()) {
        if (mixed) {
            std::vector<float> twf(tw.begin(), tw.end());
            RC_TRY(upload((float **)&P->d_fftw, twf));
        } else {
            RC_TRY(upload((double **)&P->d_fftw, tw));
        }
    }
    for (auto &c : P->fft) {
        if (c.own) continue;
        size_t len = (size_t)c.nf;
        rocfft_plan_description desc = nullptr;
        rocfft_plan_description_create(&desc);
        rocfft_plan_description_set_data_layout(desc, rocfft_array_type_complex_interleaved,
                                                rocfft_array_type_complex_interleaved, nullptr,
                                                nullptr, 0, nullptr, (size_t)c.nf, 0, nullptr,
                                                (size_t)c.nf);
        rocfft_status s = rocfft_plan_create(&c.plan, rocfft_placement_notinplace,
                                             rocfft_transform_type_complex_inverse,
                                             mixed ? rocfft_precision_single
                                                   : rocfft_precision_double,
                                             1, &len, (size_t)c.batch, desc);
        rocfft_plan_description_destroy(desc);
        if (s != rocfft_status_success) {
            nufht_set_error("rocfft_plan_create(len=%d, batch=%lld) failed: %d", c.nf, c.batch, (int)s);
            return -520;
        }
        rocfft_execution_info_create(&c.info);
        rocfft_plan_get_work_buffer_size(c.plan, &c.work_bytes);
        if (c.work_bytes) {
            HIP_TRY(hipMalloc(&c.work, c.work_bytes));
            rocfft_execution_info_set_work_buffer(c.info, c.work, c.work_bytes);
        }
    }
    clk.mark("FFT tables + rocFFT plans");
    // the second stream of the near-field stages (default priority)
    if (!P->lproj.empty() || !P->dir.empty()) {
        HIP_TRY(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// program execution
// The asymptotic stages in two parts: the fused small boxes, grid zeroing and
// spreading (part 0), then the FFTs and interpolation (part 1).
template <typename R, int EM>
static void launch_asy(Program *P, hipStream_t st, int part)
{
    if (part == 0) {
        asy_fused<R, EM>(P, st);
        asy_grid_zero<R>(P, st);
        asy_spread<R, EM>(P, st);
    } else {
        asy_fft<R>(P, st);
        asy_interp<R, EM>(P, st);
    }
}

template <typename R>
static void launch_asy_e(Program *P, hipStream_t st, int part)
{
    switch (P->emax) {
    case 6: launch_asy<R, 6>(P, st, part); break;
    case 10: launch_asy<R, 10>(P, st, part); break;
    case 14: launch_asy<R, 14>(P, st, part); break;
    default: launch_asy<R, 22>(P, st, part); break;
    }
}

// Local-expansion contractions in FP32 in mixed precision (the Bessel
// products b_l and the basis recurrence stay FP64), FP64 otherwise.
template <int NLT> static void launch_loc_proj_n(Program *P, hipStream_t st, double *coef)
{
    if (P->prec == 2)
        hipLaunchKernelGGL((k_loc_proj<NLT, float>), dim3((unsigned)P->lproj.size()), dim3(256),
                           LocProjLds<NLT>::bytes, st, P->d_tfs, P->d_lbox, P->d_lproj, coef);
    else
        hipLaunchKernelGGL((k_loc_proj<NLT, double>), dim3((unsigned)P->lproj.size()), dim3(256),
                           LocProjLds<NLT>::bytes, st, P->d_tfs, P->d_lbox, P->d_lproj, coef);
}

static void launch_loc_proj(Program *P, hipStream_t st, double *coef)
{
    switch (P->loc_nlt) {
    case 1: launch_loc_proj_n<1>(P, st, coef); break;
    case 2: launch_loc_proj_n<2>(P, st, coef); break;
    case 3: launch_loc_proj_n<3>(P, st, coef); break;
    case 4: launch_loc_proj_n<4>(P, st, coef); break;
    case 5: launch_loc_proj_n<5>(P, st, coef); break;
    default: launch_loc_proj_n<6>(P, st, coef); break;
    }
}

// coef: the local-expansion moments' buffer (coef, or the caller's
// scratch of at least P->ncoef doubles); it is cleared before use.
// Where the near-field stages fork off: 0 never (one stream), 1 before the
// spread (measured best: clustered step 12.04 -> 11.71 ms mixed, 15.73 ->
// 15.37 FP64, headline 12.03 -> 11.93; forking before the FFTs gained about
// half of that), 2 before the FFTs.  NUFHT_AMD_FORK overrides.
static int g_fork_at = [] {
    const char *e = getenv("NUFHT_AMD_FORK");
    return e ? atoi(e) : 1;
}();

extern "C" int nufht_amd_stream_fork(int at)
{
    const int prev = g_fork_at;
    if (at >= 0 && at <= 2) g_fork_at = at;
    return prev;
}

// Near-field stages: local projection and evaluation, then the direct sums.
static int run_near_field(Program *P, hipStream_t st, double *coef)
{
    if (!P->lproj.empty()) {
        {
            StageTimer tm(ST_LOCP, st);
            launch_loc_proj(P, st, coef);
            if (!P->lfold.empty())
                hipLaunchKernelGGL(k_loc_fold, dim3((unsigned)P->lfold.size()), dim3(256), 0, st,
                                   P->d_tfs, P->d_lbox, P->d_lfold, coef, P->d_cheb);
        }
        {
            StageTimer tm(ST_LOCE, st);
            if (!P->leval1.empty()) {
                if (P->prec == 2)
                    hipLaunchKernelGGL(k_loc_eval1<float>, dim3((unsigned)P->leval1.size()),
                                       dim3(P->loc_tpb), P->loc_lds_e, st, P->d_tfs, P->d_lbox,
                                       P->d_leval1, coef, P->d_cheb);
                else
                    hipLaunchKernelGGL(k_loc_eval1<double>, dim3((unsigned)P->leval1.size()),
                                       dim3(P->loc_tpb), P->loc_lds_e, st, P->d_tfs, P->d_lbox,
                                       P->d_leval1, coef, P->d_cheb);
            }
            if (!P->leval1h.empty()) {
                if (P->prec == 2)
                    hipLaunchKernelGGL(k_loc_eval1<float>, dim3((unsigned)P->leval1h.size()),
                                       dim3(256), 0, st, P->d_tfs, P->d_lbox, P->d_leval1h, coef,
                                       P->d_cheb);
                else
                    hipLaunchKernelGGL(k_loc_eval1<double>, dim3((unsigned)P->leval1h.size()),
                                       dim3(256), 0, st, P->d_tfs, P->d_lbox, P->d_leval1h, coef,
                                       P->d_cheb);
            }
            if (!P->leval.empty()) {
                const bool f32 = P->prec == 2;
                const void *fn = P->loc_bmax == 32
                                     ? (f32 ? (const void *)k_loc_eval<32, float> : (const void *)k_loc_eval<32, double>)
                                     : (f32 ? (const void *)k_loc_eval<16, float> : (const void *)k_loc_eval<16, double>);
                void *args[] = {(void *)&P->d_tfs, (void *)&P->d_lbox, (void *)&P->d_leval,
                                (void *)&P->d_lev_boxes, (void *)&coef, (void *)&P->d_cheb};
                HIP_TRY(hipLaunchKernel(fn, dim3((unsigned)P->leval.size()), dim3(P->loc_tpb), args,
                                        P->loc_lds_e, st));
            }
        }
    }
    if (!P->dir.empty()) {
        StageTimer tm(ST_DIR, st);
        // items grouped by column class (4, 8, 16), one launch per class
        const int *o = P->dir_off;
        if (o[1] > o[0])
            hipLaunchKernelGGL(k_dir<4>, dim3((unsigned)(o[1] - o[0])), dim3(256), 0, st, P->d_tfs,
                               P->d_dir + o[0], P->d_dir_rng, P->d_jtab);
        if (o[2] > o[1])
            hipLaunchKernelGGL(k_dir<8>, dim3((unsigned)(o[2] - o[1])), dim3(256), 0, st, P->d_tfs,
                               P->d_dir + o[1], P->d_dir_rng, P->d_jtab);
        if (o[3] > o[2])
            hipLaunchKernelGGL(k_dir<16>, dim3((unsigned)(o[3] - o[2])), dim3(256), 0, st, P->d_tfs,
                               P->d_dir + o[2], P->d_dir_rng, P->d_jtab);
    }
    return 0;
}

// coef: the local-expansion moments' buffer (P->d_coef, or the caller's
// scratch of at least P->ncoef doubles); it is cleared before use.
//
// The near-field stages and the asymptotic stages only meet in their
// additions into g (atomic, order-free), so when a program has both, the
// near-field stages run on the program's second stream, forked from `st`
// after g and the moments are zeroed and joined back before the call ends:
// they fill the cycles the latency-bound spread, FFT and interpolation
// kernels leave idle.  Inside a graph capture the fork and join are captured
// as dependencies like any others.
static int run_program(Program *P, hipStream_t st, double *coef)
{
    if (P->desc_dirty) {
        HIP_TRY(hipMemcpyAsync(P->d_tfs, P->tfs.data(), P->tfs.size() * sizeof(TfDesc),
                               hipMemcpyHostToDevice, st));
        P->desc_dirty = false;
    }
    if (!P->zero.empty()) {
        StageTimer tm(ST_ZERO, st);
        hipLaunchKernelGGL(k_zero, dim3((unsigned)P->zero.size()), dim3(256), 0, st, P->d_tfs,
                           P->d_zero);
        if (P->ncoef)
            hipLaunchKernelGGL(k_fill0, dim3((unsigned)std::min<long long>(1024, (P->ncoef + 255) / 256)),
                               dim3(256), 0, st, coef, P->ncoef);
    }
    const bool asy = !P->aspread.empty() || !P->adense.empty() || !P->funits.empty();
    const bool near = !P->lproj.empty() || !P->dir.empty();
    const bool fork = asy && near && P->side && g_fork_at > 0;
    if (!fork) RC_TRY(run_near_field(P, st, coef));
    for (int part = 0; part < 2; ++part) {
        if (fork && part == g_fork_at - 1) {
            HIP_TRY(hipEventRecord(P->ev_fork, st));
            HIP_TRY(hipStreamWaitEvent(P->side, P->ev_fork, 0));
            RC_TRY(run_near_field(P, P->side, coef));
            HIP_TRY(hipEventRecord(P->ev_join, P->side));
        }
        if (asy) {
            if (P->prec == 2) launch_asy_e<float>(P, st, part);
            else launch_asy_e<double>(P, st, part);
        }
    }
    if (fork) HIP_TRY(hipStreamWaitEvent(st, P->ev_join, 0));
    HIP_TRY(hipGetLastError());
    return 0;
}

static void set_io(Program *P, int k, const double *cs, long long ld_cs, double *gs, long long ld_gs)
{
    TfDesc &T = P->tfs[k];
    if (T.cs != cs || T.gs != gs || T.ld_cs != ld_cs || T.ld_gs != ld_gs) {
        T.cs = cs;
        T.gs = gs;
        T.ld_cs = ld_cs;
        T.ld_gs = ld_gs;
        P->desc_dirty = true;
    }
}

static int get_program(const NufhtPlan *plan, EngineState *E, int B, Program **out)
{
    RC_TRY(ensure_dev(plan, E));
    int prec = effective_prec(plan);
    if (!E->prog || E->prog_B != B || E->prog_prec != prec || E->prog_knobs != E->knobs) {
        free_program(E->prog);
        E->prog = nullptr;
        Program *P = new Program();
        P->knobs = E->knobs;
        std::vector<TfInput> in{{plan, B}};
        int rc = build_program(P, in);
        if (rc) { free_program(P); return rc; }
        E->prog = P;
        E->prog_B = B;
        E->prog_prec = prec;
        E->prog_knobs = P->knobs;
    }
    *out = E->prog;
    return 0;
}

// Sorted-order execution on device buffers (rows in ascending grid order).
static int exec_sorted(const NufhtPlan *plan, const double *d_cs, long long ld_cs, double *d_gs,
                       long long ld_gs, int B, hipStream_t st)
{
    EngineState *E = engine_of(plan);
    Program *P;
    RC_TRY(get_program(plan, E, B, &P));
    RC_TRY(order_after_last(E, st));
    set_io(P, 0, d_cs, ld_cs, d_gs, ld_gs);
    RC_TRY(run_program(P, st, P->d_coef));
    return mark_done(E, st);
}

static int gather_rows(const double *src, long long ld_src, double *dst, long long ld_dst,
                       const int *perm, int rows, int cols, hipStream_t st)
{
    long long tot = (long long)rows * cols;
    if (tot == 0) return 0;
    StageTimer tm(ST_PERM, st);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, src,
                       ld_src, dst, ld_dst, perm, rows, cols);
    return 0;
}

static int scatter_rows(const double *src, long long ld_src, double *dst, long long ld_dst,
                        const int *perm, int rows, int cols, hipStream_t st)
{
    long long tot = (long long)rows * cols;
    if (tot == 0) return 0;
    StageTimer tm(ST_PERM, st);
    hipLaunchKernelGGL(k_scatter_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, src,
                       ld_src, dst, ld_dst, perm, rows, cols);
    return 0;
}

// Device-pointer execution in the caller's row order.
static int exec_device(const NufhtPlan *plan, const double *d_cs, long long ld_cs, double *d_gs,
                       long long ld_gs, int B, hipStream_t st)
{
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    EngineState *E = engine_of(plan);
    RC_TRY(ensure_dev(plan, E));
    const double *cs_use = d_cs;
    long long ld_use = ld_cs;
    if (h->perm_r) {
        RC_TRY(ensure_buf(&E->d_in, &E->in_cap, (size_t)plan->m * B));
        RC_TRY(gather_rows(d_cs, ld_cs, E->d_in, plan->m, E->dev.d_perm_r, plan->m, B, st));
        cs_use = E->d_in;
        ld_use = plan->m;
    }
    if (h->perm_w) {
        RC_TRY(ensure_buf(&E->d_out, &E->out_cap, (size_t)plan->n * B));
        RC_TRY(exec_sorted(plan, cs_use, ld_use, E->d_out, plan->n, B, st));
        RC_TRY(scatter_rows(E->d_out, plan->n, d_gs, ld_gs, E->dev.d_perm_w, plan->n, B, st));
        return 0;
    }
    return exec_sorted(plan, cs_use, ld_use, d_gs, ld_gs, B, st);
}

extern "C" int nufht_batch_device(const NufhtPlan *plan, const double *d_cs, int ld_cs, double *d_gs,
                                  int ld_gs, int batch, void *stream)
{
    if (!plan || !plan->impl || !d_cs || !d_gs || batch <= 0) return -400;
    if (ld_cs < plan->m || ld_gs < plan->n) return -401;
    EngineState *E = engine_of(plan);
    std::lock_guard<std::mutex> g(E->mu);
    RC_TRY(order_after_last(E, (hipStream_t)stream));
    int rc = exec_device(plan, d_cs, ld_cs, d_gs, ld_gs, batch, (hipStream_t)stream);
    if (!rc) rc = mark_done(E, (hipStream_t)stream);
    return rc;
}

// ---------------------------------------------------------------------------
// Allocation-free execution (nufht_batch_device_noalloc).  The kernels'
// scratch, i.e. the local-expansion moments (partial sums, cleared by a fill
// kernel at the start of every call) and the sorted-order staging of
// callers whose grids are not ascending, lives in the static device array of
// scratch.hip instead of buffers the library allocates.
extern "C" int nufht_batch_device_noalloc(const NufhtPlan *plan, const double *d_cs, int ld_cs,
                                          double *d_gs, int ld_gs, int batch, void *stream)
{
    if (!plan || !plan->impl || !d_cs || !d_gs || batch <= 0) return -400;
    if (ld_cs < plan->m || ld_gs < plan->n) return -401;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    const int m = plan->m, n = plan->n, B = batch;
    const long long n_in = h->perm_r ? (long long)m * B : 0, n_out = h->perm_w ? (long long)n * B : 0;
    double *base;
    long long cap;
    RC_TRY(nufht_scratch_array(&base, &cap));
    if (n_in + n_out > cap) {
        nufht_set_error("nufht_batch_device_noalloc: staging of %lld doubles exceeds the scratch "
                        "array's %lld (m, n = %d, %d unsorted nodes, batch %d)",
                        n_in + n_out, cap, m, n, B);
        return -550;
    }
    EngineState *E = (EngineState *)h->dev;
    if (!E) {
        nufht_set_error("nufht_batch_device_noalloc: the plan has no device state yet; run "
                        "nufht_batch_device once with batch %d first", B);
        return -551;
    }
    std::lock_guard<std::mutex> g(E->mu);
    if (!E->dev_ready || !E->prog || E->prog_B != B || E->prog_prec != effective_prec(plan) ||
        E->prog_knobs != E->knobs) {
        nufht_set_error("nufht_batch_device_noalloc: the plan has no launch program for batch %d "
                        "with its current precision and knobs; run nufht_batch_device once with "
                        "this batch first", B);
        return -551;
    }
    Program *P = E->prog;
    if (P->ncoef + n_in + n_out > cap) {
        nufht_set_error("nufht_batch_device_noalloc: %lld local-expansion moments and %lld "
                        "staged values exceed the scratch array's %lld doubles",
                        P->ncoef, n_in + n_out, cap);
        return -550;
    }
    double *coef = base, *s_in = base + P->ncoef, *s_out = s_in + n_in;
    hipStream_t st = (hipStream_t)stream;
    RC_TRY(order_after_last(E, st));
    const double *cs_use = d_cs;
    long long ld_in = ld_cs;
    if (h->perm_r) {
        RC_TRY(gather_rows(d_cs, ld_cs, s_in, m, E->dev.d_perm_r, m, B, st));
        cs_use = s_in;
        ld_in = m;
    }
    set_io(P, 0, cs_use, ld_in, h->perm_w ? s_out : d_gs, h->perm_w ? n : ld_gs);
    RC_TRY(run_program(P, st, coef));
    if (h->perm_w) RC_TRY(scatter_rows(s_out, n, d_gs, ld_gs, E->dev.d_perm_w, n, B, st));
    RC_TRY(mark_done(E, st));
    return 0;
}

extern "C" int nufht_engine_batch_host(const NufhtPlan *plan, const double *cs, int ld_cs, double *gs,
                                       int ld_gs, int batch, NufhtScratch *scratch)
{
    (void)scratch;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    EngineState *E = engine_of(plan);
    std::lock_guard<std::mutex> g(E->mu);
    int m = plan->m, n = plan->n;
    RC_TRY(ensure_dev(plan, E));
    RC_TRY(ensure_buf(&E->d_in, &E->in_cap, (size_t)m * batch));
    RC_TRY(ensure_buf(&E->d_out, &E->out_cap, (size_t)n * batch));
    hipStream_t st = nullptr;
    RC_TRY(order_after_last(E, st));
    if (h->perm_r) {
        std::vector<double> tmp((size_t)m * batch);
        for (int t = 0; t < batch; ++t)
            for (int j = 0; j < m; ++j) tmp[(size_t)t * m + j] = cs[(size_t)t * ld_cs + h->perm_r[j]];
        HIP_TRY(hipMemcpy(E->d_in, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy2D(E->d_in, (size_t)m * sizeof(double), cs, (size_t)ld_cs * sizeof(double),
                            (size_t)m * sizeof(double), batch, hipMemcpyHostToDevice));
    }
    RC_TRY(exec_sorted(plan, E->d_in, m, E->d_out, n, batch, st));
    if (h->perm_w) {
        std::vector<double> tmp((size_t)n * batch);
        HIP_TRY(hipMemcpy(tmp.data(), E->d_out, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int t = 0; t < batch; ++t)
            for (int i = 0; i < n; ++i) gs[(size_t)t * ld_gs + h->perm_w[i]] = tmp[(size_t)t * n + i];
    } else {
        HIP_TRY(hipMemcpy2D(gs, (size_t)ld_gs * sizeof(double), E->d_out, (size_t)n * sizeof(double),
                            (size_t)n * sizeof(double), batch, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

// ---------------------------------------------------------------------------
// derived plans: adjoint and order+1 transforms (built on the sorted grids)
static int derived(const NufhtPlan *plan, int which, NufhtPlan **out)
{
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    NufhtPlan **slot = which == 0 ? &h->derived.adjoint
                     : which == 1 ? &h->derived.adjoint_up : &h->derived.forward_up;
    if (!*slot) {
        NufhtPlan *d = (NufhtPlan *)calloc(1, sizeof(NufhtPlan));
        if (!d) return -501;
        double nu = which == 0 ? plan->nu : std::fabs(plan->nu) + 1.0;
        int rc = (which == 2) ? nufht_plan_init(d, nu, h->r, plan->m, h->w, plan->n, &h->opt)
                              : nufht_plan_init(d, nu, h->w, plan->n, h->r, plan->m, &h->opt);
        if (rc) { free(d); return rc; }
        ((NufhtImplHost *)d->impl)->precision = h->precision;
        *slot = d;
    }
    *out = *slot;
    return 0;
}

extern "C" int nufht_adjoint_device(const NufhtPlan *plan, const double *d_ys, int ld_ys, double *d_xs,
                                    int ld_xs, int batch, void *stream)
{
    if (!plan || !plan->impl || !d_ys || !d_xs || batch <= 0) return -400;
    if (ld_ys < plan->n || ld_xs < plan->m) return -401;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    EngineState *E = engine_of(plan);
    std::lock_guard<std::mutex> g(E->mu);
    hipStream_t st = (hipStream_t)stream;
    RC_TRY(ensure_dev(plan, E));
    RC_TRY(order_after_last(E, st));
    NufhtPlan *adj;
    RC_TRY(derived(plan, 0, &adj));
    EngineState *EA = engine_of(adj);
    const double *ys = d_ys;
    long long ldy = ld_ys;
    if (h->perm_w) {
        RC_TRY(ensure_buf(&E->d_tmp[0], &E->tmp_cap[0], (size_t)plan->n * batch));
        RC_TRY(gather_rows(d_ys, ld_ys, E->d_tmp[0], plan->n, E->dev.d_perm_w, plan->n, batch, st));
        ys = E->d_tmp[0];
        ldy = plan->n;
    }
    int rc;
    {
        std::lock_guard<std::mutex> ga(EA->mu);
        if (h->perm_r) {
            RC_TRY(ensure_buf(&E->d_tmp[1], &E->tmp_cap[1], (size_t)plan->m * batch));
            rc = exec_sorted(adj, ys, ldy, E->d_tmp[1], plan->m, batch, st);
            if (!rc) rc = scatter_rows(E->d_tmp[1], plan->m, d_xs, ld_xs, E->dev.d_perm_r, plan->m, batch, st);
        } else {
            rc = exec_sorted(adj, ys, ldy, d_xs, ld_xs, batch, st);
        }
    }
    if (!rc) rc = mark_done(E, st);
    return rc;
}

extern "C" int nufht_adjoint(const NufhtPlan *plan, const double *ys, int ld_ys, double *xs, int ld_xs,
                             int batch)
{
    if (!plan || !plan->impl || !ys || !xs || batch <= 0) return -400;
    if (ld_ys < plan->n || ld_xs < plan->m) return -401;
    double *dy = nullptr, *dx = nullptr;
    HIP_TRY(hipMalloc((void **)&dy, sizeof(double) * plan->n * batch));
    HIP_TRY(hipMalloc((void **)&dx, sizeof(double) * plan->m * batch));
    int rc = 0;
    if (hipMemcpy2D(dy, plan->n * sizeof(double), ys, (size_t)ld_ys * sizeof(double),
                    plan->n * sizeof(double), batch, hipMemcpyHostToDevice) != hipSuccess) rc = -510;
    if (!rc) rc = nufht_adjoint_device(plan, dy, plan->n, dx, plan->m, batch, nullptr);
    if (!rc && hipMemcpy2D(xs, (size_t)ld_xs * sizeof(double), dx, plan->m * sizeof(double),
                           plan->m * sizeof(double), batch, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -510;
    hipFree(dy);
    hipFree(dx);
    return rc;
}

// Node gradients via J'_nu(x) = (nu/x) J_nu(x) - J_{nu+1}(x), DLMF 10.6.2.
extern "C" int nufht_grad_nodes_device(const NufhtPlan *plan, const double *d_cs, int ld_cs,
                                       const double *d_rho, int ld_rho, int batch, double *d_grad_r,
                                       double *d_grad_w, void *stream)
{
    if (!plan || !plan->impl || !d_cs || !d_rho || batch <= 0) return -400;
    if (ld_cs < plan->m || ld_rho < plan->n) return -401;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    double anu = std::fabs(plan->nu);
    if ((std::fabs(anu - std::round(anu)) < 1e-12 && anu + 1.0 > 200.0) ||
        (std::fabs(anu - std::round(anu)) >= 1e-12 && anu + 1.0 > 9.5)) {
        nufht_set_error("node gradient needs order |nu|+1 within the supported range");
        return -2;
    }
    EngineState *E = engine_of(plan);
    std::lock_guard<std::mutex> g(E->mu);
    hipStream_t st = (hipStream_t)stream;
    int m = plan->m, n = plan->n, B = batch;
    RC_TRY(ensure_dev(plan, E));
    RC_TRY(order_after_last(E, st));
    double sgn = (plan->nu < 0 && std::fabs(plan->nu - std::round(plan->nu)) < 1e-12 &&
                  (std::llround(plan->nu) & 1)) ? -1.0 : 1.0;
    // sorted copies of cs (m x B) and rho (n x B)
    RC_TRY(ensure_buf(&E->d_tmp[0], &E->tmp_cap[0], (size_t)m * B));
    RC_TRY(ensure_buf(&E->d_tmp[1], &E->tmp_cap[1], (size_t)n * B));
    double *cs_s = E->d_tmp[0], *rho_s = E->d_tmp[1];
    RC_TRY(gather_rows(d_cs, ld_cs, cs_s, m, E->dev.d_perm_r, m, B, st));
    RC_TRY(gather_rows(d_rho, ld_rho, rho_s, n, E->dev.d_perm_w, n, B, st));
    size_t big = (size_t)std::max(m, n) * B;
    RC_TRY(ensure_buf(&E->d_tmp[2], &E->tmp_cap[2], big));
    RC_TRY(ensure_buf(&E->d_tmp[3], &E->tmp_cap[3], big));
    RC_TRY(ensure_buf(&E->d_tmp[4], &E->tmp_cap[4], big));
    RC_TRY(ensure_buf(&E->d_tmp[5], &E->tmp_cap[5], (size_t)std::max(m, n) + B));
    double *P1 = E->d_tmp[2], *Q1 = E->d_tmp[3], *W = E->d_tmp[4], *aux = E->d_tmp[5];
    double *lim = aux + std::max(m, n);
    if (d_grad_r) {
        NufhtPlan *adj, *adj_up;
        RC_TRY(derived(plan, 0, &adj));
        RC_TRY(derived(plan, 1, &adj_up));
        RC_TRY(exec_sorted(adj, rho_s, n, P1, m, B, st));                 // s * H^T rho
        long long tot = (long long)n * B;
        hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, rho_s,
                           (long long)n, W, (long long)n, E->dev.d_w, n, B);
        RC_TRY(exec_sorted(adj_up, W, n, Q1, m, B, st));                  // H^T_{nu+1} (w rho)
        const double *lim0 = nullptr;
        if (h->r[0] == 0.0 && anu == 1.0) {
            hipLaunchKernelGGL(k_colsum_weighted, dim3(B), dim3(256), 0, st, rho_s, (long long)n,
                               E->dev.d_w, n, B, lim);
            lim0 = lim;
        }
        {
            StageTimer tm(ST_GRAD, st);
            hipLaunchKernelGGL(k_grad_combine, dim3((m + 255) / 256), dim3(256), 0, st, cs_s,
                               (long long)m, P1, (long long)m, Q1, (long long)m, E->dev.d_r, lim0, anu,
                               sgn, m, B, aux);
        }
        RC_TRY(scatter_rows(aux, m, d_grad_r, m, E->dev.d_perm_r, m, 1, st));
    }
    if (d_grad_w) {
        NufhtPlan *fwd_up;
        RC_TRY(derived(plan, 2, &fwd_up));
        RC_TRY(exec_sorted(plan, cs_s, m, P1, n, B, st));                 // s * H c
        long long tot = (long long)m * B;
        hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, cs_s,
                           (long long)m, W, (long long)m, E->dev.d_r, m, B);
        RC_TRY(exec_sorted(fwd_up, W, m, Q1, n, B, st));                  // H_{nu+1} (r c)
        const double *lim0 = nullptr;
        if (h->w[0] == 0.0 && anu == 1.0) {
            hipLaunchKernelGGL(k_colsum_weighted, dim3(B), dim3(256), 0, st, cs_s, (long long)m,
                               E->dev.d_r, m, B, lim);
            lim0 = lim;
        }
        {
            StageTimer tm(ST_GRAD, st);
            hipLaunchKernelGGL(k_grad_combine, dim3((n + 255) / 256), dim3(256), 0, st, rho_s,
                               (long long)n, P1, (long long)n, Q1, (long long)n, E->dev.d_w, lim0, anu,
                               sgn, n, B, aux);
        }
        RC_TRY(scatter_rows(aux, n, d_grad_w, n, E->dev.d_perm_w, n, 1, st));
    }
    HIP_TRY(hipGetLastError());
    RC_TRY(mark_done(E, st));
    return 0;
}

// ---------------------------------------------------------------------------
// precision, statistics
extern "C" int nufht_amd_plan_set_precision(NufhtPlan *plan, int mode)
{
    if (!plan || !plan->impl || mode < 0 || mode > 2) return -400;
    ((NufhtImplHost *)plan->impl)->precision = mode;
    return 0;
}

extern "C" int nufht_amd_plan_precision(const NufhtPlan *plan)
{
    if (!plan || !plan->impl) return -400;
    return effective_prec(plan);
}

extern "C" int nufht_amd_plan_stats(const NufhtPlan *plan, int batch, NufhtPlanStats *st)
{
    if (!plan || !plan->impl || !st || batch <= 0) return -400;
    EngineState *E = engine_of(plan);
    std::lock_guard<std::mutex> g(E->mu);
    Program *P;
    RC_TRY(get_program(plan, E, batch, &P));
    *st = P->stats;
    return 0;
}

// ---------------------------------------------------------------------------
// uneven multi-transform batches
struct NufhtMulti {
    std::vector<const NufhtPlan *> plans;
    std::vector<int> B;
    Knobs knobs;
    Program *P = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<const void *> gkey;
    // sorted-order staging of transforms whose caller grids are not
    // ascending: coefficients gathered in, outputs scattered back
    std::vector<double *> st_in, st_out;
};

static void multi_destroy(NufhtMulti *mt)
{
    if (mt->gexec) hipGraphExecDestroy(mt->gexec);
    if (mt->graph) hipGraphDestroy(mt->graph);
    free_program(mt->P);
    for (double *p : mt->st_in) hipFree(p);
    for (double *p : mt->st_out) hipFree(p);
    delete mt;
}

// One pass of the batch on `st`: row gathers of unsorted transforms, the
// program, row scatters back to the callers' order.
static int multi_run(NufhtMulti *mt, const double *const *d_cs, const int *ld_cs,
                     double *const *d_gs, const int *ld_gs, hipStream_t st)
{
    for (size_t k = 0; k < mt->plans.size(); ++k)
        if (mt->st_in[k]) {
            const NufhtPlan *p = mt->plans[k];
            RC_TRY(gather_rows(d_cs[k], ld_cs[k], mt->st_in[k], p->m,
                               engine_of(p)->dev.d_perm_r, p->m, mt->B[k], st));
        }
    RC_TRY(run_program(mt->P, st, mt->P->d_coef));
    for (size_t k = 0; k < mt->plans.size(); ++k)
        if (mt->st_out[k]) {
            const NufhtPlan *p = mt->plans[k];
            RC_TRY(scatter_rows(mt->st_out[k], p->n, d_gs[k], ld_gs[k],
                                engine_of(p)->dev.d_perm_w, p->n, mt->B[k], st));
        }
    return 0;
}

extern "C" int nufht_multi_create(NufhtMulti **out, int count, const NufhtPlan *const *plans,
                                  const int *batch)
{
    if (!out || count <= 0 || !plans || !batch) return -400;
    NufhtMulti *mt = new NufhtMulti();
    std::vector<TfInput> in;
    for (int k = 0; k < count; ++k) {
        if (!plans[k] || !plans[k]->impl || batch[k] <= 0) { multi_destroy(mt); return -400; }
        EngineState *E = engine_of(plans[k]);
        int rc = ensure_dev(plans[k], E);
        if (rc) { multi_destroy(mt); return rc; }
        mt->plans.push_back(plans[k]);
        mt->B.push_back(batch[k]);
        in.push_back({plans[k], batch[k]});
        const NufhtImplHost *h = (const NufhtImplHost *)plans[k]->impl;
        double *si = nullptr, *so = nullptr;
        if (h->perm_r &&
            hipMalloc((void **)&si, sizeof(double) * plans[k]->m * batch[k]) != hipSuccess)
            si = nullptr, rc = -510;
        if (!rc && h->perm_w &&
            hipMalloc((void **)&so, sizeof(double) * plans[k]->n * batch[k]) != hipSuccess)
            so = nullptr, rc = -510;
        mt->st_in.push_back(si);
        mt->st_out.push_back(so);
        if (rc) {
            nufht_set_error("multi-transform staging: out of device memory");
            multi_destroy(mt);
            return rc;
        }
    }
    mt->P = new Program();
    mt->P->knobs = mt->knobs;
    int rc = build_program(mt->P, in);
    if (rc) { multi_destroy(mt); return rc; }
    *out = mt;
    return 0;
}

// Rebuilds the multi-transform program with mt->knobs (dropping a captured
// graph; the next graph launch captures again).
static int multi_rebuild(NufhtMulti *mt)
{
    std::vector<TfInput> in;
    for (size_t k = 0; k < mt->plans.size(); ++k) in.push_back({mt->plans[k], mt->B[k]});
    Program *P = new Program();
    P->knobs = mt->knobs;
    int rc = build_program(P, in);
    if (rc) { free_program(P); return rc; }
    if (mt->gexec) { hipGraphExecDestroy(mt->gexec); mt->gexec = nullptr; }
    if (mt->graph) { hipGraphDestroy(mt->graph); mt->graph = nullptr; }
    mt->gkey.clear();
    free_program(mt->P);
    mt->P = P;
    return 0;
}

extern "C" int nufht_multi_execute(NufhtMulti *mt, const double *const *d_cs, const int *ld_cs,
                                   double *const *d_gs, const int *ld_gs, void *stream, int use_graph)
{
    if (!mt || !d_cs || !d_gs || !ld_cs || !ld_gs) return -400;
    hipStream_t st = (hipStream_t)stream;
    std::vector<const void *> key;
    for (size_t k = 0; k < mt->plans.size(); ++k) {
        const NufhtPlan *p = mt->plans[k];
        if (!d_cs[k] || !d_gs[k]) return -400;
        if (ld_cs[k] < p->m || ld_gs[k] < p->n) return -401;
        // unsorted transforms read and write the staging buffers
        set_io(mt->P, (int)k, mt->st_in[k] ? mt->st_in[k] : d_cs[k],
               mt->st_in[k] ? p->m : ld_cs[k], mt->st_out[k] ? mt->st_out[k] : d_gs[k],
               mt->st_out[k] ? p->n : ld_gs[k]);
        key.push_back(d_cs[k]);
        key.push_back(d_gs[k]);
        key.push_back((const void *)(intptr_t)ld_cs[k]);
        key.push_back((const void *)(intptr_t)ld_gs[k]);
    }
    if (!use_graph || g_prof_on) {
        int rc = multi_run(mt, d_cs, ld_cs, d_gs, ld_gs, st);
        return rc;
    }
    if (mt->P->desc_dirty) {
        HIP_TRY(hipMemcpyAsync(mt->P->d_tfs, mt->P->tfs.data(), mt->P->tfs.size() * sizeof(TfDesc),
                               hipMemcpyHostToDevice, st));
        mt->P->desc_dirty = false;
    }
    if (!mt->gexec || key != mt->gkey) {
        if (mt->gexec) { hipGraphExecDestroy(mt->gexec); mt->gexec = nullptr; }
        if (mt->graph) { hipGraphDestroy(mt->graph); mt->graph = nullptr; }
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = multi_run(mt, d_cs, ld_cs, d_gs, ld_gs, cap);
        hipGraph_t gr = nullptr;
        hipError_t ce = hipStreamEndCapture(cap, &gr);
        hipStreamDestroy(cap);
        if (rc) return rc;
        if (ce != hipSuccess) {
            nufht_set_error("graph capture failed: %s", hipGetErrorString(ce));
            return -530;
        }
        HIP_TRY(hipGraphInstantiate(&mt->gexec, gr, nullptr, nullptr, 0));
        mt->graph = gr;
        mt->gkey = key;
    }
    HIP_TRY(hipGraphLaunch(mt->gexec, st));
    return 0;
}

extern "C" int nufht_multi_stats(const NufhtMulti *mt, NufhtPlanStats *st)
{
    if (!mt || !mt->P || !st) return -400;
    *st = mt->P->stats;
    return 0;
}

extern "C" void nufht_multi_free(NufhtMulti *mt)
{
    if (mt) multi_destroy(mt);
}

// ---------------------------------------------------------------------------
// knobs and the autotuner
static void knobs_to_c(const Knobs &k, int prec, NufhtAmdKnobs *o)
{
    o->dense_cell = k.dense_cell;
    o->dense_split = k.dense_split;
    o->spread_split = k.spread_split;
    o->interp_mfma = k.interp_mfma;
    o->own_fft = k.own_fft;
    o->spread_sched = k.spread_sched;
    o->fused_max = k.fused_max;
    o->precision = prec;
}

extern "C" int nufht_amd_get_knobs(const NufhtPlan *plan, NufhtAmdKnobs *k)
{
    if (!plan || !plan->impl || !k) return -400;
    knobs_to_c(engine_of(plan)->knobs, ((NufhtImplHost *)plan->impl)->precision, k);
    return 0;
}

static int knobs_check(const NufhtAmdKnobs *k)
{
    if (!(k->dense_cell > 0.0) || (k->dense_split != 0 && k->dense_split < kSQ) ||
        k->spread_split < 64 ||
        k->interp_mfma < 0 || k->interp_mfma > 2 || (k->own_fft != 0 && k->own_fft != 1) ||
        k->spread_sched < -1 || k->spread_sched > 1 || k->fused_max < 0 ||
        (k->precision != NUFHT_PREC_FP64 && k->precision != NUFHT_PREC_AUTO &&
         k->precision != NUFHT_PREC_MIXED)) {
        nufht_set_error("invalid knobs");
        return -400;
    }
    return 0;
}

extern "C" int nufht_amd_set_knobs(NufhtPlan *plan, const NufhtAmdKnobs *k)
{
    if (!plan || !plan->impl || !k) return -400;
    RC_TRY(knobs_check(k));
    EngineState *E = engine_of(plan);
    E->knobs.dense_cell = k->dense_cell;
    E->knobs.dense_split = k->dense_split;
    E->knobs.spread_split = k->spread_split;
    E->knobs.interp_mfma = k->interp_mfma;
    E->knobs.own_fft = k->own_fft;
    E->knobs.spread_sched = k->spread_sched;
    E->knobs.fused_max = k->fused_max;
    ((NufhtImplHost *)plan->impl)->precision = k->precision;
    return 0;
}

extern "C" int nufht_multi_get_knobs(const NufhtMulti *mt, NufhtAmdKnobs *k)
{
    if (!mt || !k) return -400;
    knobs_to_c(mt->knobs, NUFHT_PREC_AUTO, k);
    return 0;
}

// Candidate values per knob, tried in this order by coordinate descent.
static const double kTuneDense[] = {2.0, 4.0, 8.0, 16.0, 32.0, 1e300};
static const int kTuneDenseSplit[] = {512, 1024, 2048, 4096};
static const int kTuneSpreadSplit[] = {1024, 2048, 4096};
static const int kTuneFusedMax[] = {0, 128, 512};

// Times `run` (which enqueues one call on st): one untimed call (builds the
// program), then the mean of three.  Returns a negative code on failure.
template <class Run> static int tune_time(Run run, hipStream_t st, float *ms)
{
    RC_TRY(run());
    HIP_TRY(hipStreamSynchronize(st));
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, st));
    int rc = 0;
    for (int r = 0; r < 3 && !rc; ++r) rc = run();
    HIP_TRY(hipEventRecord(b, st));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    hipEventElapsedTime(&t, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = t / 3.0f;
    return rc;
}

// Coordinate descent over the knobs of `cur` with `apply` (sets knobs, which
// the next call builds) and `run`; a candidate replaces the incumbent when it
// is at least 2 % faster.
template <class Apply, class Run>
static int tune_knobs(Knobs &cur, Apply apply, Run run, hipStream_t st)
{
    float best;
    apply(cur);
    RC_TRY(tune_time(run, st, &best));
    auto consider = [&](Knobs cand) -> int {
        if (cand == cur) return 0;
        apply(cand);
        float ms;
        int rc = tune_time(run, st, &ms);
        if (rc == 0 && ms < 0.98f * best) {
            best = ms;
            cur = cand;
        }
        apply(cur);
        return 0;
    };
    for (double v : kTuneDense) { Knobs c = cur; c.dense_cell = v; consider(c); }
    for (int v : {1, 2, 0}) { Knobs c = cur; c.interp_mfma = v; consider(c); }
    for (int v : kTuneDenseSplit) { Knobs c = cur; c.dense_split = v; consider(c); }
    for (int v : kTuneSpreadSplit) { Knobs c = cur; c.spread_split = v; consider(c); }
    for (int v = 1; v >= 0; --v) { Knobs c = cur; c.own_fft = v; consider(c); }
    for (int v = -1; v <= 1; ++v) { Knobs c = cur; c.spread_sched = v; consider(c); }
    for (int v : kTuneFusedMax) { Knobs c = cur; c.fused_max = v; consider(c); }
    apply(cur);
    return 0;
}

extern "C" int nufht_amd_autotune(NufhtPlan *plan, int batch, void *stream)
{
    if (!plan || !plan->impl || batch <= 0) return -400;
    NufhtImplHost *h = (NufhtImplHost *)plan->impl;
    EngineState *E = engine_of(plan);
    hipStream_t st = (hipStream_t)stream;
    double *dc = nullptr, *dg = nullptr;
    HIP_TRY(hipMalloc((void **)&dc, sizeof(double) * plan->m * batch));
    HIP_TRY(hipMalloc((void **)&dg, sizeof(double) * plan->n * batch));
    HIP_TRY(hipMemset(dc, 0, sizeof(double) * plan->m * batch));
    auto run = [&]() { return nufht_batch_device(plan, dc, plan->m, dg, plan->n, batch, st); };
    int rc = 0;
    // precision first (both meet tol at tol >= 1e-6), then the geometry
    if (plan->tol >= 1e-6) {
        int modes[2] = {NUFHT_PREC_MIXED, NUFHT_PREC_FP64};
        float best = 1e30f;
        int best_mode = h->precision;
        for (int q = 0; q < 2 && !rc; ++q) {
            h->precision = modes[q];
            float ms;
            rc = tune_time(run, st, &ms);
            if (!rc && ms < best) { best = ms; best_mode = modes[q]; }
        }
        h->precision = best_mode;
    }
    if (!rc) {
        Knobs cur = E->knobs;
        rc = tune_knobs(cur, [&](const Knobs &k) { E->knobs = k; }, run, st);
    }
    hipFree(dc);
    hipFree(dg);
    return rc;
}

extern "C" int nufht_multi_autotune(NufhtMulti *mt, void *stream)
{
    if (!mt) return -400;
    hipStream_t st = (hipStream_t)stream;
    size_t n = mt->plans.size();
    std::vector<double *> dc(n, nullptr), dg(n, nullptr);
    std::vector<const double *> pc(n);
    std::vector<double *> pg(n);
    std::vector<int> lc(n), lg(n);
    int rc = 0;
    for (size_t k = 0; k < n && !rc; ++k) {
        const NufhtPlan *p = mt->plans[k];
        if (hipMalloc((void **)&dc[k], sizeof(double) * p->m * mt->B[k]) != hipSuccess ||
            hipMalloc((void **)&dg[k], sizeof(double) * p->n * mt->B[k]) != hipSuccess ||
            hipMemset(dc[k], 0, sizeof(double) * p->m * mt->B[k]) != hipSuccess) {
            nufht_set_error("autotune buffers: out of device memory");
            rc = -510;
        }
        pc[k] = dc[k];
        pg[k] = dg[k];
        lc[k] = p->m;
        lg[k] = p->n;
    }
    if (!rc) {
        auto run = [&]() {
            return nufht_multi_execute(mt, pc.data(), lc.data(), pg.data(), lg.data(), st, 0);
        };
        Knobs cur = mt->knobs;
        int brc = 0;
        rc = tune_knobs(cur,
                        [&](const Knobs &k) {
                            if (k != mt->knobs || !mt->P) {
                                mt->knobs = k;
                                int r = multi_rebuild(mt);
                                if (r && !brc) brc = r;
                            }
                        },
                        run, st);
        if (!rc) rc = brc;
    }
    for (size_t k = 0; k < n; ++k) {
        hipFree(dc[k]);
        hipFree(dg[k]);
    }
    return rc;
}
