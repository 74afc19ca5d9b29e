// rvsp_plan.hip -- the device half of a SpGEMM structure (the "plan") and
// the C ABI the host core calls (rvsp_hip_abi.h), plus the public device
// and timing helpers of rv_sparse_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rv_sparse_hip.h"
#include "rvsp_device.hpp"
#include "rvsp_hip_abi.h"
#include "rvsp_kernels.hpp"

using namespace rvsp;

namespace {

int g_fail_allocs = 0;  // test hook: number of allocations to fail

rvsp_status_t map_err(hipError_t e)
{
    switch (e) {
    case hipSuccess: return RVSP_SUCCESS;
    case hipErrorOutOfMemory: return RVSP_ERROR_ALLOCATION_FAILED;
    case hipErrorInvalidValue: return RVSP_ERROR_INVALID_ARGUMENT;
    default: return RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            if (getenv("RVSP_HIP_VERBOSE"))                                                       \
                fprintf(stderr, "rvsp: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                                      \
            return map_err(e_);                                                                   \
        }                                                                                         \
    } while (0)

bool g_pool_ready = false;

rvsp_status_t device_ready()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RVSP_ERROR_UNSUPPORTED_BACKEND;
    if (!g_pool_ready) {
        // keep freed stream-ordered allocations cached in the pool, so
        // repeated products reuse memory instead of returning it to the OS
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            // memory freed on one stream may serve allocations on another
            int on = 1;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowOpportunistic, &on);
            (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &on);
            (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseFollowEventDependencies, &on);
        }
        g_pool_ready = true;
    }
    return RVSP_SUCCESS;
}

// Blocks freed on a one-shot stream (each host thread has its own, see
// rvsp_hip_oneshot_stream), kept for later calls on that same stream.
// Stream order makes the reuse safe without events: every earlier use of a
// block was queued on the stream before its free, so idle blocks are kept
// per stream and only handed out on it. The stream-ordered pool alone does
// not keep them: at a stream synchronisation the runtime hands most freed
// blocks back to the OS whatever the release threshold, and re-mapping C's
// arrays and the big-row scratch each call costs about a millisecond per
// few GB, with occasional 50-150 ms stalls.
struct BlockInfo {
    size_t size;
    hipStream_t stream;  // the one-shot stream it was allocated on
};
struct BlockCache {
    std::mutex mu;
    std::unordered_map<hipStream_t, std::multimap<size_t, void *>> idle;  // per stream
    std::unordered_map<void *, BlockInfo> blocks;  // every block the cache handed out
    size_t idle_bytes = 0;
    size_t limit = 0;  // idle bytes kept at most (all streams): a quarter of device memory
};
BlockCache g_cache;

size_t cache_round(size_t bytes)
{
    const size_t g = bytes >= ((size_t)1 << 24) ? ((size_t)2 << 20) : (size_t)4096;
    return (bytes + g - 1) / g * g;
}

// hands a stream's idle blocks back to the pool (caller holds the lock)
void cache_flush_locked(hipStream_t st)
{
    auto it = g_cache.idle.find(st);
    if (it == g_cache.idle.end()) return;
    for (auto &e : it->second) {
        g_cache.blocks.erase(e.second);
        g_cache.idle_bytes -= e.first;
        (void)hipFreeAsync(e.second, st);
    }
    it->second.clear();
}

// whether st is the calling thread's one-shot stream
bool is_oneshot(hipStream_t st) { return st && st == (hipStream_t)rvsp_hip_oneshot_stream(); }

// The calling thread's stream for synchronous helpers (never the legacy
// null stream, whose operations fail while another thread captures).
hipStream_t thread_stream() { return (hipStream_t)rvsp_hip_oneshot_stream(); }

rvsp_status_t dalloc(void **p, size_t bytes, hipStream_t st)
{
    *p = nullptr;
    if (g_fail_allocs > 0) {
        --g_fail_allocs;
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    if (bytes == 0) bytes = 256;
    if (!is_oneshot(st)) {
        HIP_TRY(hipMallocAsync(p, bytes, st));
        return RVSP_SUCCESS;
    }
    bytes = cache_round(bytes);
    std::lock_guard<std::mutex> lk(g_cache.mu);
    if (g_cache.limit == 0) {
        size_t fr = 0, tot = 0;
        g_cache.limit = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 4 : ((size_t)1 << 32);
    }
    // best fit among this stream's idle blocks, wasting at most an eighth
    auto &idle = g_cache.idle[st];
    auto it = idle.lower_bound(bytes);
    if (it != idle.end() && it->first - bytes <= it->first / 8) {
        *p = it->second;
        g_cache.idle_bytes -= it->first;
        idle.erase(it);
        return RVSP_SUCCESS;
    }
    hipError_t e = hipMallocAsync(p, bytes, st);
    if (e == hipErrorOutOfMemory && g_cache.idle_bytes > 0) {
        (void)hipGetLastError();  // recovered below: not a launch error
        // every thread's idle blocks go back to the pool, each freed on the
        // stream that last used it, which is then drained
        for (auto &kv : g_cache.idle) {
            if (kv.second.empty()) continue;
            cache_flush_locked(kv.first);
            (void)hipStreamSynchronize(kv.first);
        }
        e = hipMallocAsync(p, bytes, st);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return map_err(e);
    }
    g_cache.blocks[*p] = BlockInfo{bytes, st};
    return RVSP_SUCCESS;
}

void dfree(void *p, hipStream_t st)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        auto it = g_cache.blocks.find(p);
        if (it != g_cache.blocks.end()) {
            if (st == it->second.stream && is_oneshot(st)) {
                auto &idle = g_cache.idle[st];
                idle.emplace(it->second.size, p);
                g_cache.idle_bytes += it->second.size;
                // over the limit: this stream's smallest idle blocks go back first
                while (g_cache.idle_bytes > g_cache.limit && !idle.empty()) {
                    auto v = idle.begin();
                    g_cache.idle_bytes -= v->first;
                    g_cache.blocks.erase(v->second);
                    (void)hipFreeAsync(v->second, st);
                    idle.erase(v);
                }
                return;
            }
            g_cache.blocks.erase(it);  // freed elsewhere: the pool takes it
        }
    }
    (void)hipFreeAsync(p, st);
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Rank-based reuse passes are on unless RVSP_REUSE_RANKS=0 (read once).
bool reuse_ranks_off()
{
    static int off = -1;
    if (off < 0) {
        const char *e = getenv("RVSP_REUSE_RANKS");
        off = e && e[0] == '0' ? 1 : 0;
    }
    return off == 1;
}

// Count-pass bitmaps of one-pass-bitmap rows are kept for the fill unless
// RVSP_BMS_SPILL=0 (read once).
bool bms_spill_off()
{
    static int off = -1;
    if (off < 0) {
        const char *e = getenv("RVSP_BMS_SPILL");
        off = e && e[0] == '0' ? 1 : 0;
    }
    return off == 1;
}

// Big-row numeric window shape: the tuned one unless RVSP_BIG_WINDOW names
// another launchable one as "<outputs>x<chunks>" (read at every analysis).
int big_window_shape(int tuned)
{
    int shape = tuned >= 0 && tuned < kBigShapeCount ? tuned : 0;
    const char *e = getenv("RVSP_BIG_WINDOW");
    int o = 0, c = 0;
    if (e && sscanf(e, "%dx%d", &o, &c) == 2)
        for (int i = 0; i < kBigShapeCount; ++i)
            if (kBigShapes[i][0] == o && kBigShapes[i][1] == c) shape = i;
    return shape;
}

// Phase timing: when on, every analyse / fill / numeric call records a pair
// of HIP events on its own stream; rvsp_hip_get_timing resolves them.
enum Phase { PH_ANALYSE = 0, PH_FILL = 1, PH_NUMERIC = 2, PH_COUNT = 4 };
struct PendingTime {
    int phase;
    hipEvent_t a, b;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<PendingTime> g_pending;
double g_ms[PH_COUNT];
long long g_calls[PH_COUNT];

struct PhaseScope {
    hipStream_t st;
    int phase;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseScope(hipStream_t s, int ph) : st(s), phase(ph)
    {
        if (!g_timing) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, st);
    }
    ~PhaseScope()
    {
        if (!a) return;
        (void)hipEventRecord(b, st);
        std::lock_guard<std::mutex> lk(g_tmu);
        g_pending.push_back({phase, a, b});
    }
};

} // namespace

struct rvsp_hip_plan {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int32_t m = 0, k = 0, n = 0, nnz_a = 0, nnz_b = 0;
    const int32_t *arp = nullptr, *aci = nullptr, *brp = nullptr, *bci = nullptr;

    void *rowblock = nullptr;  // one allocation carved into the per-row arrays
    int64_t *ops = nullptr;
    int32_t *lo = nullptr, *hi = nullptr;
    uint8_t *sbin = nullptr, *nbin = nullptr;
    uint8_t *hbin = nullptr, *fbin = nullptr;  // learned sizing: heuristic bins, recount list
    int32_t *cnt = nullptr, *crp = nullptr, *sperm = nullptr, *nperm = nullptr;
    int64_t *tsum = nullptr;
    AnalysisCounters *ctr = nullptr;
    int32_t *binoff = nullptr;  // [0,64) bin offsets, [64,128) scatter cursors
    int *err = nullptr;
    bool operands_checked = false;  // the next analysis skips the device check

    void *bigblock = nullptr;  // big-row bitmaps, chunk counts, windows, segments
    void *wordblock = nullptr; // B's (word, bits) pairs for the big-row count (analysis only)
    int4 *cursors = nullptr;
    BigArgs big{};
    int nseg = 0;

    AnalysisCounters hctr;
    int32_t sym_count[SB_COUNT], sym_off[SB_COUNT];
    int32_t num_count[NB_COUNT], num_off[NB_COUNT];
    int64_t nnzc = 0, products = 0;
    int64_t sized_rows = 0, fallback_rows = 0;  // learned sizing: rows moved, rows recounted
    // host sources of the bin-offset uploads (symbolic, recount, numeric);
    // each stays untouched until the stream has passed its copy
    int32_t h_off[3][64];
    bool analysed = false;

    int2 *bcv = nullptr;  // B's (column, value) pairs, rewritten by every numeric pass
    uint32_t *bms_spill = nullptr;  // count-pass bitmaps of the one-pass-bitmap rows
    bool bms_spill_ok = false;      // the fill's bin lists are the count's (no recount)

    // reused structure: per-product output ranks of the LDS-path rows
    void *pmblock = nullptr;  // per-row product offsets (int64), scan scratch
    uint8_t *pmap = nullptr;  // 1- or 2-byte ranks per row, at prow[row] bytes
    int64_t *prow = nullptr;
    int64_t pm_bytes = 0;
    bool pm_ready = false, pm_skip = false;

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void *gkey[4] = {nullptr, nullptr, nullptr, nullptr};

    // arena mode (the allocation-free one-shot): the analysis and the
    // passes carve their device arrays from [arena, arena + arena_cap)
    // instead of allocating, and run every row bin on the caller's stream
    char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    bool single_stream = false;

    TuneParams tp{};  // launch parameters of the current analysis (tuning table)
};

// Device arrays of a plan: carved from its arena when it has one (released
// all at once by plan_release), else from the allocator above.
static rvsp_status_t palloc(rvsp_hip_plan *p, void **out, size_t bytes)
{
    if (!p->arena) return dalloc(out, bytes, p->stream);
    *out = nullptr;
    const size_t need = align256(bytes ? bytes : 256);
    if (p->arena_used > p->arena_cap || need > p->arena_cap - p->arena_used)
        return RVSP_ERROR_SCRATCH_EXCEEDED;
    *out = p->arena + p->arena_used;
    p->arena_used += need;
    return RVSP_SUCCESS;
}

static void pfree(rvsp_hip_plan *p, void *ptr)
{
    if (!p->arena) dfree(ptr, p->stream);
}

// Whether a pass of `work` units runs its row bins on side streams (the
// tuned thresholds: tp.fan_min_products, tp.fan_min_fill).
static bool fan_out(const rvsp_hip_plan *p, int64_t work, int64_t min_work)
{
    return !p->single_stream && work >= min_work;
}

static void drop_graph(rvsp_hip_plan *p)
{
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    for (auto &k : p->gkey) k = nullptr;
}

static void plan_release(rvsp_hip_plan *p)
{
    pfree(p, p->bcv);
    p->bcv = nullptr;
    pfree(p, p->wordblock);
    p->wordblock = nullptr;
    pfree(p, p->bms_spill);
    p->bms_spill = nullptr;
    p->bms_spill_ok = false;
    dfree(p->pmblock, p->stream);
    dfree(p->pmap, p->stream);
    p->pmblock = nullptr;
    p->pmap = nullptr;
    p->prow = nullptr;
    p->pm_bytes = 0;
    p->pm_ready = p->pm_skip = false;
    pfree(p, p->rowblock);
    pfree(p, p->bigblock);
    pfree(p, p->cursors);
    p->arena_used = 0;
    p->rowblock = nullptr;
    p->err = nullptr;  // it lived in the row block
    p->bigblock = nullptr;
    p->cursors = nullptr;
    p->big = BigArgs{};
    p->nseg = 0;
    p->analysed = false;
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
    for (auto &k : p->gkey) k = nullptr;
}

extern "C" {

rvsp_status_t rvsp_hip_plan_create(rvsp_hip_plan_t **out, void *stream)
{
    if (!out) return RVSP_ERROR_NULL_POINTER;
    *out = nullptr;
    rvsp_status_t st = device_ready();
    if (st != RVSP_SUCCESS) return st;
    rvsp_hip_plan *p = new (std::nothrow) rvsp_hip_plan();
    if (!p) return RVSP_ERROR_ALLOCATION_FAILED;
    if (stream) {
        p->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
            delete p;
            return RVSP_ERROR_UNSUPPORTED_BACKEND;
        }
        p->own_stream = true;
    }
    *out = p;
    return RVSP_SUCCESS;
}

void rvsp_hip_plan_destroy(rvsp_hip_plan_t *p)
{
    if (!p) return;
    plan_release(p);
    (void)hipStreamSynchronize(p->stream);
    if (p->own_stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

void *rvsp_hip_plan_stream(const rvsp_hip_plan_t *p) { return p ? (void *)p->stream : nullptr; }

void *rvsp_hip_oneshot_stream(void)
{
    // one per host thread (kept for the thread's lifetime), so one-shots
    // and batches of different threads do not queue behind each other
    thread_local hipStream_t st = nullptr;
    if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
    return (void *)st;
}

void rvsp_hip_plan_operands_checked(rvsp_hip_plan_t *p)
{
    if (p) p->operands_checked = true;
}

const int32_t *rvsp_hip_plan_row_ptr(const rvsp_hip_plan_t *p) { return p ? p->crp : nullptr; }

const int64_t *rvsp_hip_plan_op_counts(const rvsp_hip_plan_t *p) { return p ? p->ops : nullptr; }

rvsp_status_t rvsp_hip_plan_analyse(rvsp_hip_plan_t *p, int32_t m, int32_t k, int32_t n,
                                    int32_t nnz_a, const int32_t *arp, const int32_t *aci,
                                    int32_t nnz_b, const int32_t *brp, const int32_t *bci,
                                    rvsp_select_t select, int64_t *c_nnz, int64_t *products)
{
    if (!p || !arp || !aci || !brp || !bci) return RVSP_ERROR_NULL_POINTER;
    if (m <= 0 || k <= 0 || n <= 0 || nnz_a < 0 || nnz_b < 0) return RVSP_ERROR_INVALID_ARGUMENT;
    hipStream_t st = p->stream;
    plan_release(p);
    p->m = m;
    p->k = k;
    p->n = n;
    p->nnz_a = nnz_a;
    p->nnz_b = nnz_b;
    p->tp = tune_for(nnz_a, nnz_b, k, n, /*may_load=*/p->arena == nullptr);
    p->arp = arp;
    p->aci = aci;
    p->brp = brp;
    p->bci = bci;

    // ---- one carved allocation for everything sized by m
    const size_t M = (size_t)m;
    const int ntiles = scan_tiles(m);
    size_t off = 0;
    const size_t o_ops = off; off = align256(off + M * 8);
    const size_t o_lo = off; off = align256(off + M * 4);
    const size_t o_hi = off; off = align256(off + M * 4);
    const size_t o_sbin = off; off = align256(off + M);
    const size_t o_nbin = off; off = align256(off + M);
    const size_t o_hbin = off; off = align256(off + M);
    const size_t o_fbin = off; off = align256(off + M);
    const size_t o_cnt = off; off = align256(off + M * 4);
    const size_t o_crp = off; off = align256(off + (M + 1) * 4);
    const size_t o_sperm = off; off = align256(off + M * 4);
    const size_t o_nperm = off; off = align256(off + M * 4);
    const size_t o_tsum = off; off = align256(off + (size_t)ntiles * 8);
    const size_t o_ctr = off; off = align256(off + sizeof(AnalysisCounters));
    const size_t o_binoff = off; off = align256(off + 128 * 4);
    const size_t o_border = off; off = align256(off + kBigOrderWork * 4);
    const size_t o_err = off; off = align256(off + 4);
    const size_t o_bst = off; off = align256(off + (size_t)k * sizeof(int4));
    const size_t o_parts = off; off = align256(off + (size_t)kStatsParts * sizeof(StatPart));
    const size_t o_segoff = off; off = align256(off + (M + 1) * 4);
    rvsp_status_t s = palloc(p, &p->rowblock, off);
    if (s != RVSP_SUCCESS) return s;
    char *base = (char *)p->rowblock;
    p->ops = (int64_t *)(base + o_ops);
    p->lo = (int32_t *)(base + o_lo);
    p->hi = (int32_t *)(base + o_hi);
    p->sbin = (uint8_t *)(base + o_sbin);
    p->nbin = (uint8_t *)(base + o_nbin);
    p->hbin = (uint8_t *)(base + o_hbin);
    p->fbin = (uint8_t *)(base + o_fbin);
    p->cnt = (int32_t *)(base + o_cnt);
    p->crp = (int32_t *)(base + o_crp);
    p->sperm = (int32_t *)(base + o_sperm);
    p->nperm = (int32_t *)(base + o_nperm);
    p->tsum = (int64_t *)(base + o_tsum);
    p->ctr = (AnalysisCounters *)(base + o_ctr);
    p->binoff = (int32_t *)(base + o_binoff);
    int32_t *border = (int32_t *)(base + o_border);
    p->err = (int *)(base + o_err);
    int4 *bst = (int4 *)(base + o_bst);
    StatPart *parts = (StatPart *)(base + o_parts);
    int32_t *segoff = (int32_t *)(base + o_segoff);

    PhaseScope timer(st, PH_ANALYSE);
    HIP_TRY(hipMemsetAsync(p->ctr, 0, sizeof(AnalysisCounters), st));
    HIP_TRY(hipMemsetAsync(p->binoff, 0, 128 * 4, st));
    HIP_TRY(hipMemsetAsync(p->err, 0, 4, st));

    // ---- 0. operands the host has not checked: canonical CSR, on the
    // device, before any kernel trusts their indices. The host reads the
    // verdict with the row statistics' counters (one synchronisation less):
    // the statistics clamp every index they read through, and nothing else
    // runs before the verdict is read.
    const bool checked = p->operands_checked;
    p->operands_checked = false;
    if (!checked) {
        const bool same = arp == brp && aci == bci && m == k && k == n && nnz_a == nnz_b;
        const int32_t most = std::max(nnz_a, nnz_b);
        void *starts = nullptr;
        s = palloc(p, &starts, (((size_t)most + 32) / 32) * 4);
        if (s != RVSP_SUCCESS) return s;
        hipError_t e = launch_csr_validate(st, m, k, nnz_a, arp, aci, (uint32_t *)starts, p->err);
        if (e == hipSuccess && !same)
            e = launch_csr_validate(st, k, n, nnz_b, brp, bci, (uint32_t *)starts, p->err);
        pfree(p, starts);  // stream-ordered: reused only by later work on st
        HIP_TRY(e);
    }

    // learned selectors (null when none is loaded: heuristic rules)
    float margin = 1.0f, cmargin = 1.0f;
    const float *sizer =
        select == RVSP_SELECT_LEARNED ? selector_device_params(RVSP_SELECTOR_SIZER, &margin) : nullptr;
    const float *chooser =
        select == RVSP_SELECT_LEARNED ? selector_device_params(RVSP_SELECTOR_CHOOSER, &cmargin) : nullptr;
    p->sized_rows = p->fallback_rows = 0;

    // ---- 1. per-row product counts, spans and symbolic bins
    // (the symbolic row order, filled later, lends its m ints to the list of
    // long rows)
    launch_row_stats(st, m, nnz_a, arp, aci, k, nnz_b, brp, bci, bst, p->ops, p->lo, p->hi,
                     p->sbin, p->ctr, p->tp.hash_max_u, p->sperm, parts);
    HIP_TRY(hipGetLastError());
    if (sizer) {
        launch_sel_bins(st, m, sizer, margin, p->ops, arp, p->lo, p->hi, p->sbin, p->hbin, p->ctr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipMemcpyAsync(&p->hctr, p->ctr, sizeof(AnalysisCounters), hipMemcpyDeviceToHost, st));
    int herr = 0;
    if (!checked) HIP_TRY(hipMemcpyAsync(&herr, p->err, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (herr) return RVSP_ERROR_INVALID_CSR;
    p->products = (int64_t)p->hctr.products;

    int32_t *hoff = p->h_off[0];
    memset(p->h_off, 0, sizeof(p->h_off));
    // bin lists of the symbolic pass from the counters' histogram
    auto scatter_sym = [&]() -> rvsp_status_t {
        int run = 0;
        for (int b = 0; b < SB_COUNT; ++b) {
            p->sym_count[b] = p->hctr.sym_hist[b];
            p->sym_off[b] = run;
            hoff[b] = run;
            run += p->sym_count[b];
        }
        HIP_TRY(hipMemcpyAsync(p->binoff, hoff, 64 * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(p->binoff + 64, 0, 64 * 4, st));
        launch_bin_scatter(st, m, p->sbin, p->binoff, p->binoff + 64, p->sperm, SB_COUNT);
        HIP_TRY(hipGetLastError());
        // big rows largest first (nperm is free until the numeric bins)
        HIP_TRY(launch_big_order(st, p->sperm + p->sym_off[SB_BIG], p->sym_count[SB_BIG], p->ops,
                                 border, p->nperm));
        return RVSP_SUCCESS;
    };
    s = scatter_sym();
    if (s != RVSP_SUCCESS) return s;

    // bitmaps for every row the heuristic sends to the big-row path (a
    // superset of the learned big rows: overflowing rows may return there)
    const bool any_big = p->hctr.bits_cap > 0;
    p->big = BigArgs{};
    p->big.arp = arp;
    p->big.aci = aci;
    p->big.brp = brp;
    p->big.bci = bci;
    p->big.lo = p->lo;
    p->big.hi = p->hi;
    p->big.ops = p->ops;
    p->big.cnt = p->cnt;
    p->big.bits_fill = &p->ctr->bits_fill;
    p->big.seg_products = p->tp.seg_products;
    p->big.sym_threads = p->tp.sym_big_threads;
    {
        // bitmap pass: the whole span of any row when B has few enough
        // columns (chunk-aligned spans need at most n/1024 + 1 chunks)
        const int64_t need = ((int64_t)n / kChunkCols + 1) * kChunkWords;
        p->big.pass_words = (int)(need < kBigPassWords ? need : kBigPassWords);
        const int shape = big_window_shape(p->tp.big_window);
        p->big.win_outs = kBigShapes[shape][0];
        p->big.win_chunks = kBigShapes[shape][1];
    }
    if (any_big) {
        const size_t words = (size_t)p->hctr.bits_cap;
        const size_t chunks = words / kChunkWords;
        size_t bo = 0;
        const size_t o_off = bo; bo = align256(bo + M * 8);
        const size_t o_nwin = bo; bo = align256(bo + M * 4);
        const size_t o_bits = bo; bo = align256(bo + words * 4);
        const size_t o_chunk = bo; bo = align256(bo + chunks * 4);
        const size_t o_wc0 = bo; bo = align256(bo + chunks * 4);
        const size_t o_wout = bo; bo = align256(bo + chunks * 4);
        const size_t o_segs = bo; bo = align256(bo + chunks * 16);
        const size_t o_scur = bo; bo = align256(bo + chunks * 8);
        s = palloc(p, &p->bigblock, bo);
        if (s != RVSP_SUCCESS) {
            plan_release(p);
            return s;
        }
        char *bb = (char *)p->bigblock;
        p->big.big_off = (int64_t *)(bb + o_off);
        p->big.big_nwin = (int32_t *)(bb + o_nwin);
        p->big.gbits = (uint32_t *)(bb + o_bits);
        p->big.gchunk = (int32_t *)(bb + o_chunk);
        p->big.win_chunk0 = (int32_t *)(bb + o_wc0);
        p->big.win_out = (int32_t *)(bb + o_wout);
        p->big.segs = (int4 *)(bb + o_segs);
        p->big.seg_cur = (int64_t *)(bb + o_scur);
        // B's rows as (word, bits) pairs for the count (freed once counted)
        const size_t NB = (size_t)nnz_b;
        size_t wo = 0;
        const size_t o_flag = wo; wo = align256(wo + NB * 4);
        const size_t o_pidx = wo; wo = align256(wo + (NB + 1) * 4);
        const size_t o_wts = wo; wo = align256(wo + (size_t)scan_tiles(nnz_b) * 8);
        const size_t o_wtot = wo; wo = align256(wo + 8);
        const size_t o_bwm = wo; wo = align256(wo + NB * 8);
        const size_t o_bwp = wo; wo = align256(wo + ((size_t)k + 1) * 4);
        s = palloc(p, &p->wordblock, wo);
        if (s != RVSP_SUCCESS) {
            plan_release(p);
            return s;
        }
        char *wb = (char *)p->wordblock;
        launch_big_words(st, k, nnz_b, brp, bci, (int32_t *)(wb + o_flag), (int32_t *)(wb + o_pidx),
                         (int64_t *)(wb + o_wts), (int64_t *)(wb + o_wtot), (int2 *)(wb + o_bwm),
                         (int32_t *)(wb + o_bwp));
        HIP_TRY(hipGetLastError());
        p->big.bwm = (const int2 *)(wb + o_bwm);
        p->big.bwp = (const int32_t *)(wb + o_bwp);
    }

    // ---- 2. exact nnz(C_i)
    if (p->sym_count[SB_BMS] > 0 && !bms_spill_off()) {
        // count-pass bitmaps of the one-pass-bitmap rows, read back by the fill
        s = palloc(p, (void **)&p->bms_spill, (size_t)p->sym_count[SB_BMS] * kBmsWords * 4);
        if (s != RVSP_SUCCESS) {
            plan_release(p);
            return s;
        }
    }
    p->bms_spill_ok = p->bms_spill != nullptr;
    SymArgs sa{arp,    aci,    brp,    bci,    p->lo,  p->hi,       p->cnt,
               p->crp, nullptr, p->big, p->err, p->ops, p->bms_spill};
    // rows without products get no kernel, so their count must start at 0
    HIP_TRY(hipMemsetAsync(p->cnt, 0, M * 4, st));
    const bool sized = sizer && p->hctr.moved > 0;
    // (the bins' products spread them over the streams; a learned sizing
    // moves rows after they were counted, so its bins go in turn)
    launch_sym_count(st, sa, p->sperm, p->sym_count, p->sym_off, sized,
                     fan_out(p, p->products, p->tp.fan_min_products),
                     sized ? nullptr : p->hctr.sym_ops);
    HIP_TRY(hipGetLastError());
    if (sized) {
        // rows whose learned table overflowed go back to their heuristic
        // bin and are counted again there; the bin lists are rebuilt
        launch_sel_fallback(st, m, p->cnt, p->sbin, p->hbin, p->fbin, p->ctr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&p->hctr, p->ctr, sizeof(AnalysisCounters), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        int nfb = 0;
        for (int b = 0; b < SB_COUNT; ++b) nfb += p->hctr.fb_hist[b];
        p->sized_rows = p->hctr.moved;
        p->fallback_rows = nfb;
        // a recount rebuilds the bin lists: bitmap slots no longer match rows
        if (nfb > 0) p->bms_spill_ok = false;
        if (nfb > 0) {
            int32_t fcount[SB_COUNT], foff[SB_COUNT];
            int32_t *fho = p->h_off[1];
            int run = 0;
            for (int b = 0; b < SB_COUNT; ++b) {
                fcount[b] = b == SB_EMPTY ? m - nfb : p->hctr.fb_hist[b];
                foff[b] = run;
                fho[b] = run;
                run += fcount[b];
            }
            HIP_TRY(hipMemcpyAsync(p->binoff, fho, 64 * 4, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemsetAsync(p->binoff + 64, 0, 64 * 4, st));
            // nperm is free until the numeric bins are scattered
            launch_bin_scatter(st, m, p->fbin, p->binoff, p->binoff + 64, p->nperm, SB_COUNT);
            HIP_TRY(hipGetLastError());
            launch_sym_count(st, sa, p->nperm, fcount, foff, false, fan_out(p, p->products, p->tp.fan_min_products));
            HIP_TRY(hipGetLastError());
            s = scatter_sym();
            if (s != RVSP_SUCCESS) return s;
        }
    }
    // the counts are done: B's word pairs are no longer read
    pfree(p, p->wordblock);
    p->wordblock = nullptr;
    p->big.bwm = nullptr;
    p->big.bwp = nullptr;
    // ---- 3. row pointer
    launch_scan(st, m, p->cnt, p->tsum, (int64_t *)&p->ctr->nnz_total, p->crp);
    HIP_TRY(hipGetLastError());
    // ---- 4. big-row windows and segments, numeric bins
    if (p->sym_count[SB_BIG] > 0) {
        // (nperm, free until the numeric bins are scattered, holds the
        // per-row segment counts; the row-pointer scan's tiles are done)
        launch_big_plan(st, p->sperm + p->sym_off[SB_BIG], p->sym_count[SB_BIG], p->big, p->crp,
                        p->ctr, p->nperm, segoff, p->tsum);
        HIP_TRY(hipGetLastError());
    }
    launch_num_bins(st, m, p->crp, p->ops, arp, p->sbin, p->lo, p->hi, chooser,
                    cmargin > 1.0f ? log2f(cmargin) : 0.0f, p->products, p->nbin, p->ctr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&p->hctr, p->ctr, sizeof(AnalysisCounters), hipMemcpyDeviceToHost, st));
    // the counts' error word (a full hash table) comes back in the same
    // synchronisation; no later analysis kernel raises one
    HIP_TRY(hipMemcpyAsync(&herr, p->err, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (herr) return RVSP_ERROR_INVALID_CSR;
    p->nnzc = p->hctr.nnz_total;
    p->nseg = (int)p->hctr.seg_total;
    p->big.nsegs = p->nseg;
    if (p->nseg > 0) {
        s = palloc(p, (void **)&p->cursors, (size_t)p->hctr.cur_total * sizeof(int4));
        if (s != RVSP_SUCCESS) {
            plan_release(p);
            return s;
        }
        p->big.cursors = p->cursors;
    }
    hoff = p->h_off[2];
    {
        int run = 0;
        for (int b = 0; b < NB_COUNT; ++b) {
            p->num_count[b] = p->hctr.num_hist[b];
            p->num_off[b] = run;
            hoff[b] = run;
            run += p->num_count[b];
        }
    }
    HIP_TRY(hipMemcpyAsync(p->binoff, hoff, 64 * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(p->binoff + 64, 0, 64 * 4, st));
    launch_bin_scatter(st, m, p->nbin, p->binoff, p->binoff + 64, p->nperm, NB_COUNT);
    HIP_TRY(hipGetLastError());
    if (c_nnz) *c_nnz = p->nnzc;
    if (products) *products = p->products;
    if (p->nnzc > INT32_MAX) return RVSP_ERROR_ALLOCATION_FAILED;
    s = palloc(p, (void **)&p->bcv, (size_t)nnz_b * sizeof(int2));
    if (s != RVSP_SUCCESS) {
        plan_release(p);
        return s;
    }
    p->analysed = true;
    return RVSP_SUCCESS;
}

// Products per symbolic bin for the fill's stream balance (null after a
// learned sizing, which moves rows between bins after they were counted).
static const unsigned long long *bin_ops(const rvsp_hip_plan *p)
{
    return p->sized_rows > 0 ? nullptr : p->hctr.sym_ops;
}

rvsp_status_t rvsp_hip_plan_fill(rvsp_hip_plan_t *p, int32_t *cci)
{
    if (!p || !cci) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    SymArgs sa{p->arp, p->aci, p->brp, p->bci, p->lo,  p->hi,  p->cnt,
               p->crp, cci,    p->big, p->err, p->ops, p->bms_spill_ok ? p->bms_spill : nullptr};
    PhaseScope timer(p->stream, PH_FILL);
    launch_sym_fill(p->stream, sa, p->sperm, p->sym_count, p->sym_off, false,
                    fan_out(p, p->nnzc, p->tp.fan_min_fill), bin_ops(p));
    HIP_TRY(hipGetLastError());
    return RVSP_SUCCESS;
}

static NumArgs num_args(rvsp_hip_plan *p, const float *aval, const float *bval, const int32_t *cci,
                        float *cval)
{
    return NumArgs{p->arp, p->aci, aval,    p->brp, p->bci, bval,
                   p->bcv, p->crp, cci,     cval,   p->lo,  p->err};
}

static void queue_numeric(rvsp_hip_plan *p, const float *aval, const float *bval,
                          const int32_t *cci, float *cval, int32_t *big_cols = nullptr)
{
    BigArgs big = p->big;
    big.cols_out = big_cols;
    // kernels that gather B's columns read the (column, value) pairs; the
    // rank kernels read only values, so a pass with ranks and no big rows
    // skips the pairs
    if (!p->pm_ready || p->nseg > 0) launch_interleave(p->stream, p->nnz_b, p->bci, bval, p->bcv);
    const NumArgs a = num_args(p, aval, bval, cci, cval);
    if (p->pm_ready)
        launch_numeric_ranked(p->stream, a, p->nperm, p->num_count, p->num_off, big, p->nseg,
                              p->prow, p->pmap, fan_out(p, p->products, p->tp.fan_min_products),
                              p->hctr.num_ops);
    else
        launch_numeric(p->stream, a, p->nperm, p->num_count, p->num_off, big, p->nseg,
                       fan_out(p, p->products, p->tp.fan_min_products), p->hctr.num_ops);
}

rvsp_status_t rvsp_hip_plan_numeric(rvsp_hip_plan_t *p, const float *aval, const float *bval,
                                    const int32_t *cci, float *cval)
{
    if (!p || !aval || !bval || !cci || !cval) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    PhaseScope timer(p->stream, PH_NUMERIC);
    queue_numeric(p, aval, bval, cci, cval);
    HIP_TRY(hipGetLastError());
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_fill_numeric(rvsp_hip_plan_t *p, const float *aval,
                                         const float *bval, int32_t *cci, float *cval)
{
    if (!p || !aval || !bval || !cci || !cval) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    {
        SymArgs sa{p->arp, p->aci, p->brp, p->bci, p->lo,  p->hi,  p->cnt,
                   p->crp, cci,    p->big, p->err, p->ops, p->bms_spill_ok ? p->bms_spill : nullptr};
        PhaseScope timer(p->stream, PH_FILL);
        launch_sym_fill(p->stream, sa, p->sperm, p->sym_count, p->sym_off, /*skip_big=*/true,
                        fan_out(p, p->nnzc, p->tp.fan_min_fill), bin_ops(p));
        HIP_TRY(hipGetLastError());
    }
    PhaseScope timer(p->stream, PH_NUMERIC);
    queue_numeric(p, aval, bval, cci, cval, cci);
    HIP_TRY(hipGetLastError());
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_numeric_typed(rvsp_hip_plan_t *p, rvsp_dtype_t in_dtype,
                                          const void *aval, const void *bval, const int32_t *cci,
                                          void *cval)
{
    if (!p || !aval || !bval || !cci || !cval) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    if (in_dtype != RVSP_DTYPE_FP64 && in_dtype != RVSP_DTYPE_INT8)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    hipStream_t st = p->stream;
    PhaseScope timer(st, PH_NUMERIC);
    // non-empty, non-big rows: the numeric bins before NB_WIN, contiguous in nperm
    const int32_t *rows = p->nperm + p->num_off[NB_HASH0];
    const int nrows = p->num_off[NB_WIN] - p->num_off[NB_HASH0];
    rvsp_status_t s;
    if (in_dtype == RVSP_DTYPE_FP64) {
        PairD *pairs = nullptr;
        if ((s = dalloc((void **)&pairs, (size_t)p->nnz_b * sizeof(PairD), st)) != RVSP_SUCCESS)
            return s;
        launch_pairs_typed(st, p->nnz_b, p->bci, (const double *)bval, pairs);
        const NumArgsD a{p->arp, p->aci, (const double *)aval, p->brp, p->bci,
                         (const double *)bval, pairs, p->crp, cci, (double *)cval, p->lo, p->err};
        launch_numeric_typed(st, a, rows, nrows, p->big, p->nseg);
        dfree(pairs, st);
    } else {
        int2 *pairs = nullptr;
        if ((s = dalloc((void **)&pairs, (size_t)p->nnz_b * sizeof(int2), st)) != RVSP_SUCCESS)
            return s;
        launch_pairs_typed(st, p->nnz_b, p->bci, (const int8_t *)bval, pairs);
        const NumArgsI a{p->arp, p->aci, (const int8_t *)aval, p->brp, p->bci,
                         (const int8_t *)bval, pairs, p->crp, cci, (int32_t *)cval, p->lo, p->err};
        launch_numeric_typed(st, a, rows, nrows, p->big, p->nseg);
        dfree(pairs, st);
    }
    HIP_TRY(hipGetLastError());
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_drop_zeros(rvsp_hip_plan_t *p, rvsp_dtype_t value_dtype,
                                       const int32_t *cci, const void *cval, int32_t *rp_out,
                                       int32_t *ci_out, void *cv_out, int32_t *nnz_out)
{
    if (!p || !cci || !cval || !rp_out || !ci_out || !cv_out || !nnz_out)
        return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    if (value_dtype != RVSP_DTYPE_FP64 && value_dtype != RVSP_DTYPE_INT32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    hipStream_t st = p->stream;
    const size_t M = (size_t)p->m;
    const int ntiles = scan_tiles(p->m);
    size_t off = 0;
    const size_t o_cnt = off; off = align256(off + M * 4);
    const size_t o_ts = off; off = align256(off + (size_t)ntiles * 8);
    const size_t o_tot = off; off = align256(off + 8);
    void *scratch = nullptr;
    rvsp_status_t s = dalloc(&scratch, off, st);
    if (s != RVSP_SUCCESS) return s;
    char *b = (char *)scratch;
    if (value_dtype == RVSP_DTYPE_FP64)
        launch_drop_zeros(st, p->m, p->crp, cci, (const double *)cval, (int32_t *)(b + o_cnt),
                          (int64_t *)(b + o_ts), (int64_t *)(b + o_tot), rp_out, ci_out,
                          (double *)cv_out);
    else
        launch_drop_zeros(st, p->m, p->crp, cci, (const int32_t *)cval, (int32_t *)(b + o_cnt),
                          (int64_t *)(b + o_ts), (int64_t *)(b + o_tot), rp_out, ci_out,
                          (int32_t *)cv_out);
    HIP_TRY(hipGetLastError());
    int32_t kept = 0;
    HIP_TRY(hipMemcpyAsync(&kept, rp_out + p->m, 4, hipMemcpyDeviceToHost, st));
    dfree(scratch, st);
    HIP_TRY(hipStreamSynchronize(st));
    *nnz_out = kept;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_prepare_reuse(rvsp_hip_plan_t *p, const int32_t *cci)
{
    if (!p || !cci) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    if (p->pm_ready || p->pm_skip || reuse_ranks_off()) return RVSP_SUCCESS;
    hipStream_t st = p->stream;
    const size_t M = (size_t)p->m;
    const int ntiles = pm_scan_tiles(p->m);
    size_t off = 0;
    const size_t o_prow = off; off = align256(off + M * 8);
    const size_t o_cnt = off; off = align256(off + M * 8);
    const size_t o_tsum = off; off = align256(off + (size_t)ntiles * 8);
    const size_t o_tot = off; off = align256(off + 8);
    rvsp_status_t s = dalloc(&p->pmblock, off, st);
    if (s != RVSP_SUCCESS) {
        p->pm_skip = true;  // the plain kernels stay in use
        return RVSP_SUCCESS;
    }
    char *b = (char *)p->pmblock;
    p->prow = (int64_t *)(b + o_prow);
    int64_t *cnt = (int64_t *)(b + o_cnt), *tsum = (int64_t *)(b + o_tsum);
    int64_t *tot = (int64_t *)(b + o_tot);
    launch_pm_build(st, p->m, p->nbin, p->ops, p->arp, p->aci, p->brp, p->bci, p->crp, cci, cnt,
                    tsum, tot, p->prow, nullptr, p->err);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&p->pm_bytes, tot, 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // 1 or 2 bytes per product; keep to a quarter of the free device memory
    size_t free_b = 0, total_b = 0;
    const size_t need = (size_t)p->pm_bytes;
    if (p->pm_bytes <= 0 || hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 4 ||
        dalloc((void **)&p->pmap, need, st) != RVSP_SUCCESS) {
        dfree(p->pmblock, st);
        p->pmblock = nullptr;
        p->prow = nullptr;
        p->pm_skip = true;
        return RVSP_SUCCESS;
    }
    launch_pm_build(st, p->m, p->nbin, p->ops, p->arp, p->aci, p->brp, p->bci, p->crp, cci, cnt,
                    tsum, tot, p->prow, p->pmap, p->err);
    HIP_TRY(hipGetLastError());
    p->pm_ready = true;
    drop_graph(p);  // a captured pass would still use the plain kernels
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_numeric_graph(rvsp_hip_plan_t *p, const float *aval,
                                          const float *bval, const int32_t *cci, float *cval)
{
    if (!p || !aval || !bval || !cci || !cval) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    const void *key[4] = {aval, bval, cci, cval};
    if (p->gexec && memcmp(key, p->gkey, sizeof(key)) != 0) {
        (void)hipGraphExecDestroy(p->gexec);
        (void)hipGraphDestroy(p->graph);
        p->gexec = nullptr;
        p->graph = nullptr;
    }
    if (!p->gexec) {
        HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        queue_numeric(p, aval, bval, cci, cval);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(p->stream, &g);
        if (e != hipSuccess) return map_err(e);
        e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return map_err(e);
        }
        p->graph = g;
        memcpy(p->gkey, key, sizeof(key));
    }
    PhaseScope timer(p->stream, PH_NUMERIC);
    HIP_TRY(hipGraphLaunch(p->gexec, p->stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_check(rvsp_hip_plan_t *p)
{
    if (!p) return RVSP_ERROR_NULL_POINTER;
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (!p->err) return RVSP_SUCCESS;
    int herr = 0;
    HIP_TRY(hipMemcpyAsync(&herr, p->err, 4, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return herr ? RVSP_ERROR_INVALID_CSR : RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_get_bins(rvsp_hip_plan_t *p, uint8_t *sym_bins, uint8_t *num_bins)
{
    if (!p) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    const size_t M = (size_t)p->m;
    if (sym_bins) HIP_TRY(hipMemcpyAsync(sym_bins, p->sbin, M, hipMemcpyDeviceToHost, p->stream));
    if (num_bins) HIP_TRY(hipMemcpyAsync(num_bins, p->nbin, M, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_set_num_bins(rvsp_hip_plan_t *p, const uint8_t *num_bins)
{
    if (!p || !num_bins) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    hipStream_t st = p->stream;
    std::vector<uint8_t> sb((size_t)p->m), nb((size_t)p->m);
    HIP_TRY(hipMemcpyAsync(sb.data(), p->sbin, sb.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(nb.data(), p->nbin, nb.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int32_t hist[64] = {0};
    for (int32_t i = 0; i < p->m; ++i) {
        const int b = num_bins[i];
        // empty rows and big rows keep their bins; the rest take a hash or dense bin
        const bool fixed = nb[i] == NB_EMPTY || nb[i] == NB_WIN;
        if (fixed ? b != nb[i] : (b < NB_HASH0 || b >= NB_WIN)) return RVSP_ERROR_INVALID_ARGUMENT;
        ++hist[b];
    }
    int32_t hoff[64] = {0};
    int run = 0;
    for (int b = 0; b < NB_COUNT; ++b) {
        p->num_count[b] = hist[b];
        p->num_off[b] = run;
        hoff[b] = run;
        run += hist[b];
    }
    HIP_TRY(hipMemcpyAsync(p->nbin, num_bins, (size_t)p->m, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->binoff, hoff, 64 * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(p->binoff + 64, 0, 64 * 4, st));
    launch_bin_scatter(st, p->m, p->nbin, p->binoff, p->binoff + 64, p->nperm, NB_COUNT);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    drop_graph(p);
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_features(rvsp_hip_plan_t *p, float *feats, float *ratio)
{
    if (!p || !feats || !ratio) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed) return RVSP_ERROR_INVALID_ARGUMENT;
    launch_sel_features(p->stream, p->m, p->ops, p->arp, p->lo, p->hi, p->crp, feats, ratio);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(p->stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_selector_stats(const rvsp_hip_plan_t *p, int64_t *sized_rows,
                                           int64_t *fallback_rows)
{
    if (!p) return RVSP_ERROR_NULL_POINTER;
    if (sized_rows) *sized_rows = p->sized_rows;
    if (fallback_rows) *fallback_rows = p->fallback_rows;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_plan_time_bin(rvsp_hip_plan_t *p, int32_t bin, const int32_t *rows,
                                     int32_t nrows, const float *aval, const float *bval,
                                     const int32_t *cci, float *cval, int32_t reps, float *ms)
{
    if (!p || !rows || !aval || !bval || !cci || !cval || !ms) return RVSP_ERROR_NULL_POINTER;
    if (!p->analysed || nrows <= 0 || reps <= 0 || bin < NB_HASH0 || bin >= NB_WIN)
        return RVSP_ERROR_INVALID_ARGUMENT;
    hipStream_t st = p->stream;
    const NumArgs a = num_args(p, aval, bval, cci, cval);
    HIP_TRY(hipMemsetAsync(p->err, 0, 4, st));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    rvsp_status_t s = RVSP_SUCCESS;
    launch_interleave(st, p->nnz_b, p->bci, bval, p->bcv);
    (void)launch_num_bin(st, a, bin, rows, nrows);  // warm-up
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) (void)launch_num_bin(st, a, bin, rows, nrows);
    (void)hipEventRecord(e1, st);
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    else
        *ms /= (float)reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (s != RVSP_SUCCESS) return s;
    int herr = 0;
    HIP_TRY(hipMemcpyAsync(&herr, p->err, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemsetAsync(p->err, 0, 4, st));
    return herr ? RVSP_ERROR_INVALID_ARGUMENT : RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_stream_sync(void *stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_copy_async(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return RVSP_SUCCESS;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_alloc_async(void **ptr, size_t bytes, void *stream)
{
    if (!ptr) return RVSP_ERROR_NULL_POINTER;
    rvsp_status_t s = device_ready();
    if (s != RVSP_SUCCESS) return s;
    return dalloc(ptr, bytes, (hipStream_t)stream);
}

rvsp_status_t rvsp_hip_free_async(void *ptr, void *stream)
{
    dfree(ptr, (hipStream_t)stream);
    return RVSP_SUCCESS;
}

void rvsp_hip_fail_next_allocs(int n) { g_fail_allocs = n; }

void rvsp_hip_set_timing(int on)
{
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing = on != 0;
}

rvsp_status_t rvsp_hip_get_timing(double *ms, long long *calls, int n)
{
    if (!ms || !calls) return RVSP_ERROR_NULL_POINTER;
    std::vector<PendingTime> todo;
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        todo.swap(g_pending);
    }
    for (auto &t : todo) {
        float e = 0.0f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&e, t.a, t.b) == hipSuccess) {
            g_ms[t.phase] += e;
            g_calls[t.phase] += 1;
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (int i = 0; i < n && i < PH_COUNT; ++i) {
        ms[i] = g_ms[i];
        calls[i] = g_calls[i];
    }
    return RVSP_SUCCESS;
}

void rvsp_hip_reset_timing(void)
{
    double ms[PH_COUNT];
    long long c[PH_COUNT];
    (void)rvsp_hip_get_timing(ms, c, PH_COUNT);
    for (int i = 0; i < PH_COUNT; ++i) {
        g_ms[i] = 0.0;
        g_calls[i] = 0;
    }
}

// ---- public helpers (rv_sparse_hip.h) ------------------------------------

rvsp_status_t rvsp_hip_available(int *count)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
    return c > 0 ? RVSP_SUCCESS : RVSP_ERROR_UNSUPPORTED_BACKEND;
}

rvsp_status_t rvsp_hip_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return RVSP_ERROR_NULL_POINTER;
    *ptr = nullptr;
    rvsp_status_t s = device_ready();
    if (s != RVSP_SUCCESS) return s;
    if (g_fail_allocs > 0) {
        --g_fail_allocs;
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    // stream-ordered on this thread's stream, never on the legacy null
    // stream: a copy there fails while any other thread captures a graph
    const hipStream_t ts = thread_stream();
    HIP_TRY(hipMallocAsync(ptr, bytes ? bytes : 256, ts));
    HIP_TRY(hipStreamSynchronize(ts));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_free(void *ptr)
{
    if (!ptr) return RVSP_SUCCESS;
    const hipStream_t ts = thread_stream();
    HIP_TRY(hipFreeAsync(ptr, ts));
    HIP_TRY(hipStreamSynchronize(ts));
    return RVSP_SUCCESS;
}

static rvsp_status_t sync_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (bytes == 0) return RVSP_SUCCESS;
    if (!dst || !src) return RVSP_ERROR_NULL_POINTER;
    const hipStream_t ts = thread_stream();
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, ts));
    HIP_TRY(hipStreamSynchronize(ts));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return sync_copy(dst, src, bytes, hipMemcpyHostToDevice);
}

rvsp_status_t rvsp_hip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return sync_copy(dst, src, bytes, hipMemcpyDeviceToHost);
}

rvsp_status_t rvsp_hip_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    return sync_copy(dst, src, bytes, hipMemcpyDeviceToDevice);
}

rvsp_status_t rvsp_hip_memset(void *dst, int value, size_t bytes)
{
    if (bytes == 0) return RVSP_SUCCESS;
    if (!dst) return RVSP_ERROR_NULL_POINTER;
    const hipStream_t ts = thread_stream();
    HIP_TRY(hipMemsetAsync(dst, value, bytes, ts));
    HIP_TRY(hipStreamSynchronize(ts));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return RVSP_SUCCESS;
}

struct rvsp_hip_event {
    hipEvent_t ev;
};

rvsp_status_t rvsp_hip_event_create(rvsp_hip_event_t *ev)
{
    if (!ev) return RVSP_ERROR_NULL_POINTER;
    *ev = nullptr;
    rvsp_status_t s = device_ready();
    if (s != RVSP_SUCCESS) return s;
    rvsp_hip_event *e = new (std::nothrow) rvsp_hip_event();
    if (!e) return RVSP_ERROR_ALLOCATION_FAILED;
    if (hipEventCreate(&e->ev) != hipSuccess) {
        delete e;
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    *ev = e;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_event_record(rvsp_hip_event_t ev, void *stream)
{
    if (!ev) return RVSP_ERROR_NULL_POINTER;
    HIP_TRY(hipEventRecord(ev->ev, (hipStream_t)stream));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_event_synchronize(rvsp_hip_event_t ev)
{
    if (!ev) return RVSP_ERROR_NULL_POINTER;
    HIP_TRY(hipEventSynchronize(ev->ev));
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_event_elapsed_ms(rvsp_hip_event_t a, rvsp_hip_event_t b, float *ms)
{
    if (!a || !b || !ms) return RVSP_ERROR_NULL_POINTER;
    HIP_TRY(hipEventElapsedTime(ms, a->ev, b->ev));
    return RVSP_SUCCESS;
}

void rvsp_hip_event_destroy(rvsp_hip_event_t ev)
{
    if (!ev) return;
    (void)hipEventDestroy(ev->ev);
    delete ev;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free device one-shot (rvsp_spgemm_csr_device_noalloc).
//
// Its plan arrays (per-row counts and bins, big-row bitmaps, windows and
// cursors, B's interleaved pairs, the operand check's row marks) live in
// `scratch`, a fixed device array of RVSP_NOALLOC_SCRATCH_BYTES that every
// GPU the library runs on holds until the process ends. Host code never
// passes `scratch` itself (in host code the name stands for a host-side
// shadow, not device memory): on a device's first call its device address
// is looked up with hipGetSymbolAddress, with that device current, and kept
// per device; kernels and copies get that address. The host-side plan is a
// static object per device too, so a call allocates neither device nor host
// memory. A per-device lock serialises calls from several host threads;
// calls on one GPU from several processes would share nothing, since each
// process holds its own copy of the array.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocDevices = 16;

static __device__ __align__(256) unsigned char scratch[RVSP_NOALLOC_SCRATCH_BYTES];

struct NoAllocSlot {
    std::mutex mu;
    char *scratch_ptr = nullptr;  // device address of `scratch` on this device
    rvsp_hip_plan plan;
};
NoAllocSlot g_noalloc[kNoAllocDevices];

} // namespace

extern "C" rvsp_status_t rvsp_spgemm_csr_device_noalloc(const rvsp_csr_matrix_t *A,
                                                        const rvsp_csr_matrix_t *B,
                                                        rvsp_csr_matrix_t *C, void *stream)
{
    if (!A || !B || !C) return RVSP_ERROR_NULL_POINTER;
    if (!A->row_ptr || !A->col_idx || !A->values || !B->row_ptr || !B->col_idx || !B->values ||
        !C->row_ptr || !C->col_idx || !C->values)
        return RVSP_ERROR_NULL_POINTER;
    if (A->dtype != RVSP_DTYPE_FP32 || B->dtype != RVSP_DTYPE_FP32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (A->rows <= 0 || A->cols <= 0 || B->cols <= 0 || A->cols != B->rows || A->nnz < 0 ||
        B->nnz < 0 || C->nnz < 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    rvsp_status_t s = device_ready();
    if (s != RVSP_SUCCESS) return s;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return RVSP_ERROR_UNSUPPORTED_BACKEND;
    NoAllocSlot &slot = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(slot.mu);
    if (!slot.scratch_ptr) {
        void *ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return RVSP_ERROR_UNSUPPORTED_BACKEND;
        slot.scratch_ptr = (char *)ptr;
    }
    rvsp_hip_plan *p = &slot.plan;
    p->stream = (hipStream_t)stream;
    p->own_stream = false;
    p->arena = slot.scratch_ptr;
    p->arena_cap = RVSP_NOALLOC_SCRATCH_BYTES;
    p->single_stream = true;
    int64_t nnz = 0, products = 0;
    // heuristic row bins: the learned selectors upload their weights on
    // first use, which allocates
    s = rvsp_hip_plan_analyse(p, A->rows, A->cols, B->cols, A->nnz, A->row_ptr, A->col_idx, B->nnz,
                              B->row_ptr, B->col_idx, RVSP_SELECT_HEURISTIC, &nnz, &products);
    if (s == RVSP_SUCCESS && nnz > C->nnz) {
        // too small an output: report the size needed
        C->nnz = (int32_t)nnz;
        s = RVSP_ERROR_INVALID_ARGUMENT;
    }
    if (s == RVSP_SUCCESS) {
        const hipError_t e = hipMemcpyAsync(C->row_ptr, p->crp, ((size_t)A->rows + 1) * 4,
                                            hipMemcpyDeviceToDevice, p->stream);
        s = e == hipSuccess ? RVSP_SUCCESS : map_err(e);
    }
    if (s == RVSP_SUCCESS)
        s = rvsp_hip_plan_fill_numeric(p, (const float *)A->values, (const float *)B->values,
                                       C->col_idx, (float *)C->values);
    // also waits for the queued work when something above failed
    const rvsp_status_t chk = rvsp_hip_plan_check(p);
    if (s == RVSP_SUCCESS) s = chk;
    plan_release(p);
    if (s != RVSP_SUCCESS) return s;
    C->rows = A->rows;
    C->cols = B->cols;
    C->nnz = (int32_t)nnz;
    C->dtype = RVSP_DTYPE_FP32;
    C->format = RVSP_FORMAT_CSR;
    C->owns_data = 0;
    return RVSP_SUCCESS;
}
