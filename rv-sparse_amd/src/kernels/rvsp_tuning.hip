// rvsp_tuning.hip -- the autotuned launch-parameter table.
//
// Parameters that decide how a product's work is binned and scheduled, and
// never what it computes: the symbolic hash limit (rows whose distinct-
// column bound exceeds it leave the LDS hash tables for the one-pass bitmap
// or the big-row path), the big-row numeric window shape, the big-row count
// block size, products per big-row segment, and the work above which a
// pass runs its row bins on side streams. Every setting gives bit-identical
// results; only the time changes.
//
// One entry per workload class, picked on the host before any kernel from
// the operand shapes alone: the estimated product count nnz(A) * nnz(B) / k
// in four size buckets, times whether B is wide enough (n > 16384 columns)
// for big rows. The table is written by the autotuner (rvsp/autotune.py,
// which times every candidate on the GPU) and loaded on first use from the
// file named by RVSP_TUNING_TABLE, else from tuning/table.txt beside the
// library's directory; without one, every class uses the built-in defaults.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <mutex>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>

#include "rv_sparse_hip.h"
#include "rvsp_kernels.hpp"

using namespace rvsp;

namespace {

constexpr int kClasses = RVSP_TUNING_CLASSES;

rvsp_tuning_t defaults()
{
    rvsp_tuning_t t;
    t.big_window_outputs = kBigShapes[0][0];
    t.big_window_chunks = kBigShapes[0][1];
    t.sym_big_threads = 512;
    t.seg_products = kSegProducts;
    t.hash_max_u = kSymHashMaxU;
    t.fan_min_products = kFanMinProducts;
    t.fan_min_fill = kFanMinFill;
    return t;
}

struct Table {
    std::mutex mu;
    bool loaded = false;
    rvsp_tuning_t cls[kClasses];
    bool forced = false;
    rvsp_tuning_t force;
};
Table g_table;

int shape_index(int outs, int chunks)
{
    for (int i = 0; i < kBigShapeCount; ++i)
        if (kBigShapes[i][0] == outs && kBigShapes[i][1] == chunks) return i;
    return -1;
}

bool valid(const rvsp_tuning_t &t)
{
    return shape_index(t.big_window_outputs, t.big_window_chunks) >= 0 &&
           (t.sym_big_threads == 256 || t.sym_big_threads == 512 || t.sym_big_threads == 1024) &&
           t.seg_products >= 1024 && t.seg_products <= (1 << 22) && t.hash_max_u >= 64 &&
           t.hash_max_u <= kSymHashMaxU && t.fan_min_products >= 0 && t.fan_min_fill >= 0;
}

// Text: '#' comments; one line per class:
//   class outputs chunks sym_big_threads seg_products hash_max_u fan_min_products fan_min_fill
rvsp_status_t parse(const char *path, rvsp_tuning_t out[kClasses])
{
    FILE *f = fopen(path, "r");
    if (!f) return RVSP_ERROR_INVALID_ARGUMENT;
    for (int c = 0; c < kClasses; ++c) out[c] = defaults();
    bool seen[kClasses] = {};
    char line[512];
    rvsp_status_t s = RVSP_SUCCESS;
    while (s == RVSP_SUCCESS && fgets(line, sizeof(line), f)) {
        const char *q = line;
        while (*q == ' ' || *q == '\t') ++q;
        if (*q == '#' || *q == '\n' || *q == '\0') continue;
        int c = -1;
        rvsp_tuning_t t;
        long long fp = 0, ff = 0;
        if (sscanf(q, "%d %d %d %d %d %d %lld %lld", &c, &t.big_window_outputs,
                   &t.big_window_chunks, &t.sym_big_threads, &t.seg_products, &t.hash_max_u, &fp,
                   &ff) != 8 ||
            c < 0 || c >= kClasses || seen[c]) {
            s = RVSP_ERROR_INVALID_ARGUMENT;
            break;
        }
        t.fan_min_products = fp;
        t.fan_min_fill = ff;
        if (!valid(t)) {
            s = RVSP_ERROR_INVALID_ARGUMENT;
            break;
        }
        out[c] = t;
        seen[c] = true;
    }
    fclose(f);
    return s;
}

std::string default_path()
{
    Dl_info info;
    if (!dladdr((void *)&parse, &info) || !info.dli_fname) return std::string();
    std::string lib(info.dli_fname);
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../tuning/table.txt";
}

// first use: the file named by RVSP_TUNING_TABLE, else the shipped table,
// else the defaults (caller holds the lock)
void ensure_loaded()
{
    if (g_table.loaded) return;
    g_table.loaded = true;
    for (int c = 0; c < kClasses; ++c) g_table.cls[c] = defaults();
    const char *env = getenv("RVSP_TUNING_TABLE");
    const std::string path = env ? std::string(env) : default_path();
    rvsp_tuning_t t[kClasses];
    if (!path.empty() && parse(path.c_str(), t) == RVSP_SUCCESS)
        memcpy(g_table.cls, t, sizeof(t));
}

} // namespace

namespace rvsp {

TuneParams tune_for(int64_t nnz_a, int64_t nnz_b, int32_t k, int32_t n, bool may_load)
{
    const int c = rvsp_tuning_class(nnz_a, nnz_b, k, n);
    rvsp_tuning_t t;
    {
        std::lock_guard<std::mutex> lk(g_table.mu);
        if (may_load) ensure_loaded();
        t = g_table.forced ? g_table.force : g_table.loaded ? g_table.cls[c] : defaults();
    }
    TuneParams p;
    p.big_window = shape_index(t.big_window_outputs, t.big_window_chunks);
    p.sym_big_threads = t.sym_big_threads;
    p.seg_products = t.seg_products;
    p.hash_max_u = t.hash_max_u;
    p.fan_min_products = t.fan_min_products;
    p.fan_min_fill = t.fan_min_fill;
    return p;
}

} // namespace rvsp

extern "C" {

int32_t rvsp_tuning_class(int64_t nnz_a, int64_t nnz_b, int32_t k, int32_t n)
{
    // estimated products: every A entry meets a B row of mean length
    const double est = k > 0 ? (double)nnz_a * ((double)nnz_b / (double)k) : 0.0;
    const int size = est < (double)(1 << 21) ? 0 : est < (double)(1 << 25) ? 1
                     : est < (double)(1 << 29) ? 2 : 3;
    const int wide = n > kBmsWords * 32 ? 1 : 0;  // columns enough for big rows
    return size * 2 + wide;
}

rvsp_status_t rvsp_tuning_get(int32_t cls, rvsp_tuning_t *out)
{
    if (!out) return RVSP_ERROR_NULL_POINTER;
    if (cls < 0 || cls >= kClasses) return RVSP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(g_table.mu);
    ensure_loaded();
    *out = g_table.cls[cls];
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_tuning_set(int32_t cls, const rvsp_tuning_t *t)
{
    if (!t) return RVSP_ERROR_NULL_POINTER;
    if (cls < 0 || cls >= kClasses || !valid(*t)) return RVSP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(g_table.mu);
    ensure_loaded();
    g_table.cls[cls] = *t;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_tuning_force(const rvsp_tuning_t *t)
{
    std::lock_guard<std::mutex> lk(g_table.mu);
    if (!t) {
        g_table.forced = false;
        return RVSP_SUCCESS;
    }
    if (!valid(*t)) return RVSP_ERROR_INVALID_ARGUMENT;
    g_table.force = *t;
    g_table.forced = true;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_tuning_load_file(const char *path)
{
    if (!path) return RVSP_ERROR_NULL_POINTER;
    rvsp_tuning_t t[kClasses];
    const rvsp_status_t s = parse(path, t);
    if (s != RVSP_SUCCESS) return s;
    std::lock_guard<std::mutex> lk(g_table.mu);
    memcpy(g_table.cls, t, sizeof(t));
    g_table.loaded = true;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_tuning_defaults(rvsp_tuning_t *out)
{
    if (!out) return RVSP_ERROR_NULL_POINTER;
    *out = defaults();
    return RVSP_SUCCESS;
}

} // extern "C"
