// rvsp_selector.hip -- the learned row selectors: two small MLPs with two
// ReLU hidden layers of 16, evaluated and trained by the kernels here.
//
// Per-row features (products ops = sum of |B_k| over A_i, A-row length
// alen, product column span, and, once the symbolic pass has run, the
// exact nnz(C_i)):
//   f0 = log2(1 + alen) / 16      f1 = log2(1 + ops) / 32
//   f2 = log2(1 + span) / 32      f3 = log2(1 + ops / alen) / 16
//   f4 = log2(1 + nnz) / 32
//
// SIZER (f0..f3 -> 1): y ~ log2(nnz(C_i) / u_i), u_i = min(ops_i, span_i),
//   trained on the exact counts the symbolic pass produces. The analysis
//   gives row i a symbolic hash table of H >= 2 u_i 2^y margin slots when
//   that is smaller than what the heuristic bins give it; a row whose exact
//   count then exceeds H/2 is recounted in its heuristic bin, so a
//   misprediction costs time only, never a different result.
// CHOOSER (f0..f4 -> 8): y_c ~ log2(time per row in ns) of numeric
//   accumulator c = 4*kind + log2(G)-3 (kind 0: LDS hash table sized from
//   nnz, kind 1: dense LDS window sized from the span; G lanes per row),
//   trained on kernel timings of rows grouped by their features. The
//   numeric binning takes the feasible accumulator with the smallest
//   predicted time. Every accumulator gives bit-identical values.
//
// Training: blocks of TB rows run forward and backward (thread = row) with
// activations in LDS, then sum each parameter's gradient over the block's
// rows in a fixed order (thread = parameter); one more kernel adds the
// block partials in block order and takes an Adam step. Deterministic for
// a given seed.
#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <math.h>
#include <mutex>
#include <new>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

#include "rv_sparse_hip.h"
#include "rvsp_device.hpp"
#include "rvsp_hip_abi.h"
#include "rvsp_kernels.hpp"
#include "rvsp_mlp.hpp"

namespace rvsp {

constexpr int TB = 128;  // samples per training block

// Features of every row (m x 5) and the sizer target (NaN for rows
// without products).
__global__ __launch_bounds__(256) void k_sel_features(int32_t m, const int64_t *__restrict__ ops,
                                                      const int32_t *__restrict__ arp,
                                                      const int32_t *__restrict__ lo,
                                                      const int32_t *__restrict__ hi,
                                                      const int32_t *__restrict__ crp,
                                                      float *__restrict__ feats,
                                                      float *__restrict__ ratio)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t span = ops[i] ? (int64_t)hi[i] - lo[i] + 1 : 0;
    const int nz = crp[i + 1] - crp[i];
    float f[kFeat];
    row_features(ops[i], arp[i + 1] - arp[i], span, nz, f);
#pragma unroll
    for (int j = 0; j < kFeat; ++j) feats[i * kFeat + j] = f[j];
    const int64_t u = ops[i] < span ? ops[i] : span;
    ratio[i] = ops[i] ? log2f((float)max(nz, 1) / (float)u) : __int_as_float(0x7fc00000);
}

// Forward + backward over one block of samples: per-block gradient
// partials (NP entries), then the block's squared-error sum and its number
// of measured targets. A NaN target is unmeasured: no loss, no gradient.
template <class Net, int NF, int NO>
__global__ __launch_bounds__(TB) void k_sel_grad(const float *__restrict__ params,
                                                 const float *__restrict__ feats,
                                                 const float *__restrict__ target,
                                                 const int32_t *__restrict__ idx, int32_t n,
                                                 float inv_n, float *__restrict__ partial)
{
    constexpr int NP = Net::NP, H1 = Net::H1, H2 = Net::H2;
    __shared__ float P[NP];
    __shared__ float sF[TB][NF], sH1[TB][H1], sH2[TB][H2], sD1[TB][H1], sD2[TB][H2], sDY[TB][NO];
    __shared__ float sL[TB], sC[TB];
    const int t = threadIdx.x;
    for (int p = t; p < NP; p += TB) P[p] = params[p];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * TB + t;
    const bool valid = r < n;
    float f[NF], tg[NO];
#pragma unroll
    for (int j = 0; j < NF; ++j) f[j] = 0.0f;
#pragma unroll
    for (int o = 0; o < NO; ++o) tg[o] = __int_as_float(0x7fc00000);
    if (valid) {
        const int64_t s = idx ? idx[r] : r;
#pragma unroll
        for (int j = 0; j < NF; ++j) f[j] = feats[s * kFeat + j];
#pragma unroll
        for (int o = 0; o < NO; ++o) tg[o] = target[s * NO + o];
    }
    float z1[H1], h1[H1], z2[H2], h2[H2], y[NO], dy[NO];
    Net::forward(P, f, z1, h1, z2, h2, y);
    float lsum = 0.0f, lcnt = 0.0f;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        const bool meas = tg[o] == tg[o];  // false for NaN
        const float err = meas ? y[o] - tg[o] : 0.0f;
        dy[o] = 2.0f * err * inv_n;  // d(inv_n * sum of squared errors)/dy
        lsum += err * err;
        lcnt += meas ? 1.0f : 0.0f;
    }
    float d2[H2], d1[H1];
#pragma unroll
    for (int i = 0; i < H2; ++i) {
        float g = 0.0f;
#pragma unroll
        for (int o = 0; o < NO; ++o) g += P[Net::W3 + o * H2 + i] * dy[o];
        d2[i] = z2[i] > 0.0f ? g : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < H1; ++j) {
        float g = 0.0f;
#pragma unroll
        for (int i = 0; i < H2; ++i) g += P[Net::W2 + i * H1 + j] * d2[i];
        d1[j] = z1[j] > 0.0f ? g : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < NF; ++j) sF[t][j] = f[j];
#pragma unroll
    for (int i = 0; i < H1; ++i) {
        sH1[t][i] = h1[i];
        sD1[t][i] = d1[i];
    }
#pragma unroll
    for (int i = 0; i < H2; ++i) {
        sH2[t][i] = h2[i];
        sD2[t][i] = d2[i];
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) sDY[t][o] = dy[o];
    sL[t] = lsum;
    sC[t] = lcnt;
    __syncthreads();
    // thread = parameter: its gradient summed over the block's samples in order
    for (int p = t; p < NP + 2; p += TB) {
        float g = 0.0f;
        if (p < Net::B1) {
            const int i = (p - Net::W1) / NF, j = (p - Net::W1) % NF;
            for (int s = 0; s < TB; ++s) g += sD1[s][i] * sF[s][j];
        } else if (p < Net::W2) {
            const int i = p - Net::B1;
            for (int s = 0; s < TB; ++s) g += sD1[s][i];
        } else if (p < Net::B2) {
            const int i = (p - Net::W2) / H1, j = (p - Net::W2) % H1;
            for (int s = 0; s < TB; ++s) g += sD2[s][i] * sH1[s][j];
        } else if (p < Net::W3) {
            const int i = p - Net::B2;
            for (int s = 0; s < TB; ++s) g += sD2[s][i];
        } else if (p < Net::B3) {
            const int o = (p - Net::W3) / H2, i = (p - Net::W3) % H2;
            for (int s = 0; s < TB; ++s) g += sDY[s][o] * sH2[s][i];
        } else if (p < NP) {
            const int o = p - Net::B3;
            for (int s = 0; s < TB; ++s) g += sDY[s][o];
        } else if (p == NP) {
            for (int s = 0; s < TB; ++s) g += sL[s];
        } else {
            for (int s = 0; s < TB; ++s) g += sC[s];
        }
        partial[(int64_t)blockIdx.x * (NP + 2) + p] = g;
    }
}

// Adds the block partials in block order into grad[NP + 2]; with apply,
// takes one Adam step on params.
template <int NP>
__global__ __launch_bounds__(512) void k_sel_reduce_adam(const float *__restrict__ partial,
                                                         int nblocks, float *__restrict__ grad,
                                                         float *__restrict__ params,
                                                         float *__restrict__ m1,
                                                         float *__restrict__ m2, float lr,
                                                         float b1t, float b2t, int apply)
{
    for (int p = threadIdx.x; p < NP + 2; p += 512) {
        float g = 0.0f;
        for (int b = 0; b < nblocks; ++b) g += partial[(int64_t)b * (NP + 2) + p];
        grad[p] = g;
        if (!apply || p >= NP) continue;
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        m1[p] = beta1 * m1[p] + (1.0f - beta1) * g;
        m2[p] = beta2 * m2[p] + (1.0f - beta2) * g * g;
        const float mh = m1[p] / (1.0f - b1t), vh = m2[p] / (1.0f - b2t);
        params[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

template <class Net, int NF, int NO>
__global__ __launch_bounds__(256) void k_sel_predict(const float *__restrict__ params,
                                                     const float *__restrict__ feats, int32_t n,
                                                     float *__restrict__ pred)
{
    __shared__ float P[Net::NP];
    for (int p = threadIdx.x; p < Net::NP; p += 256) P[p] = params[p];
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    float f[NF], z1[16], h1[16], z2[16], h2[16], y[NO];
#pragma unroll
    for (int j = 0; j < NF; ++j) f[j] = feats[r * kFeat + j];
    Net::forward(P, f, z1, h1, z2, h2, y);
#pragma unroll
    for (int o = 0; o < NO; ++o) pred[r * NO + o] = y[o];
}

// Learned symbolic bins. Records every row's heuristic bin in hbin, then
// moves a row to an LDS hash table sized from the sizer's prediction when
// that table's footprint is below the heuristic bin's (hash: 4H bytes of
// keys; one-pass bitmap: its fixed 4*kBmsWords bytes; big rows: bitmaps in
// HBM, never cheaper). Keeps the symbolic histogram in step.
__global__ __launch_bounds__(256) void k_sel_bins(int32_t m, const float *__restrict__ params,
                                                  float margin, const int64_t *__restrict__ ops,
                                                  const int32_t *__restrict__ arp,
                                                  const int32_t *__restrict__ lo,
                                                  const int32_t *__restrict__ hi,
                                                  uint8_t *__restrict__ sbin,
                                                  uint8_t *__restrict__ hbin,
                                                  AnalysisCounters *__restrict__ ctr)
{
    __shared__ float P[Sizer::NP];
    __shared__ int hist[SB_COUNT];
    __shared__ int moved;
    for (int p = threadIdx.x; p < Sizer::NP; p += 256) P[p] = params[p];
    if (threadIdx.x < SB_COUNT) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) moved = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < m) {
        const int b = sbin[i];
        hbin[i] = (uint8_t)b;
        if (b != SB_EMPTY) {
            const int64_t span = (int64_t)hi[i] - lo[i] + 1;
            const int64_t u = ops[i] < span ? ops[i] : span;
            float f[kFeat], z1[16], h1[16], z2[16], h2[16], y;
            row_features(ops[i], arp[i + 1] - arp[i], span, 0, f);
            Sizer::forward(P, f, z1, h1, z2, h2, &y);
            const float pred = (float)u * exp2f(fminf(y, 0.0f)) * margin;
            const float want = fminf(2.0f * pred, 2.0f * (float)u);
            if (want <= 4096.0f) {
                const int H = max(64, next_pow2((int)ceilf(fmaxf(want, 1.0f))));
                int64_t cur;
                if (b >= SB_H64 && b <= SB_H4096)
                    cur = 4LL * (64 << (b - SB_H64));
                else if (b == SB_BMS)
                    cur = 4LL * kBmsWords;
                else
                    cur = INT64_MAX;
                if (4LL * H < cur) {
                    const int nb = SB_H64 + (ilog2(H) - 6);
                    sbin[i] = (uint8_t)nb;
                    atomicAdd(&hist[b], -1);
                    atomicAdd(&hist[nb], 1);
                    atomicAdd(&moved, 1);
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < SB_COUNT && hist[threadIdx.x])
        atomicAdd(&ctr->sym_hist[threadIdx.x], hist[threadIdx.x]);
    if (threadIdx.x == 0 && moved) atomicAdd(&ctr->moved, moved);
}

// After a learned count: a row moved by k_sel_bins whose exact count does
// not fit its table (count > H/2; the overflow-checking count kernel
// reports H + 1) returns to its heuristic bin; fbin lists such rows by
// that bin for the recount (SB_EMPTY for every other row).
__global__ __launch_bounds__(256) void k_sel_fallback(int32_t m, const int32_t *__restrict__ cnt,
                                                      uint8_t *__restrict__ sbin,
                                                      const uint8_t *__restrict__ hbin,
                                                      uint8_t *__restrict__ fbin,
                                                      AnalysisCounters *__restrict__ ctr)
{
    __shared__ int hist[SB_COUNT], fh[SB_COUNT];
    __shared__ int back;
    if (threadIdx.x < SB_COUNT) {
        hist[threadIdx.x] = 0;
        fh[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) back = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < m) {
        const int b = sbin[i], h = hbin[i];
        int f = SB_EMPTY;
        if (b != h && b >= SB_H64 && b <= SB_H4096 && cnt[i] > (32 << (b - SB_H64))) {
            sbin[i] = (uint8_t)h;
            f = h;
            atomicAdd(&hist[b], -1);
            atomicAdd(&hist[h], 1);
            atomicAdd(&fh[h], 1);
            atomicAdd(&back, 1);
        }
        fbin[i] = (uint8_t)f;
    }
    __syncthreads();
    if (threadIdx.x == 0 && back) atomicAdd(&ctr->moved, -back);
    if (threadIdx.x < SB_COUNT) {
        if (hist[threadIdx.x]) atomicAdd(&ctr->sym_hist[threadIdx.x], hist[threadIdx.x]);
        if (fh[threadIdx.x]) atomicAdd(&ctr->fb_hist[threadIdx.x], fh[threadIdx.x]);
    }
}

void launch_sel_bins(hipStream_t st, int32_t m, const float *params, float margin,
                     const int64_t *ops, const int32_t *arp, const int32_t *lo, const int32_t *hi,
                     uint8_t *sbin, uint8_t *hbin, AnalysisCounters *ctr)
{
    k_sel_bins<<<(int)(((int64_t)m + 255) / 256), 256, 0, st>>>(m, params, margin, ops, arp, lo,
                                                                  hi, sbin, hbin, ctr);
}

void launch_sel_fallback(hipStream_t st, int32_t m, const int32_t *cnt, uint8_t *sbin,
                         const uint8_t *hbin, uint8_t *fbin, AnalysisCounters *ctr)
{
    k_sel_fallback<<<(int)(((int64_t)m + 255) / 256), 256, 0, st>>>(m, cnt, sbin, hbin, fbin, ctr);
}

void launch_sel_features(hipStream_t st, int32_t m, const int64_t *ops, const int32_t *arp,
                         const int32_t *lo, const int32_t *hi, const int32_t *crp, float *feats,
                         float *ratio)
{
    k_sel_features<<<(int)(((int64_t)m + 255) / 256), 256, 0, st>>>(m, ops, arp, lo, hi, crp,
                                                                      feats, ratio);
}

} // namespace rvsp

using namespace rvsp;

namespace {

struct SelectorSlot {
    int np;
    float *dev = nullptr;
    std::vector<float> host;
    float margin = 1.0f;
    bool ready = false;
    bool tried_default = false;
};

std::mutex g_sel_mu;
SelectorSlot g_slots[2] = {{Sizer::NP}, {Chooser::NP}};

bool valid_which(int which) { return which == RVSP_SELECTOR_SIZER || which == RVSP_SELECTOR_CHOOSER; }

// Selector work runs on the calling thread's stream, never on the legacy
// null stream, whose operations fail while another thread captures a graph.
hipStream_t sel_stream() { return (hipStream_t)rvsp_hip_oneshot_stream(); }
hipError_t sel_malloc(void **p, size_t bytes) { return hipMallocAsync(p, bytes, sel_stream()); }
void sel_free(void *p)
{
    if (p) (void)hipFreeAsync(p, sel_stream());
}
hipError_t sel_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, sel_stream());
    return e == hipSuccess ? hipStreamSynchronize(sel_stream()) : e;
}
hipError_t sel_zero(void *p, size_t bytes) { return hipMemsetAsync(p, 0, bytes, sel_stream()); }

rvsp_status_t slot_set(SelectorSlot &s, const float *params, float margin)
{
    if (!s.dev && hipMalloc((void **)&s.dev, (size_t)s.np * 4) != hipSuccess) {
        s.dev = nullptr;
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    if (sel_copy(s.dev, params, (size_t)s.np * 4, hipMemcpyHostToDevice) != hipSuccess)
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    s.host.assign(params, params + s.np);
    s.margin = margin;
    s.ready = true;
    return RVSP_SUCCESS;
}

// Text file: the margin, then the parameters, whitespace separated.
rvsp_status_t read_params(const char *path, int np, std::vector<float> &p, float &margin)
{
    FILE *f = fopen(path, "r");
    if (!f) return RVSP_ERROR_INVALID_ARGUMENT;
    p.assign((size_t)np, 0.0f);
    int ok = fscanf(f, "%f", &margin) == 1;
    for (int i = 0; ok && i < np; ++i) ok = fscanf(f, "%f", &p[i]) == 1;
    float extra;
    if (ok && fscanf(f, "%f", &extra) == 1) ok = 0;  // too many numbers
    fclose(f);
    return ok && margin > 0.0f ? RVSP_SUCCESS : RVSP_ERROR_INVALID_ARGUMENT;
}

const char *const kDefaultFile[2] = {"sizer.txt", "chooser.txt"};
const char *const kEnvFile[2] = {"RVSP_SELECTOR_SIZER", "RVSP_SELECTOR_CHOOSER"};

// The default weights ship beside the library: <libdir>/../selector/.
std::string default_path(int which)
{
    Dl_info info;
    if (!dladdr((void *)&read_params, &info) || !info.dli_fname) return std::string();
    std::string lib(info.dli_fname);
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
    return dir + "/../selector/" + kDefaultFile[which];
}

template <class Net, int NF, int NO>
void launch_grad(const float *dp, const float *feats, const float *target, const int32_t *idx,
                 int n, float *part)
{
    const int nb = (n + TB - 1) / TB;
    k_sel_grad<Net, NF, NO><<<nb, TB, 0, sel_stream()>>>(dp, feats, target, idx, n, 1.0f / (float)n, part);
}

void grad_launch(int which, const float *dp, const float *feats, const float *target,
                 const int32_t *idx, int n, float *part)
{
    if (which == RVSP_SELECTOR_SIZER)
        launch_grad<Sizer, 4, 1>(dp, feats, target, idx, n, part);
    else
        launch_grad<Chooser, 5, 8>(dp, feats, target, idx, n, part);
}

void reduce_launch(int which, const float *part, int nb, float *g, float *dp, float *m1, float *m2,
                   float lr, float b1t, float b2t, int apply)
{
    if (which == RVSP_SELECTOR_SIZER)
        k_sel_reduce_adam<Sizer::NP><<<1, 512, 0, sel_stream()>>>(part, nb, g, dp, m1, m2, lr, b1t, b2t, apply);
    else
        k_sel_reduce_adam<Chooser::NP><<<1, 512, 0, sel_stream()>>>(part, nb, g, dp, m1, m2, lr, b1t, b2t, apply);
}

} // namespace

// Device parameters of a selector for the analysis, or NULL when none is
// loaded: the set or loaded ones, else (tried once per process) the file
// named by RVSP_SELECTOR_SIZER / RVSP_SELECTOR_CHOOSER, else the default
// file shipped beside the library.
const float *rvsp::selector_device_params(int which, float *margin)
{
    if (!valid_which(which)) return nullptr;
    std::lock_guard<std::mutex> lk(g_sel_mu);
    SelectorSlot &s = g_slots[which];
    if (!s.ready && !s.tried_default) {
        s.tried_default = true;
        const char *env = getenv(kEnvFile[which]);
        const std::string path = env ? std::string(env) : default_path(which);
        std::vector<float> p;
        float mg = 0.0f;
        if (!path.empty() && read_params(path.c_str(), s.np, p, mg) == RVSP_SUCCESS)
            (void)slot_set(s, p.data(), mg);
    }
    if (!s.ready) return nullptr;
    *margin = s.margin;
    return s.dev;
}

extern "C" {

int32_t rvsp_selector_param_count(int32_t which)
{
    return valid_which(which) ? g_slots[which].np : -1;
}

rvsp_status_t rvsp_selector_set(int32_t which, const float *params, int32_t n, float margin)
{
    if (!params) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which) || n != g_slots[which].np || !(margin > 0.0f))
        return RVSP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(g_sel_mu);
    return slot_set(g_slots[which], params, margin);
}

rvsp_status_t rvsp_selector_clear(int32_t which)
{
    if (!valid_which(which)) return RVSP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(g_sel_mu);
    g_slots[which].ready = false;
    g_slots[which].tried_default = true;  // stays off until set or loaded again
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_selector_get(int32_t which, float *params, int32_t n, float *margin)
{
    if (!params || !margin) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which) || n != g_slots[which].np) return RVSP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(g_sel_mu);
    const SelectorSlot &s = g_slots[which];
    if (!s.ready) return RVSP_ERROR_INVALID_ARGUMENT;
    memcpy(params, s.host.data(), (size_t)s.np * 4);
    *margin = s.margin;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_selector_load_file(int32_t which, const char *path)
{
    if (!path) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which)) return RVSP_ERROR_INVALID_ARGUMENT;
    std::vector<float> p;
    float margin = 0.0f;
    rvsp_status_t s = read_params(path, g_slots[which].np, p, margin);
    if (s != RVSP_SUCCESS) return s;
    std::lock_guard<std::mutex> lk(g_sel_mu);
    return slot_set(g_slots[which], p.data(), margin);
}

rvsp_status_t rvsp_selector_predict(int32_t which, const float *feats, int32_t n,
                                    const float *params_host, float *pred)
{
    if (!feats || !params_host || !pred) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which) || n < 0) return RVSP_ERROR_INVALID_ARGUMENT;
    if (n == 0) return RVSP_SUCCESS;
    const int np = g_slots[which].np;
    float *dp = nullptr;
    if (sel_malloc((void **)&dp, (size_t)np * 4) != hipSuccess) return RVSP_ERROR_ALLOCATION_FAILED;
    rvsp_status_t s = RVSP_SUCCESS;
    if (sel_copy(dp, params_host, (size_t)np * 4, hipMemcpyHostToDevice) != hipSuccess)
        s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    if (s == RVSP_SUCCESS) {
        const int grid = (n + 255) / 256;
        if (which == RVSP_SELECTOR_SIZER)
            k_sel_predict<Sizer, 4, 1><<<grid, 256, 0, sel_stream()>>>(dp, feats, n, pred);
        else
            k_sel_predict<Chooser, 5, 8><<<grid, 256, 0, sel_stream()>>>(dp, feats, n, pred);
        if (hipStreamSynchronize(sel_stream()) != hipSuccess) s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    sel_free(dp);
    return s;
}

// Gradient, at params_host, of (sum over samples and measured outputs of
// the squared error) / n, and that loss.
rvsp_status_t rvsp_selector_grad(int32_t which, const float *feats, const float *target, int32_t n,
                                 const float *params_host, float *grad_host, float *loss)
{
    if (!feats || !target || !params_host || !grad_host) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which) || n <= 0) return RVSP_ERROR_INVALID_ARGUMENT;
    const int np = g_slots[which].np;
    const int nb = (n + TB - 1) / TB;
    float *dp = nullptr, *part = nullptr, *g = nullptr;
    rvsp_status_t s = RVSP_SUCCESS;
    if (sel_malloc((void **)&dp, (size_t)np * 4) != hipSuccess ||
        sel_malloc((void **)&part, (size_t)nb * (np + 2) * 4) != hipSuccess ||
        sel_malloc((void **)&g, (size_t)(np + 2) * 4) != hipSuccess)
        s = RVSP_ERROR_ALLOCATION_FAILED;
    if (s == RVSP_SUCCESS &&
        sel_copy(dp, params_host, (size_t)np * 4, hipMemcpyHostToDevice) != hipSuccess)
        s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    if (s == RVSP_SUCCESS) {
        grad_launch(which, dp, feats, target, nullptr, n, part);
        reduce_launch(which, part, nb, g, nullptr, nullptr, nullptr, 0.0f, 0.0f, 0.0f, 0);
        std::vector<float> hg((size_t)np + 2);
        if (sel_copy(hg.data(), g, hg.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
            s = RVSP_ERROR_UNSUPPORTED_BACKEND;
        } else {
            memcpy(grad_host, hg.data(), (size_t)np * 4);
            if (loss) *loss = hg[np] / (float)n;
        }
    }
    sel_free(dp);
    sel_free(part);
    sel_free(g);
    return s;
}

// Adam on minibatches of `batch` samples, drawn by a seeded shuffle per
// epoch. params_host is the start point and receives the result; loss_out
// gets the last epoch's loss per sample.
rvsp_status_t rvsp_selector_train(int32_t which, const float *feats, const float *target,
                                  int32_t n, int32_t epochs, int32_t batch, float lr,
                                  uint64_t seed, float *params_host, float *loss_out)
{
    if (!feats || !target || !params_host) return RVSP_ERROR_NULL_POINTER;
    if (!valid_which(which) || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f))
        return RVSP_ERROR_INVALID_ARGUMENT;
    const int np = g_slots[which].np;
    const int bsz = ((min(batch, n) + TB - 1) / TB) * TB;
    const int nbmax = bsz / TB;
    float *dp = nullptr, *part = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    int32_t *didx = nullptr;
    rvsp_status_t s = RVSP_SUCCESS;
    if (sel_malloc((void **)&dp, (size_t)np * 4) != hipSuccess ||
        sel_malloc((void **)&part, (size_t)nbmax * (np + 2) * 4) != hipSuccess ||
        sel_malloc((void **)&g, (size_t)(np + 2) * 4) != hipSuccess ||
        sel_malloc((void **)&m1, (size_t)np * 4) != hipSuccess ||
        sel_malloc((void **)&m2, (size_t)np * 4) != hipSuccess ||
        sel_malloc((void **)&didx, (size_t)n * 4) != hipSuccess)
        s = RVSP_ERROR_ALLOCATION_FAILED;
    std::vector<int32_t> perm;
    if (s == RVSP_SUCCESS) {
        perm.resize((size_t)n);
        for (int i = 0; i < n; ++i) perm[i] = i;
        if (sel_copy(dp, params_host, (size_t)np * 4, hipMemcpyHostToDevice) != hipSuccess ||
            sel_zero(m1, (size_t)np * 4) != hipSuccess ||
            sel_zero(m2, (size_t)np * 4) != hipSuccess)
            s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    float b1t = 1.0f, b2t = 1.0f;
    double last = 0.0;
    float lb = 0.0f;
    for (int e = 0; e < epochs && s == RVSP_SUCCESS; ++e) {
        for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with a 64-bit LCG
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            const int j = (int)((st >> 33) % (uint64_t)(i + 1));
            const int32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        if (sel_copy(didx, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) {
            s = RVSP_ERROR_UNSUPPORTED_BACKEND;
            break;
        }
        double lsum = 0.0;
        for (int off = 0; off < n; off += bsz) {
            const int cnt = min(bsz, n - off);
            const int nb = (cnt + TB - 1) / TB;
            b1t *= 0.9f;
            b2t *= 0.999f;
            grad_launch(which, dp, feats, target, didx + off, cnt, part);
            reduce_launch(which, part, nb, g, dp, m1, m2, lr, b1t, b2t, 1);
            if (sel_copy(&lb, g + np, 4, hipMemcpyDeviceToHost) != hipSuccess) {
                s = RVSP_ERROR_UNSUPPORTED_BACKEND;
                break;
            }
            lsum += lb;
        }
        last = lsum / (double)n;
    }
    if (s == RVSP_SUCCESS &&
        sel_copy(params_host, dp, (size_t)np * 4, hipMemcpyDeviceToHost) != hipSuccess)
        s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    if (loss_out) *loss_out = (float)last;
    sel_free(dp);
    sel_free(part);
    sel_free(g);
    sel_free(m1);
    sel_free(m2);
    sel_free(didx);
    return s;
}

} // extern "C"
