// rvsp_numeric.hip -- numeric pass of C = A * B (FP32) on gfx950.
//
// Arithmetic contract (rv-sparse gustavson_core_f32.inc:31-63 with
// accum_scalar_f32.c:67-70, built without FMA contraction): every C entry
// starts at +0.0f and receives fl(a_ik * b_kj) added with one rounding, in
// ascending A-row position (ascending k) order; C entries are drained in
// C's stored column order. The kernels below keep exactly that order, so
// the values are bit-identical to the scalar kernel:
//   * one row (or one column window of a row) belongs to lanes of ONE wave;
//   * the A entries of the row are walked in stored order, and for each
//     the lanes cover B_k's entries, whose columns are distinct;
//   * each contribution is one LDS add (a read-modify-write, lds_add, where
//     that is faster than the LDS atomic: no lane of a k-step shares a slot
//     with another), and a wave's LDS operations execute in issue order, so
//     successive k-steps reach an accumulator slot in k order. A compiler
//     fence between k-steps keeps the issue order equal to program order.
// Rows are binned (exact nnz(C_i) from the symbolic pass) to the accumulator
// with the smallest LDS footprint (or the learned chooser's pick):
//   hash    nnz(C_i) <= 2730: LDS table of H >= 1.5 nnz slots, pre-loaded
//           with C_i's columns, G lanes per row (G ~ mean |B_k|)
//   dense   an LDS window of W >= span columns indexed by column
//   rank    span <= 16384 and nnz(C_i) <= 4096: a bitmap of C_i's columns
//           with a popcount prefix maps a column to its output position
//   big     rows the symbolic pass kept a bitmap for (rvsp_bigrow.hip)
#include <stdlib.h>

#include <mutex>

#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

// Walks one row's A entries in stored order with G lanes over each B_k and
// calls add(col, fl(a*b)) for every product. Loads for four k-steps are
// issued before any of their adds, so up to eight loads per lane are in
// flight; the adds of each k-step are then issued in k order, separated by
// compiler fences (B is read-only, so hoisting its loads is safe); B is read
// as (column, value) pairs, one 8-byte load per entry. A B_k longer than G
// entries is finished kTailBatches chunks at a time, their loads issued
// together, so a long B row costs one round trip per kTailBatches * G
// entries rather than per G. The
// entry of a k-step reaches the group's lanes through a scalar register
// (G = 64: one row per wave) or through the group's LDS table M of G
// entries (one 16-byte load per k-step instead of lane permutes).
constexpr int kTailBatches = 4;

template <int G, typename Add>
__device__ __forceinline__ void walk_row_ordered(const NumArgs &a, int as, int ae, KMeta *M,
                                                 Add add)
{
    const int lane = threadIdx.x % G;
    for (int base = as; base < ae; base += G) {
        const int ap = base + lane;
        float av = 0.0f;
        int bs = 0, be = 0;
        if (ap < ae) {
            const int k = a.aci[ap];
            av = a.aval[ap];
            bs = a.brp[k];
            be = a.brp[k + 1];
        }
        if constexpr (G < 64) {
            M[lane] = KMeta{av, bs, be, 0};
            wave_sync();
        }
        const int nch = min(G, ae - base);
        for (int j0 = 0; j0 < nch; j0 += 4) {
            float au[4], vu[4];
            int su[4], eu[4];
            int32_t cu[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(j0 + u, G - 1);
                if constexpr (G == 64) {  // one row per wave: j is wave-uniform
                    au[u] = lane_bcast(av, j);
                    su[u] = lane_bcast(bs, j);
                    eu[u] = lane_bcast(be, j);
                } else {
                    const KMeta e = M[j];
                    au[u] = e.av;
                    su[u] = e.bs;
                    eu[u] = e.be;
                }
                const int bp = su[u] + lane;
                const bool ok = j0 + u < nch && bp < eu[u];
                const int2 cv = ok ? a.bcv[bp] : make_int2(-1, 0);
                cu[u] = cv.x;
                vu[u] = __int_as_float(cv.y);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (j0 + u < nch) {  // uniform within the group
                    if (cu[u] >= 0) add(cu[u], __fmul_rn(au[u], vu[u]));
                    // the rest of B_k, kTailBatches chunks of G entries per
                    // round: one k, distinct columns, so any add order works
                    for (int bp0 = su[u] + G; bp0 < eu[u]; bp0 += kTailBatches * G) {
                        int2 tv[kTailBatches];
#pragma unroll
                        for (int t = 0; t < kTailBatches; ++t) {
                            const int bp = bp0 + t * G + lane;
                            tv[t] = bp < eu[u] ? a.bcv[bp] : make_int2(-1, 0);
                        }
#pragma unroll
                        for (int t = 0; t < kTailBatches; ++t)
                            if (tv[t].x >= 0) add(tv[t].x, __fmul_rn(au[u], __int_as_float(tv[t].y)));
                    }
                    compiler_fence();
                }
            }
        }
        if constexpr (G < 64) wave_sync();  // the table is rewritten for the next chunk
    }
}

// Calls fn(q, cci[q]) for every q in [cs, ce) with the G lanes of a row,
// eight loads per lane issued before any call, so reading C_i's columns
// costs one dependent load per 8 G columns instead of one per G.
template <int G, typename Fn>
__device__ __forceinline__ void for_each_c_col(const int32_t *__restrict__ cci, int cs, int ce,
                                               Fn fn)
{
    const int lane = threadIdx.x % G;
    for (int q0 = cs; q0 < ce; q0 += 8 * G) {
        int32_t cc[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * G + lane;
            cc[u] = q < ce ? cci[q] : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * G + lane;
            if (q < ce) fn(q, cc[u]);
        }
    }
}

template <int G, int H, int R>
__global__ __launch_bounds__(R *G) void k_num_hash(const int32_t *__restrict__ rows, int32_t nrows,
                                                   NumArgs a)
{
    __shared__ int32_t keys[R * H];
    __shared__ float vals[R * H];
    __shared__ KMeta meta[G < 64 ? R * G : 1];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    int32_t *T = keys + g * H;
    float *V = vals + g * H;
    for (int s = lane; s < H; s += G) {
        T[s] = -1;
        V[s] = 0.0f;
    }
    __syncthreads();
    const int cs = valid ? a.crp[row] : 0, ce = valid ? a.crp[row + 1] : 0;
    for_each_c_col<G>(a.cci, cs, ce, [&](int, int32_t c) { hash_insert_slot<H>(T, c, a.err); });
    __syncthreads();
    if (valid) {
        int *err = a.err;
        walk_row_ordered<G>(a, a.arp[row], a.arp[row + 1], meta + (G < 64 ? g * G : 0),
                            [&](int32_t col, float prod) {
                                const int s = hash_find<H>(T, col, err);
                                if (s >= 0) lds_add(&V[s], prod);
                            });
    }
    __syncthreads();
    for_each_c_col<G>(a.cci, cs, ce, [&](int q, int32_t c) {
        const int s = hash_find<H>(T, c, a.err);
        a.cval[q] = s >= 0 ? V[s] : 0.0f;
    });
}

// Dense window: G lanes per row, R rows per block, an LDS accumulator of W
// columns over [lo_i, lo_i + W) indexed directly by column. Same ordered
// walk as the hash kernel, no probing; drained through C_i's columns.
template <int G, int W, int R>
__global__ __launch_bounds__(R *G) void k_num_dense(const int32_t *__restrict__ rows, int32_t nrows,
                                                    NumArgs a)
{
    __shared__ float accs[R * W];
    __shared__ KMeta meta[G < 64 ? R * G : 1];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    float *S = accs + g * W;
    for (int i = lane; i < W; i += G) S[i] = 0.0f;
    __syncthreads();
    const int32_t wlo = valid ? a.lo[row] : 0;
    if (valid) {
        int *err = a.err;
        walk_row_ordered<G>(a, a.arp[row], a.arp[row + 1], meta + (G < 64 ? g * G : 0),
                            [&](int32_t col, float prod) {
                                const int32_t c = col - wlo;
                                // the LDS atomic here: dense rows of rv-sparse's
                                // generated inputs walk long B rows in many
                                // batches per k-step, and a read-modify-write
                                // per batch serialises them (generated 16384^2
                                // numeric 7.4 -> 8.7 ms)
                                if ((uint32_t)c < (uint32_t)W)
                                    atomicAdd(&S[c], prod);
                                else
                                    atomicOr(err, kErrWindow);
                            });
    }
    __syncthreads();
    const int cs = valid ? a.crp[row] : 0, ce = valid ? a.crp[row + 1] : 0;
    for_each_c_col<G>(a.cci, cs, ce, [&](int q, int32_t col) {
        const int32_t c = col - wlo;
        a.cval[q] = (uint32_t)c < (uint32_t)W ? S[c] : 0.0f;
    });
}

// Rank window: G lanes per row (32 or 64), R rows per block. The row's
// output columns (C_i's structure) are set in an LDS bitmap over its product
// span, each word gets the popcount of the words before it, and a product
// at column c adds into acc[rank(c)] (two LDS reads and a popcount). The
// accumulator holds CAP >= nnz(C_i) output positions, so it drains straight
// into C_i's values. Same ordered walk as the hash kernel.
template <int G, int CAP, int R>
__global__ __launch_bounds__(R *G) void k_num_rankw(const int32_t *__restrict__ rows, int32_t nrows,
                                                    NumArgs a)
{
    constexpr int WPL = kRankWords / G;  // bitmap words per lane in the prefix
    __shared__ uint32_t Wd[R * kRankWords];
    __shared__ uint16_t Pd[R * kRankWords];
    __shared__ float accs[R * CAP];
    __shared__ KMeta meta[G < 64 ? R * G : 1];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    uint32_t *Wr = Wd + g * kRankWords;
    uint16_t *Pr = Pd + g * kRankWords;
    float *S = accs + g * CAP;
    const int cs = valid ? a.crp[row] : 0;
    const int nz = valid ? min(a.crp[row + 1] - cs, CAP) : 0;
    const int32_t wlo = valid ? a.lo[row] : 0;
    for (int i = lane; i < kRankWords; i += G) Wr[i] = 0u;
    for (int i = lane; i < nz; i += G) S[i] = 0.0f;
    __syncthreads();
    for_each_c_col<G>(a.cci, cs, cs + nz, [&](int, int32_t col) {
        const int32_t c = col - wlo;
        if ((uint32_t)c < (uint32_t)kRankMaxCols)
            atomicOr(&Wr[c >> 5], 1u << (c & 31));
        else
            atomicOr(a.err, kErrWindow);
    });
    __syncthreads();
    {
        int local = 0;
#pragma unroll
        for (int j = 0; j < WPL; ++j) local += __popc(Wr[lane * WPL + j]);
        int incl = local;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int y = __shfl_up(incl, o, G);
            if (lane >= o) incl += y;
        }
        int run = incl - local;
#pragma unroll
        for (int j = 0; j < WPL; ++j) {
            Pr[lane * WPL + j] = (uint16_t)run;
            run += __popc(Wr[lane * WPL + j]);
        }
    }
    __syncthreads();
    if (valid) {
        int *err = a.err;
        walk_row_ordered<G>(a, a.arp[row], a.arp[row + 1], meta + (G < 64 ? g * G : 0),
                            [&](int32_t col, float prod) {
                                const int32_t c = col - wlo;
                                if ((uint32_t)c < (uint32_t)kRankMaxCols) {
                                    const int wi = c >> 5;
                                    const uint32_t word = Wr[wi];
                                    const int pos =
                                        (int)Pr[wi] + __popc(word & ((1u << (c & 31)) - 1u));
                                    if (((word >> (c & 31)) & 1u) && pos < nz)
                                        lds_add(&S[pos], prod);
                                    else
                                        atomicOr(err, kErrProbeMiss);
                                } else {
                                    atomicOr(err, kErrWindow);
                                }
                            });
    }
    __syncthreads();
    for (int q = lane; q < nz; q += G) a.cval[cs + q] = S[q];
}

// Rows per block for a kernel needing `rowb` LDS bytes per row: the count
// (whole waves, at most 256 threads) that keeps the most rows resident in a
// CU (its 160 KB of LDS, and at most 32 waves), among blocks of at most
// 64 KB (the larger count on a tie); rows too big for that take the
// smallest count that fits a CU.
constexpr int kLdsPerCu = 163840;
constexpr int kWavesPerCu = 32;

template <int G>
constexpr int rows_for_lds(int rowb)
{
    for (int pass = 0; pass < 2; ++pass) {
        const int cap = pass == 0 ? 65536 : kLdsPerCu;
        int best = 0, best_rows = 0;
        for (int r = 1; r * G <= 256; ++r) {
            if ((r * G) % 64 != 0 || r * rowb > cap) continue;
            const int by_lds = (kLdsPerCu / (r * rowb)) * r;
            const int by_waves = kWavesPerCu * 64 / G;
            const int resident = by_lds < by_waves ? by_lds : by_waves;
            if (resident > best_rows || (pass == 0 && resident == best_rows)) {
                best = r;
                best_rows = resident;
            }
        }
        if (best > 0) return best;
    }
    return 64 / G;
}

template <int G, int CAP>
static void launch_rankw(hipStream_t st, const NumArgs &a, const int32_t *rows, int nrows)
{
    constexpr int ROWB = kRankWords * 6 + CAP * 4 + (G < 64 ? G * 16 : 0);
    constexpr int R = rows_for_lds<G>(ROWB);
    static_assert((R * G) % 64 == 0, "block must be whole waves");
    k_num_rankw<G, CAP, R><<<(nrows + R - 1) / R, R * G, 0, st>>>(rows, nrows, a);
}

template <int G, int H>
static void launch_hash(hipStream_t st, const NumArgs &a, const int32_t *rows, int nrows)
{
    constexpr int R = rows_for_lds<G>(8 * H + (G < 64 ? G * 16 : 0));
    static_assert((R * G) % 64 == 0, "block must be whole waves");
    const int grid = (nrows + R - 1) / R;
    k_num_hash<G, H, R><<<grid, R * G, 0, st>>>(rows, nrows, a);
}

template <int G, int W>
static void launch_dense(hipStream_t st, const NumArgs &a, const int32_t *rows, int nrows)
{
    constexpr int R = rows_for_lds<G>(4 * W + (G < 64 ? G * 16 : 0));
    static_assert((R * G) % 64 == 0, "block must be whole waves");
    const int grid = (nrows + R - 1) / R;
    k_num_dense<G, W, R><<<grid, R * G, 0, st>>>(rows, nrows, a);
}

#define RVSP_HASH_CASE(HI, GI, H, G)                                                            \
    case NB_HASH0 + 4 * (HI) + (GI):                                                            \
        launch_hash<G, H>(st, a, rows, nr);                                                     \
        return true;
#define RVSP_RANK_CASE(CI, CAP, G)                                                              \
    case NB_RANK0 + 2 * (CI) + ((G) == 64 ? 1 : 0):                                             \
        launch_rankw<G, CAP>(st, a, rows, nr);                                                  \
        return true;
#define RVSP_DENSE_CASE(WI, GI, W, G)                                                           \
    case NB_DENSE0 + 4 * (WI) + (GI):                                                           \
        launch_dense<G, W>(st, a, rows, nr);                                                    \
        return true;

bool launch_num_bin(hipStream_t st, const NumArgs &a, int bin, const int32_t *rows, int nr)
{
    if (nr <= 0) return true;
    switch (bin) {
        RVSP_HASH_CASE(0, 0, 64, 8)
        RVSP_HASH_CASE(0, 1, 64, 16)
        RVSP_HASH_CASE(0, 2, 64, 32)
        RVSP_HASH_CASE(0, 3, 64, 64)
        RVSP_HASH_CASE(1, 0, 128, 8)
        RVSP_HASH_CASE(1, 1, 128, 16)
        RVSP_HASH_CASE(1, 2, 128, 32)
        RVSP_HASH_CASE(1, 3, 128, 64)
        RVSP_HASH_CASE(2, 0, 256, 8)
        RVSP_HASH_CASE(2, 1, 256, 16)
        RVSP_HASH_CASE(2, 2, 256, 32)
        RVSP_HASH_CASE(2, 3, 256, 64)
        RVSP_HASH_CASE(3, 0, 512, 8)
        RVSP_HASH_CASE(3, 1, 512, 16)
        RVSP_HASH_CASE(3, 2, 512, 32)
        RVSP_HASH_CASE(3, 3, 512, 64)
        RVSP_HASH_CASE(4, 0, 1024, 8)
        RVSP_HASH_CASE(4, 1, 1024, 16)
        RVSP_HASH_CASE(4, 2, 1024, 32)
        RVSP_HASH_CASE(4, 3, 1024, 64)
        RVSP_HASH_CASE(5, 1, 2048, 16)
        RVSP_HASH_CASE(5, 2, 2048, 32)
        RVSP_HASH_CASE(5, 3, 2048, 64)
        RVSP_HASH_CASE(6, 2, 4096, 32)
        RVSP_HASH_CASE(6, 3, 4096, 64)
        RVSP_DENSE_CASE(0, 0, 512, 8)
        RVSP_DENSE_CASE(0, 1, 512, 16)
        RVSP_DENSE_CASE(0, 2, 512, 32)
        RVSP_DENSE_CASE(0, 3, 512, 64)
        RVSP_DENSE_CASE(1, 0, 1024, 8)
        RVSP_DENSE_CASE(1, 1, 1024, 16)
        RVSP_DENSE_CASE(1, 2, 1024, 32)
        RVSP_DENSE_CASE(1, 3, 1024, 64)
        RVSP_DENSE_CASE(2, 0, 2048, 8)
        RVSP_DENSE_CASE(2, 1, 2048, 16)
        RVSP_DENSE_CASE(2, 2, 2048, 32)
        RVSP_DENSE_CASE(2, 3, 2048, 64)
        RVSP_DENSE_CASE(3, 1, 4096, 16)
        RVSP_DENSE_CASE(3, 2, 4096, 32)
        RVSP_DENSE_CASE(3, 3, 4096, 64)
        RVSP_DENSE_CASE(4, 2, 8192, 32)
        RVSP_DENSE_CASE(4, 3, 8192, 64)
        RVSP_DENSE_CASE(5, 3, 16384, 64)
        RVSP_RANK_CASE(0, 512, 32)
        RVSP_RANK_CASE(0, 512, 64)
        RVSP_RANK_CASE(1, 1024, 32)
        RVSP_RANK_CASE(1, 1024, 64)
        RVSP_RANK_CASE(2, 2048, 32)
        RVSP_RANK_CASE(2, 2048, 64)
        RVSP_RANK_CASE(3, 4096, 32)
        RVSP_RANK_CASE(3, 4096, 64)
    default:
        return false;
    }
}

__global__ __launch_bounds__(256) void k_interleave(int32_t nnz, const int32_t *__restrict__ bci,
                                                    const float *__restrict__ bval,
                                                    int2 *__restrict__ bcv)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nnz) bcv[i] = make_int2(bci[i], __float_as_int(bval[i]));
}

void launch_interleave(hipStream_t st, int32_t nnz, const int32_t *bci, const float *bval,
                       int2 *bcv)
{
    if (nnz > 0) k_interleave<<<(int)(((int64_t)nnz + 255) / 256), 256, 0, st>>>(nnz, bci, bval, bcv);
}

namespace {
constexpr int kFanDevices = 16;
struct FanPool {
    std::mutex mu;
    bool tried = false, ok = false;
    hipStream_t side[3] = {};
    hipEvent_t fork = nullptr, join[3] = {};
};
FanPool g_fan[kFanDevices];

int fan_streams()
{
    static const int n = [] {
        const char *e = getenv("RVSP_NUM_STREAMS");
        const int v = e ? atoi(e) : 4;
        return v < 1 ? 1 : v > 4 ? 4 : v;
    }();
    return n;
}
}  // namespace

StreamFan::StreamFan(hipStream_t st, bool want) : main_(st)
{
    const int nwant = want ? fan_streams() : 1;
    if (nwant <= 1) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kFanDevices) return;
    FanPool &p = g_fan[dev];
    p.mu.lock();
    if (!p.tried) {
        p.tried = true;
        bool good = hipEventCreateWithFlags(&p.fork, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 3 && good; ++i)
            good = hipStreamCreateWithFlags(&p.side[i], hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&p.join[i], hipEventDisableTiming) == hipSuccess;
        p.ok = good;
    }
    int k = 0;
    if (p.ok && hipEventRecord(p.fork, st) == hipSuccess) {
        for (int i = 0; i < nwant - 1; ++i) {
            if (hipStreamWaitEvent(p.side[i], p.fork, 0) != hipSuccess) break;
            side_[k] = p.side[i];
            join_[k] = p.join[i];
            ++k;
        }
    }
    if (k == 0) {
        p.mu.unlock();
        return;
    }
    n_ = 1 + k;
    dev_ = dev;
}

hipStream_t StreamFan::take(unsigned long long work)
{
    int best = 0;
    for (int i = 1; i < n_; ++i)
        if (load_[i] < load_[best]) best = i;
    load_[best] += work;
    return best == 0 ? main_ : side_[best - 1];
}

int StreamFan::assign(int lo, int hi, const int32_t *count, const unsigned long long *work,
                      int *order, hipStream_t *sts)
{
    // Behind big rows (preloaded work) the bins go over the side streams in
    // bin order: the big rows outlast them whatever the spread, and largest
    // first queued many short launches behind the long bins, past the end
    // of the big rows (generated 32768^2 numeric 5.45 -> 6.0-6.2 ms). With
    // no big rows the pass is the bins alone: largest first onto the least
    // loaded stream, the caller's included (generated 4096^2 numeric 0.77
    // -> 0.55 ms, stencil reuse per call 0.302 -> 0.280 ms).
    const bool lpt = work && load_[0] == 0 && n_ > 1;  // one stream: bin order
    const int n = bins_by_work(lo, hi, count, lpt ? work : nullptr, order);
    for (int i = 0; i < n; ++i) sts[i] = lpt ? take(work[order[i]]) : pick(i);
    return n;
}

int bins_by_work(int lo, int hi, const int32_t *count, const unsigned long long *work, int *order)
{
    int n = 0;
    for (int b = lo; b < hi; ++b) {
        if (count[b] <= 0) continue;
        int j = n++;
        // insertion by descending work (stable: equal work keeps bin order)
        while (work && j > 0 && work[order[j - 1]] < work[b]) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = b;
    }
    return n;
}

StreamFan::~StreamFan()
{
    if (dev_ < 0) return;
    for (int i = 0; i < n_ - 1; ++i) {
        // a join that cannot be queued is made on the host instead
        if (hipEventRecord(join_[i], side_[i]) != hipSuccess ||
            hipStreamWaitEvent(main_, join_[i], 0) != hipSuccess)
            (void)hipStreamSynchronize(side_[i]);
    }
    g_fan[dev_].mu.unlock();
}

void launch_numeric(hipStream_t st, const NumArgs &a, const int32_t *perm, const int32_t *count,
                    const int32_t *off, const BigArgs &big, int nsegs, bool fan,
                    const unsigned long long *work)
{
    StreamFan f(st, fan);
    launch_num_big(st, nsegs, a, big);  // the longest work first, on the caller's stream
    if (work) f.preload(work[NB_WIN]);
    int order[NB_COUNT];
    hipStream_t sts[NB_COUNT];
    const int n = f.assign(NB_HASH0, NB_WIN, count, work, order, sts);
    for (int i = 0; i < n; ++i)
        (void)launch_num_bin(sts[i], a, order[i], perm + off[order[i]], count[order[i]]);
}

} // namespace rvsp
