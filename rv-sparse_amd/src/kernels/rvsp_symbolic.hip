// rvsp_symbolic.hip -- row analysis, binning, symbolic count/fill and the
// row-pointer scan for C = A * B on gfx950.
//
// Follows the structure semantics of rv-sparse's symbolic phase
// (src/kernels/spgemm/rvsp_symbolic.c:13-138): C_i's columns are the
// distinct columns of the B rows that A_i references, stored ascending;
// op_i = sum over A_i of len(B_k) in int64; row_ptr is the exclusive scan.
// The GPU builds the distinct set in LDS (hash table or bitmap, chosen per
// row) instead of the reference's mark/touched arrays, and sorts in LDS
// (bitonic) or emits in bitmap order, so the result is the same set in the
// same (ascending) order.
#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"
#include "rvsp_mlp.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Per-row statistics: product count, column span of the products, bin.
// 16 lanes per row; lanes stride over the row's A entries.
// ---------------------------------------------------------------------------
// hmax (tuned, at most kSymHashMaxU): rows with u <= hmax use a hash table
// Rows of at least kBigRowOps products take the big-row path whatever their
// span, once A has kBigRowNnz entries: its count spreads a row over a
// 512-thread block and its numeric pass over one wave per window run, where
// the one-pass bitmap and dense kernels give a row one 64-lane group.
// (Generated 16384^2: every row's span fits the one-pass bitmap, and its
// near-dense rows held the count at 3.6 ms and the numeric pass at 7.3 ms:
// 48 -> 157 GFLOP/s at 8192, 131 at 2^17; 2048 moves the 512-product
// batch's rows too, 25.7 -> 28.1 ms. Generated 2048^2 to 8192^2, with 94 k
// entries and more: 1.16x to 1.38x; 512^2 and 1024^2, with 15 k and 41 k
// entries: 0.72x and 0.87x, the big-row path's launches outweighing what it
// spreads.)
// Rows of more candidate columns u than the class's hash limit, too wide
// for the one-pass bitmap, take the big-row path unless their span is at
// most kBigSpanFactor * u (many repeated products over a narrow band) and
// u fits the largest hash table: a 27-point stencil 64^3 row bounds 729
// candidates over about 16.6 k columns (125 outputs) and stays in the hash
// tables, where R-MAT's and the mixed product's rows of that size span
// 2^18 to 2^20 columns and move. (With a hash limit of 512 in their class,
// the stencil held 1300 GFLOP/s with the span test and fell to 333
// without it; at the shipped limits the test moves no measured input.)
constexpr int64_t kBigSpanFactor = 64;
constexpr int64_t kBigRowOps = 8192;
constexpr int32_t kBigRowNnz = 65536;

// big_ops: kBigRowOps, or INT64_MAX below kBigRowNnz entries of A
__device__ __forceinline__ int sym_bin_of(int64_t ops, int32_t lo, int32_t hi, int hmax,
                                          int64_t big_ops)
{
    if (ops == 0) return SB_EMPTY;
    if (ops >= big_ops) return SB_BIG;
    const int64_t span = (int64_t)hi - lo + 1;
    const int64_t u = ops < span ? ops : span;
    int H = 64;
    if (u <= hmax) H = max(64, next_pow2((int)(2 * u)));
    if (span <= 32 * kBmsWords && (u > hmax || span <= 32 * (int64_t)H)) return SB_BMS;
    if (u <= hmax) return SB_H64 + (ilog2(H) - 6);
    if (u <= kSymHashMaxU && span <= kBigSpanFactor * u)
        return SB_H64 + (ilog2(max(64, next_pow2((int)(2 * u)))) - 6);
    return SB_BIG;
}

// Row r's entry range with every bound clamped into [0, nnz]: the row
// statistics run before the host has read the device operands' check, so
// their reads stay in bounds on operands that fail it (the host then
// discards the statistics and reports the operands invalid).
__device__ __forceinline__ int2 clamped_row(const int32_t *__restrict__ rp, int64_t r, int32_t nnz)
{
    const int s = min(max(rp[r], 0), nnz);
    return make_int2(s, min(max(rp[r + 1], s), nnz));
}

// Per B row: its length, first column and last column (INT32_MAX and -1
// when empty), so a row walk reads one 16-byte record per A entry instead
// of the row bounds and then both end columns.
__global__ __launch_bounds__(256) void k_brow_stats(int32_t k, int32_t nnz_b,
                                                    const int32_t *__restrict__ brp,
                                                    const int32_t *__restrict__ bci,
                                                    int4 *__restrict__ bst)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= k) return;
    const int2 b = clamped_row(brp, r, nnz_b);
    bst[r] = make_int4(b.y - b.x, b.y > b.x ? bci[b.x] : INT32_MAX, b.y > b.x ? bci[b.y - 1] : -1, 0);
}

// Sums |B_k|, and takes the smallest first and largest last column of the
// B_k, over the A entries ap0 + t * step (t < U, ap0 + t * step < end): the
// two dependent loads (A column, B row record) are each issued for all U
// entries before any is used.
// (A columns outside [0, kb) are skipped: see clamped_row.)
template <int U>
__device__ __forceinline__ void row_stats_step(const int32_t *__restrict__ aci,
                                               const int4 *__restrict__ bst, int32_t kb, int ap0,
                                               int step, int end, int64_t &ops, int32_t &lo,
                                               int32_t &hi)
{
    int k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = ap0 + u * step < end ? aci[ap0 + u * step] : -1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if ((uint32_t)k[u] < (uint32_t)kb) {
            const int4 b = bst[k[u]];
            ops += b.x;
            lo = min(lo, b.y);
            hi = max(hi, b.z);
        }
    }
}

// Per-row product count, product column span and symbolic bin. Four lanes
// per row (sixteen rows per wave) take the first 32 entries of every row in
// one step; the entries past that of the wave's longer rows are then taken
// one row at a time by the whole wave, eight entries per lane per step.
// Rows of more than kStatsLong entries are listed instead and reduced by a
// block each in k_row_stats_long: a skewed A keeps its longest rows at low
// indices (R-MAT 2^18: 15 744, then 6 600-entry rows), and one wave walking
// several of them in turn was the whole kernel's time. The blocks stride
// over groups of kStatsRowsPerBlock rows and leave their bin histogram,
// product and bitmap-word totals in a StatPart of their own, which
// k_stats_reduce sums: a global atomic per block and counter put every
// block of a 2^20-row A on the same few L2 addresses (mixed 2^20 x 2^16:
// 300 us for the pass).
static_assert(SB_COUNT <= 16, "a StatPart holds 16 bins");
constexpr int kStatsRowsPerBlock = 64;
constexpr int kStatsBlocks = kStatsParts / 2;  // per pass: the two share the partials
constexpr int kStatsLong = 512;
constexpr int kStatsLongThreads = 512;

// A row's product count and span, as both statistics kernels write them.
__device__ __forceinline__ void row_stats_finish(int64_t row, int64_t ops, int32_t lo, int32_t hi,
                                                 int64_t *__restrict__ ops_out,
                                                 int32_t *__restrict__ lo_out,
                                                 int32_t *__restrict__ hi_out)
{
    ops_out[row] = ops;
    lo_out[row] = ops ? lo : 0;
    hi_out[row] = ops ? hi : -1;
}

__global__ __launch_bounds__(256) void k_row_stats(int32_t m, int32_t nnz_a, int32_t kb,
                                                   const int32_t *__restrict__ arp,
                                                   const int32_t *__restrict__ aci,
                                                   const int4 *__restrict__ bst,
                                                   int64_t *__restrict__ ops_out,
                                                   int32_t *__restrict__ lo_out,
                                                   int32_t *__restrict__ hi_out,
                                                   uint8_t *__restrict__ sbin_out,
                                                   AnalysisCounters *__restrict__ ctr,
                                                   int hash_max_u, int64_t big_ops,
                                                   int32_t *__restrict__ long_rows,
                                                   StatPart *__restrict__ parts)
{
    constexpr int G = 4, U = 8;
    __shared__ int hist[SB_COUNT];
    __shared__ unsigned long long bprod, bcap, bops[SB_COUNT];
    if (threadIdx.x < SB_COUNT) {
        hist[threadIdx.x] = 0;
        bops[threadIdx.x] = 0ull;
    }
    if (threadIdx.x == 0) bprod = bcap = 0ull;
    __syncthreads();
    const int lane = threadIdx.x & 63, gl = lane % G;
    const int64_t ngroups = ((int64_t)m + kStatsRowsPerBlock - 1) / kStatsRowsPerBlock;
    for (int64_t rg = blockIdx.x; rg < ngroups; rg += gridDim.x) {  // uniform over the block
        const int64_t row = rg * kStatsRowsPerBlock + threadIdx.x / G;
        int64_t ops = 0;
        int32_t lo = INT32_MAX, hi = -1;
        int as = 0, ae = 0;
        if (row < m) {
            const int2 r = clamped_row(arp, row, nnz_a);
            as = r.x;
            ae = r.y;
        }
        const bool lng = ae - as > kStatsLong;  // the row's four lanes agree
        const uint64_t lmask = __ballot(gl == 0 && lng);
        if (lmask) {  // one list slot range per wave
            int base = 0;
            if (lane == __ffsll((unsigned long long)lmask) - 1)
                base = atomicAdd(&ctr->n_long, __popcll(lmask));
            base = __builtin_amdgcn_readlane(base, __ffsll((unsigned long long)lmask) - 1);
            if (gl == 0 && lng)
                long_rows[base + __popcll(lmask & ((1ull << lane) - 1ull))] = (int32_t)row;
        }
        if (lng) ae = as;
        const int e1 = min(ae, as + G * U);
        row_stats_step<U>(aci, bst, kb, as + gl, G, e1, ops, lo, hi);
        uint64_t longs = __ballot(gl == 0 && ae > e1);
        while (longs) {
            const int src = __ffsll((unsigned long long)longs) - 1;
            longs &= longs - 1;
            const int ls = lane_bcast(e1, src), le = lane_bcast(ae, src);
            int64_t o2 = 0;
            int32_t lo2 = INT32_MAX, hi2 = -1;
            for (int ap0 = ls + lane; ap0 < le; ap0 += U * 64)
                row_stats_step<U>(aci, bst, kb, ap0, 64, le, o2, lo2, hi2);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                o2 += __shfl_xor(o2, off, 64);
                lo2 = min(lo2, __shfl_xor(lo2, off, 64));
                hi2 = max(hi2, __shfl_xor(hi2, off, 64));
            }
            if (lane == src) {
                ops += o2;
                lo = min(lo, lo2);
                hi = max(hi, hi2);
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            ops += __shfl_xor(ops, off, G);
            lo = min(lo, __shfl_xor(lo, off, G));
            hi = max(hi, __shfl_xor(hi, off, G));
        }
        if (row < m && gl == 0 && !lng) {
            const int b = sym_bin_of(ops, lo, hi, hash_max_u, big_ops);
            row_stats_finish(row, ops, lo, hi, ops_out, lo_out, hi_out);
            sbin_out[row] = (uint8_t)b;
            atomicAdd(&hist[b], 1);
            if (b == SB_BIG) atomicAdd(&bcap, (unsigned long long)big_nchunks(lo, hi) * kChunkWords);
            if (ops) {
                atomicAdd(&bprod, (unsigned long long)ops);
                atomicAdd(&bops[b], (unsigned long long)ops);
            }
        }
    }
    __syncthreads();
    StatPart &pp = parts[blockIdx.x];
    if (threadIdx.x < SB_COUNT) {
        pp.hist[threadIdx.x] = hist[threadIdx.x];
        pp.ops[threadIdx.x] = bops[threadIdx.x];
    }
    if (threadIdx.x == 0) {
        pp.prod = bprod;
        pp.cap = bcap;
    }
}

// The listed long rows, a block of kStatsLongThreads each (the grid strides
// over the list, whose length the per-row pass left in ctr->n_long); block
// b's totals go to parts[b].
__global__ __launch_bounds__(kStatsLongThreads) void k_row_stats_long(
    int32_t nnz_a, int32_t kb, const int32_t *__restrict__ arp, const int32_t *__restrict__ aci, const int4 *__restrict__ bst,
    const int32_t *__restrict__ long_rows, int64_t *__restrict__ ops_out,
    int32_t *__restrict__ lo_out, int32_t *__restrict__ hi_out, uint8_t *__restrict__ sbin_out,
    const AnalysisCounters *__restrict__ ctr, int hash_max_u, int64_t big_ops,
    StatPart *__restrict__ parts)
{
    constexpr int U = 8, NWV = kStatsLongThreads / 64;
    __shared__ long long s_ops[NWV];
    __shared__ int s_lo[NWV], s_hi[NWV];
    __shared__ int hist[SB_COUNT];
    __shared__ unsigned long long bops[SB_COUNT];
    if (threadIdx.x < SB_COUNT) {
        hist[threadIdx.x] = 0;
        bops[threadIdx.x] = 0ull;
    }
    __syncthreads();
    unsigned long long bprod = 0, bcap = 0;  // thread 0's
    const int n = ctr->n_long;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {  // uniform over the block
        const int row = long_rows[i];
        const int2 r = clamped_row(arp, row, nnz_a);
        int64_t ops = 0;
        int32_t lo = INT32_MAX, hi = -1;
        for (int ap0 = r.x + (int)threadIdx.x; ap0 < r.y; ap0 += U * kStatsLongThreads)
            row_stats_step<U>(aci, bst, kb, ap0, kStatsLongThreads, r.y, ops, lo, hi);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            ops += __shfl_xor(ops, off, 64);
            lo = min(lo, __shfl_xor(lo, off, 64));
            hi = max(hi, __shfl_xor(hi, off, 64));
        }
        if (lane == 0) {
            s_ops[wv] = ops;
            s_lo[wv] = lo;
            s_hi[wv] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < NWV; ++w) {
                ops += s_ops[w];
                lo = min(lo, s_lo[w]);
                hi = max(hi, s_hi[w]);
            }
            const int b = sym_bin_of(ops, lo, hi, hash_max_u, big_ops);
            row_stats_finish(row, ops, lo, hi, ops_out, lo_out, hi_out);
            sbin_out[row] = (uint8_t)b;
            ++hist[b];
            bops[b] += (unsigned long long)ops;
            if (b == SB_BIG) bcap += (unsigned long long)big_nchunks(lo, hi) * kChunkWords;
            bprod += (unsigned long long)ops;
        }
        __syncthreads();  // the block's partials are read before the next row's
    }
    StatPart &pp = parts[blockIdx.x];
    if (threadIdx.x < SB_COUNT) {
        pp.hist[threadIdx.x] = hist[threadIdx.x];
        pp.ops[threadIdx.x] = bops[threadIdx.x];
    }
    if (threadIdx.x == 0) {
        pp.prod = bprod;
        pp.cap = bcap;
    }
}

// Sums the statistics passes' partials into the analysis counters (one
// block; nothing else writes these counters before it).
__global__ __launch_bounds__(256) void k_stats_reduce(const StatPart *__restrict__ parts, int n,
                                                      AnalysisCounters *__restrict__ ctr)
{
    constexpr int NV = 2 * SB_COUNT + 2;  // histogram, products, words, products per bin
    __shared__ unsigned long long s_v[4][NV];
    unsigned long long v[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b) v[b] = 0ull;
    for (int i = threadIdx.x; i < n; i += 256) {
#pragma unroll
        for (int b = 0; b < SB_COUNT; ++b) {
            v[b] += (unsigned long long)parts[i].hist[b];
            v[SB_COUNT + 2 + b] += parts[i].ops[b];
        }
        v[SB_COUNT] += parts[i].prod;
        v[SB_COUNT + 1] += parts[i].cap;
    }
#pragma unroll
    for (int b = 0; b < NV; ++b)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[b] += __shfl_xor(v[b], off, 64);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int b = 0; b < NV; ++b) s_v[threadIdx.x >> 6][b] = v[b];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const unsigned long long t = s_v[0][threadIdx.x] + s_v[1][threadIdx.x] +
                                     s_v[2][threadIdx.x] + s_v[3][threadIdx.x];
        if (threadIdx.x < SB_COUNT)
            ctr->sym_hist[threadIdx.x] = (int)t;
        else if (threadIdx.x == SB_COUNT)
            ctr->products = t;
        else if (threadIdx.x == SB_COUNT + 1)
            ctr->bits_cap = t;
        else
            ctr->sym_ops[threadIdx.x - SB_COUNT - 2] = t;
    }
}

// ---------------------------------------------------------------------------
// Scatter of rows into per-bin lists (order inside a bin is arbitrary;
// every row is processed independently, so results do not depend on it).
// A block takes kScatterRows rows, thread t rows t, t + 256, ... (so each
// atomic instruction places 64 consecutive rows and a bin's list stays
// near row order, which neighbouring stencil rows' shared B rows in L2
// reward), and counts them per bin in LDS before one global atomic per
// bin: a block of 256 rows and an atomic per bin each put every block of a
// 2^20-row matrix on the same few L2 addresses (50 us for the mixed
// 2^20 x 2^16 product).
// ---------------------------------------------------------------------------
constexpr int kScatterRows = 4096;

__global__ __launch_bounds__(256) void k_bin_scatter(int32_t m, const uint8_t *__restrict__ bin,
                                                     const int32_t *__restrict__ off,
                                                     int32_t *__restrict__ cursor,
                                                     int32_t *__restrict__ perm, int nb)
{
    constexpr int PER = kScatterRows / 256;
    __shared__ int lcount[64];
    __shared__ int lbase[64];
    if (threadIdx.x < 64) lcount[threadIdx.x] = 0;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * kScatterRows + threadIdx.x;
    uint8_t b[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) b[u] = r0 + u * 256 < m ? bin[r0 + u * 256] : (uint8_t)0xFF;
    int lpos[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) lpos[u] = b[u] != 0xFF ? atomicAdd(&lcount[b[u]], 1) : 0;
    __syncthreads();
    if (threadIdx.x < nb && lcount[threadIdx.x] > 0)
        lbase[threadIdx.x] = off[threadIdx.x] + atomicAdd(&cursor[threadIdx.x], lcount[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (b[u] != 0xFF) perm[lbase[b[u]] + lpos[u]] = (int32_t)(r0 + u * 256);
}

// ---------------------------------------------------------------------------
// Product columns of a row, in any order (the structure is a set): W <= 64
// lanes take the row's A entries as + off + step*t (t = 0, 1, ...) W at a
// time, load their B-row bounds together, then issue the first W-lane batch
// of eight entries' B rows before using any of them, so a row costs about
// 2 + entries/8 dependent loads instead of 3 per entry. The rest of a B row
// longer than W goes four W-lane batches per round.
// ---------------------------------------------------------------------------
template <int W>
__device__ __forceinline__ int group_bcast(int v, int j)
{
    if constexpr (W == 64)
        return lane_bcast(v, j);
    else
        return __shfl(v, j, W);
}

template <int W, typename Fn>
__device__ __forceinline__ void for_each_product_col(const int32_t *__restrict__ aci,
                                                     const int32_t *__restrict__ brp,
                                                     const int32_t *__restrict__ bci, int as,
                                                     int ae, int off, int step, Fn fn)
{
    static_assert(W <= 64, "a walker is at most one wave");
    const int lane = threadIdx.x % W;
    const int n = ae - as - off > 0 ? (ae - as - off + step - 1) / step : 0;
    for (int t0 = 0; t0 < n; t0 += W) {
        int bs = 0, be = 0;
        if (t0 + lane < n) {
            const int k = aci[as + off + step * (t0 + lane)];
            bs = brp[k];
            be = brp[k + 1];
        }
        const int nch = min(W, n - t0);  // uniform within the group
        for (int j0 = 0; j0 < nch; j0 += 8) {
            int su[8], eu[8];
            int32_t cu[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = min(j0 + u, W - 1);
                su[u] = group_bcast<W>(bs, j);
                eu[u] = group_bcast<W>(be, j);
                const int bp = su[u] + lane;
                cu[u] = j0 + u < nch && bp < eu[u] ? bci[bp] : -1;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (cu[u] >= 0) fn(cu[u]);
                if (j0 + u < nch)  // the rest of B_k, four chunks of W per round
                    for (int bp0 = su[u] + W; bp0 < eu[u]; bp0 += 4 * W) {
                        int32_t t4[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int bp = bp0 + t * W + lane;
                            t4[t] = bp < eu[u] ? bci[bp] : -1;
                        }
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            if (t4[t] >= 0) fn(t4[t]);
                    }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Symbolic pass, hash variant: G lanes per row, LDS key table of H slots.
// Count mode writes nnz(C_i); fill mode writes C_i's columns ascending
// (compaction, then a bitonic sort padded with INT32_MAX, over the block's
// largest row count rounded up to a power of two, at most H/2).
// Product order does not matter for the structure, so lanes take B entries
// freely.
// ---------------------------------------------------------------------------
template <int G, int H, bool FILL>
__global__ __launch_bounds__(256) void k_sym_hash(const int32_t *__restrict__ rows, int32_t nrows,
                                                  const int32_t *__restrict__ arp,
                                                  const int32_t *__restrict__ aci,
                                                  const int32_t *__restrict__ brp,
                                                  const int32_t *__restrict__ bci,
                                                  int32_t *__restrict__ cnt,
                                                  const int32_t *__restrict__ crp,
                                                  int32_t *__restrict__ cci, int *err)
{
    constexpr int R = 256 / G;
    constexpr int P = H / 2;  // capacity of the sorted list (nnz <= H/2)
    __shared__ int32_t keys[R * H];
    __shared__ int32_t list[FILL ? R * P : 1];
    __shared__ int counts[R];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    int32_t *T = keys + g * H;
    for (int s = lane; s < H; s += G) T[s] = -1;
    if (lane == 0) counts[g] = 0;
    __syncthreads();
    int mine = 0;
    if (valid) {
        // groups wider than a wave: each wave walks every (G/64)-th entry
        constexpr int WL = G < 64 ? G : 64;
        const int as = arp[row], ae = arp[row + 1];
        for_each_product_col<WL>(aci, brp, bci, as, ae, lane / WL, G / WL,
                                 [&](int32_t c) { mine += hash_insert<H>(T, c, err); });
    }
    if (mine) atomicAdd(&counts[g], mine);
    __syncthreads();
    if constexpr (!FILL) {
        if (valid && lane == 0) cnt[row] = counts[g];
        return;
    } else {
        int32_t *L = list + g * P;
        const int total = counts[g];
        __syncthreads();
        if (lane == 0) counts[g] = 0;
        for (int s = lane; s < P; s += G) L[s] = INT32_MAX;
        __syncthreads();
        for (int s = lane; s < H; s += G) {
            const int32_t key = T[s];
            if (key >= 0) {
                const int pos = atomicAdd(&counts[g], 1);
                if (pos < P) L[pos] = key;
            }
        }
        __syncthreads();
        // bitonic sort, ascending, of the first Q slots: Q = the block's
        // largest row count rounded up to a power of two (the table is
        // sized from the product count, an upper bound, so Q is often far
        // below P); the padding past a row's keys is INT32_MAX. Every
        // thread computes the same Q, so block barriers are uniform.
        int q = 2;
        for (int r = 0; r < R; ++r) q = max(q, counts[r]);
        q = min(P, next_pow2(q));
        for (int kk = 2; kk <= q; kk <<= 1) {
            for (int j = kk >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < q; i += G) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const int32_t x = L[i], y = L[ixj];
                        const bool up = (i & kk) == 0;
                        if ((x > y) == up) {
                            L[i] = y;
                            L[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (valid) {
            const int out = crp[row];
            for (int i = lane; i < total && i < P; i += G) cci[out + i] = L[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Count for rows whose table was sized by the learned sizer: the exact
// count when it fits the fill pass (<= H/2), else H + 1. A row stops
// early once its count passes H/2 (checked per A entry) or its table is
// full, so a badly underestimated row costs about H inserts.
// ---------------------------------------------------------------------------
template <int H>
__device__ __forceinline__ int hash_try_insert(int32_t *T, int32_t col)
{
    constexpr int LG = ilog2(H);
    uint32_t s = hash_slot(col, LG);
    for (int probe = 0; probe < H; ++probe) {
        int32_t cur = __hip_atomic_load(&T[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == col) return 0;
        if (cur == -1) {
            int32_t old = atomicCAS(&T[s], -1, col);
            if (old == -1) return 1;
            if (old == col) return 0;
        }
        s = (s + 1) & (H - 1);
    }
    return -1;  // full
}

template <int G, int H>
__global__ __launch_bounds__(256) void k_sym_hash_ovf(const int32_t *__restrict__ rows,
                                                      int32_t nrows,
                                                      const int32_t *__restrict__ arp,
                                                      const int32_t *__restrict__ aci,
                                                      const int32_t *__restrict__ brp,
                                                      const int32_t *__restrict__ bci,
                                                      int32_t *__restrict__ cnt)
{
    constexpr int R = 256 / G;
    constexpr int P = H / 2;
    __shared__ int32_t keys[R * H];
    __shared__ int counts[R];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    int32_t *T = keys + g * H;
    for (int s = lane; s < H; s += G) T[s] = -1;
    if (lane == 0) counts[g] = 0;
    __syncthreads();
    if (valid) {
        const int as = arp[row], ae = arp[row + 1];
        bool stop = false;
        for (int ap = as; ap < ae && !stop; ++ap) {
            if (__hip_atomic_load(&counts[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) > P)
                break;
            const int k = aci[ap];
            const int bs = brp[k], be = brp[k + 1];
            for (int bp = bs + lane; bp < be; bp += G) {
                const int r = hash_try_insert<H>(T, bci[bp]);
                if (r > 0) {
                    atomicAdd(&counts[g], 1);
                } else if (r < 0) {
                    atomicAdd(&counts[g], H);  // full: certainly over H/2
                    stop = true;
                    break;
                }
            }
        }
    }
    __syncthreads();
    if (valid && lane == 0) {
        const int c = counts[g];
        cnt[row] = c > P ? H + 1 : c;
    }
}

// ---------------------------------------------------------------------------
// Symbolic pass, bitmap variant for rows whose product columns span at most
// 32*WORDS columns: one bit per column from the row's smallest product
// column. Fill mode emits set bits in ascending order, which is C_i's column
// order by construction. (Wider rows go to the big-row path,
// rvsp_bigrow.hip.)
// ---------------------------------------------------------------------------
template <int G, int WORDS, bool FILL>
__global__ __launch_bounds__(256) void k_sym_bitmap(const int32_t *__restrict__ rows, int32_t nrows,
                                                    const int32_t *__restrict__ arp,
                                                    const int32_t *__restrict__ aci,
                                                    const int32_t *__restrict__ brp,
                                                    const int32_t *__restrict__ bci,
                                                    const int32_t *__restrict__ lo_col,
                                                    const int32_t *__restrict__ hi_col,
                                                    const int64_t *__restrict__ ops,
                                                    uint32_t *__restrict__ spill,
                                                    int32_t *__restrict__ cnt,
                                                    const int32_t *__restrict__ crp,
                                                    int32_t *__restrict__ cci)
{
    constexpr int R = 256 / G;
    constexpr int COLS = 32 * WORDS;
    constexpr int WPT = WORDS / G;  // words per lane (rounds of 64 in the fill)
    static_assert(WORDS % G == 0, "words per thread");
    static_assert(G == 64, "one row per wave");
    __shared__ uint32_t bits[R * WORDS];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    uint32_t *Bm = bits + g * WORDS;
    const int32_t base = valid ? lo_col[row] : 0;
    // words the row's product span covers (the rest stay zero)
    const int nw = valid ? min(WORDS, (int)(((int64_t)hi_col[row] - base + 32) >> 5)) : 0;
    // a row with many products per bitmap word keeps its count-pass bitmap
    // in HBM (slot ridx of the bin's list, the same in both passes), so the
    // fill pass reads it back instead of walking the products again; only
    // the span's words travel
    const bool spilled = spill && valid && ops[row] >= kBmsSpillOps;
    uint32_t *slot = spilled ? spill + ridx * WORDS : nullptr;
    if (FILL && spilled) {
        for (int w = lane; w < WORDS; w += G) Bm[w] = w < nw ? slot[w] : 0u;
    } else {
        for (int w = lane; w < WORDS; w += G) Bm[w] = 0u;
    }
    __syncthreads();
    if (valid && !(FILL && spilled)) {
        for_each_product_col<G>(aci, brp, bci, arp[row], arp[row + 1], 0, 1, [&](int32_t col) {
            const int32_t c = col - base;
            if (c >= 0 && c < COLS) atomicOr(&Bm[c >> 5], 1u << (c & 31));
        });
    }
    __syncthreads();
    if (!FILL && spilled)
        for (int w = lane; w < nw; w += G) slot[w] = Bm[w];
    if (!valid) return;  // no block barriers below: one row per wave
    if constexpr (FILL) {
        // ascending columns, 64 words per round (lane l: word 64r + l): a
        // wave scan of the round's popcounts places each lane's bits, so one
        // store instruction covers a short run of C_i and no lane expands
        // more than one word per round
        int run = crp[row];
        const int rounds = (nw + 63) / 64;  // words past the span are zero
#pragma unroll 1
        for (int r = 0; r < rounds; ++r) {
            uint32_t x = Bm[r * 64 + lane];
            const int c = __popc(x);
            int incl = c;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(incl, off, 64);
                if (lane >= off) incl += y;
            }
            int pos = run + incl - c;
            const int32_t cb = base + (r * 64 + lane) * 32;
            while (x) {
                const int b = __ffs(x) - 1;
                x &= x - 1;
                cci[pos++] = cb + b;
            }
            run += lane_bcast(incl, 63);
        }
    } else {
        int local = 0;
#pragma unroll
        for (int w = 0; w < WPT; ++w) local += __popc(Bm[w * 64 + lane]);
        const int gtotal = group_sum<64>(local);
        if (lane == 0) cnt[row] = gtotal;
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of per-row counts into C's row pointer (int64 running sum,
// int32 stored; the host rejects totals above INT32_MAX). Three launches:
// tile sums, scan of tile sums, tile scan plus offset.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;

__global__ __launch_bounds__(256) void k_scan_tiles(int32_t m, const int32_t *__restrict__ cnt,
                                                    int64_t *__restrict__ tsum)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < m) s += cnt[base + j];
    int64_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_top(int32_t ntiles, int64_t *__restrict__ tsum,
                                                  int64_t *__restrict__ total_out)
{
    int64_t carry = 0;
    for (int base = 0; base < ntiles; base += 256) {
        const int i = base + threadIdx.x;
        const int64_t v = i < ntiles ? tsum[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan(v, &tot);
        if (i < ntiles) tsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(int32_t m, const int32_t *__restrict__ cnt,
                                                    const int64_t *__restrict__ tsum,
                                                    const int64_t *__restrict__ total,
                                                    int32_t *__restrict__ crp)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < m ? cnt[base + j] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = tsum[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < m) {
            crp[base + j] = (int32_t)run;
            run += v[j];
        }
    if (blockIdx.x == 0 && threadIdx.x == 0) crp[m] = (int32_t)*total;
}

// ---------------------------------------------------------------------------
// Numeric binning from exact nnz(C_i): LDS hash (H from nnz, G from the
// mean B-row length), a single dense window, or the big-row path (rows the
// symbolic pass kept a bitmap for).
// ---------------------------------------------------------------------------
// Lanes per row: about the mean |B_k| (lanes cover one B row per k-step),
// raised so that a block of R rows still holds a full wave.
__device__ __forceinline__ int lanes_for(int64_t ops, int alen, int rows_per_block_max)
{
    const int64_t avgb = alen > 0 ? (ops + alen - 1) / alen : 1;
    int G = next_pow2((int)min<int64_t>(avgb, 64));
    G = max(8, min(64, G));
    const int gmin = rows_per_block_max >= 8 ? 8 : 64 / rows_per_block_max;
    return max(G, gmin);
}

// Candidate accumulators, by LDS bytes per row: a hash table of H >= 1.5 nnz
// slots (8 bytes each; counted twice, for its probing), a dense window of
// W >= span columns (4 bytes each), or a rank window (a 512-word bitmap over
// a span of at most 16384 columns with its 16-bit popcount prefix, 3 KB,
// plus 4 bytes per output position, CAP >= nnz). The cheapest feasible one
// wins.
__device__ __forceinline__ int num_bin_of(int nz, int64_t ops, int alen, int64_t span,
                                          bool critical)
{
    const int H = nz <= kHashMaxNnz ? max(64, next_pow2((3 * nz + 1) / 2)) : 0;
    const int Wd = span <= kDenseMaxCols ? max(512, next_pow2((int)span)) : 0;
    const int C = span <= kRankMaxCols && nz <= kRankMaxNnz ? max(512, next_pow2(nz)) : 0;
    const int64_t fh = H ? 16LL * H : INT64_MAX;
    const int64_t fd = Wd ? 4LL * Wd : INT64_MAX;
    const int64_t fr = C ? kRankWords * 6LL + 4LL * C : INT64_MAX;
    if (fr <= fh && fr <= fd) {
        const int G = critical ? 64 : max(32, lanes_for(ops, alen, 4));
        return NB_RANK0 + 2 * (ilog2(C) - 9) + (G == 64 ? 1 : 0);
    }
    if (fd <= fh) {
        const int G = critical ? 64 : lanes_for(ops, alen, 16384 / Wd);
        return NB_DENSE0 + 4 * (ilog2(Wd) - 9) + (ilog2(G) - 3);
    }
    const int G = critical ? 64 : lanes_for(ops, alen, 8192 / H);
    return NB_HASH0 + 4 * (ilog2(H) - 6) + (ilog2(G) - 3);
}

__global__ __launch_bounds__(256) void k_num_bins(int32_t m, const int32_t *__restrict__ crp,
                                                  const int64_t *__restrict__ ops,
                                                  const int32_t *__restrict__ arp,
                                                  const uint8_t *__restrict__ sbin,
                                                  const int32_t *__restrict__ lo,
                                                  const int32_t *__restrict__ hi,
                                                  const float *__restrict__ chooser,
                                                  float lg_margin, int64_t products,
                                                  uint8_t *__restrict__ nbin,
                                                  AnalysisCounters *__restrict__ ctr)
{
    __shared__ int hist[NB_COUNT];
    __shared__ unsigned long long bops[NB_COUNT];
    if (threadIdx.x < NB_COUNT) {
        hist[threadIdx.x] = 0;
        bops[threadIdx.x] = 0ull;
    }
    __syncthreads();
    // (the chooser's weights are read from its global parameters at
    // wave-uniform addresses, so they come through scalar loads into SGPRs
    // instead of one LDS read per multiply-add)
    // blocks stride over the rows, so the counters take one global atomic
    // per bin per block of a bounded grid
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256) {
        const int nz = crp[i + 1] - crp[i];
        int b;
        if (nz == 0)
            b = NB_EMPTY;
        else if (sbin[i] == SB_BIG)
            b = NB_WIN;  // segments planned by k_big_plan
        else {
            const int alen = arp[i + 1] - arp[i];
            const int64_t span = (int64_t)hi[i] - lo[i] + 1;
            // a row holding at least 1/kCriticalShare of the pass's products
            // can outlast the rest of the pass: 64 lanes for its latency, and
            // the chooser (trained on throughput per row) is not asked
            const bool critical = ops[i] * kCriticalShare >= products;
            const int hb = num_bin_of(nz, ops[i], alen, span, critical);
            b = chooser && !critical ? chooser_pick(chooser, ops[i], alen, span, nz, hb, lg_margin) : -1;
            if (b < 0) b = hb;
        }
        nbin[i] = (uint8_t)b;
        atomicAdd(&hist[b], 1);
        if (ops[i]) atomicAdd(&bops[b], (unsigned long long)ops[i]);
    }
    __syncthreads();
    if (threadIdx.x < NB_COUNT && hist[threadIdx.x]) {
        atomicAdd(&ctr->num_hist[threadIdx.x], hist[threadIdx.x]);
        if (bops[threadIdx.x]) atomicAdd(&ctr->num_ops[threadIdx.x], bops[threadIdx.x]);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int G, int H, bool FILL>
static void launch_sym_hash(hipStream_t st, const SymArgs &a, const int32_t *rows, int nrows)
{
    const int R = 256 / G;
    const int grid = (nrows + R - 1) / R;
    k_sym_hash<G, H, FILL><<<grid, 256, 0, st>>>(rows, nrows, a.arp, a.aci, a.brp, a.bci, a.cnt,
                                                  a.crp, a.cci, a.err);
}

template <int G, int H>
static void launch_sym_hash_ovf(hipStream_t st, const SymArgs &a, const int32_t *rows, int nrows)
{
    const int R = 256 / G;
    const int grid = (nrows + R - 1) / R;
    k_sym_hash_ovf<G, H><<<grid, 256, 0, st>>>(rows, nrows, a.arp, a.aci, a.brp, a.bci, a.cnt);
}

static void launch_sym_learned(hipStream_t st, const SymArgs &a, int b, const int32_t *rows, int nr)
{
    switch (b) {
    case SB_H64: launch_sym_hash_ovf<8, 64>(st, a, rows, nr); break;
    case SB_H128: launch_sym_hash_ovf<16, 128>(st, a, rows, nr); break;
    case SB_H256: launch_sym_hash_ovf<32, 256>(st, a, rows, nr); break;
    case SB_H512: launch_sym_hash_ovf<64, 512>(st, a, rows, nr); break;
    case SB_H1024: launch_sym_hash_ovf<64, 1024>(st, a, rows, nr); break;
    case SB_H2048: launch_sym_hash_ovf<128, 2048>(st, a, rows, nr); break;
    case SB_H4096: launch_sym_hash_ovf<256, 4096>(st, a, rows, nr); break;
    default: break;
    }
}

template <bool FILL>
static void launch_sym_bins(hipStream_t main, const SymArgs &a, const int32_t *perm,
                            const int32_t *count, const int32_t *off, bool learned = false,
                            bool skip_big = false, bool fan = false,
                            const unsigned long long *work = nullptr)
{
    // bins touch disjoint rows: the big rows (the longest work) first on the
    // caller's stream, the others over the streams, largest first onto the
    // least loaded when the bins' products are known, else in turn over the
    // side streams
    StreamFan f(main, fan);
    static_assert(SB_BIG == SB_COUNT - 1 && SB_BMS == SB_BIG - 1, "big rows are the last bin");
    const bool big = count[SB_BIG] > 0 && !skip_big;
    if (work && big) f.preload(work[SB_BIG]);
    int order[SB_COUNT];
    hipStream_t sts[SB_COUNT];
    order[0] = SB_BIG;
    sts[0] = main;
    const int n = (big ? 1 : 0) + f.assign(SB_H64, SB_BIG, count, work, order + (big ? 1 : 0),
                                           sts + (big ? 1 : 0));
    for (int bi = 0; bi < n; ++bi) {
        const int b = order[bi];
        const int nr = count[b];
        const hipStream_t st = sts[bi];
        const int32_t *rows = perm + off[b];
        if (!FILL && learned && b >= SB_H64 && b <= SB_H4096) {
            launch_sym_learned(st, a, b, rows, nr);
            continue;
        }
        switch (b) {
        case SB_H64: launch_sym_hash<8, 64, FILL>(st, a, rows, nr); break;
        case SB_H128: launch_sym_hash<16, 128, FILL>(st, a, rows, nr); break;
        case SB_H256: launch_sym_hash<32, 256, FILL>(st, a, rows, nr); break;
        case SB_H512: launch_sym_hash<64, 512, FILL>(st, a, rows, nr); break;
        case SB_H1024: launch_sym_hash<64, 1024, FILL>(st, a, rows, nr); break;
        case SB_H2048: launch_sym_hash<128, 2048, FILL>(st, a, rows, nr); break;
        case SB_H4096: launch_sym_hash<256, 4096, FILL>(st, a, rows, nr); break;
        case SB_BMS: {
            const int grid = (nr + 3) / 4;
            k_sym_bitmap<64, kBmsWords, FILL><<<grid, 256, 0, st>>>(
                rows, nr, a.arp, a.aci, a.brp, a.bci, a.lo, a.hi, a.ops, a.bms_spill, a.cnt,
                a.crp, a.cci);
            break;
        }
        case SB_BIG:
            if constexpr (FILL)
                launch_big_fill(st, a.big, a.crp, a.cci);
            else
                launch_sym_big(st, rows, nr, a.big);
            break;
        default: break;
        }
    }
}

void launch_row_stats(hipStream_t st, int32_t m, int32_t nnz_a, const int32_t *arp,
                      const int32_t *aci, int32_t k, int32_t nnz_b, const int32_t *brp,
                      const int32_t *bci, int4 *bst, int64_t *ops, int32_t *lo, int32_t *hi,
                      uint8_t *sbin, AnalysisCounters *ctr, int hash_max_u, int32_t *long_rows,
                      StatPart *parts)
{
    if (k > 0) k_brow_stats<<<(int)(((int64_t)k + 255) / 256), 256, 0, st>>>(k, nnz_b, brp, bci, bst);
    const int64_t ngroups = ((int64_t)m + kStatsRowsPerBlock - 1) / kStatsRowsPerBlock;
    const int g1 = (int)std::max<int64_t>(1, std::min<int64_t>(ngroups, kStatsBlocks));
    const int hmax = hash_max_u < 64 ? 64 : hash_max_u > kSymHashMaxU ? kSymHashMaxU : hash_max_u;
    const int64_t big_ops = nnz_a >= kBigRowNnz ? kBigRowOps : INT64_MAX;
    k_row_stats<<<g1, 256, 0, st>>>(m, nnz_a, k, arp, aci, bst, ops, lo, hi, sbin, ctr, hmax,
                                    big_ops, long_rows, parts);
    int g2 = 0;
    if (nnz_a > kStatsLong) {  // else no row can be long
        g2 = std::min(nnz_a / (kStatsLong + 1), kStatsBlocks);  // listed rows at most
        k_row_stats_long<<<g2, kStatsLongThreads, 0, st>>>(nnz_a, k, arp, aci, bst, long_rows, ops,
                                                           lo, hi, sbin, ctr, hmax, big_ops,
                                                           parts + g1);
    }
    k_stats_reduce<<<1, 256, 0, st>>>(parts, g1 + g2, ctr);
}

void launch_bin_scatter(hipStream_t st, int32_t m, const uint8_t *bin, const int32_t *off,
                        int32_t *cursor, int32_t *perm, int nb)
{
    if (m <= 0) return;
    const int grid = (int)(((int64_t)m + kScatterRows - 1) / kScatterRows);
    k_bin_scatter<<<grid, 256, 0, st>>>(m, bin, off, cursor, perm, nb);
}

void launch_sym_count(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                      const int32_t *off, bool learned, bool fan, const unsigned long long *work)
{
    launch_sym_bins<false>(st, a, perm, count, off, learned, false, fan, work);
}

void launch_sym_fill(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                     const int32_t *off, bool skip_big, bool fan, const unsigned long long *work)
{
    launch_sym_bins<true>(st, a, perm, count, off, false, skip_big, fan, work);
}

void launch_scan(hipStream_t st, int32_t m, const int32_t *cnt, int64_t *tsum, int64_t *total,
                 int32_t *crp)
{
    const int ntiles = (int)(((int64_t)m + kScanTile - 1) / kScanTile);
    const int nt = ntiles > 0 ? ntiles : 1;
    k_scan_tiles<<<nt, 256, 0, st>>>(m, cnt, tsum);
    k_scan_top<<<1, 256, 0, st>>>(nt, tsum, total);
    k_scan_apply<<<nt, 256, 0, st>>>(m, cnt, tsum, total, crp);
}

int scan_tiles(int32_t m)
{
    const int ntiles = (int)(((int64_t)m + kScanTile - 1) / kScanTile);
    return ntiles > 0 ? ntiles : 1;
}

void launch_num_bins(hipStream_t st, int32_t m, const int32_t *crp, const int64_t *ops,
                     const int32_t *arp, const uint8_t *sbin, const int32_t *lo,
                     const int32_t *hi, const float *chooser, float lg_margin,
                     int64_t products, uint8_t *nbin, AnalysisCounters *ctr)
{
    if (m <= 0) return;
    const int grid = (int)std::min<int64_t>(((int64_t)m + 255) / 256, 512);
    k_num_bins<<<grid, 256, 0, st>>>(m, crp, ops, arp, sbin, lo, hi, chooser, lg_margin, products,
                                     nbin, ctr);
}

} // namespace rvsp
