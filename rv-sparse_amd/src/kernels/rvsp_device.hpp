// rvsp_device.hpp -- shared device helpers for the gfx950 SpGEMM kernels.
//
// Wave64 throughout. "Group" = G consecutive lanes that cooperate on one
// output row (G divides 64, or G = 256 for a whole block).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace rvsp {

constexpr int kWave = 64;

// Error bits raised by kernels (read back by rvsp_hip_plan_check).
constexpr int kErrTableFull = 1;    // hash insert found no free slot
constexpr int kErrProbeMiss = 2;    // lookup of a column C's structure lacks
constexpr int kErrWindow = 4;       // product column outside its window
constexpr int kErrOperand = 8;      // a device operand is not canonical CSR

// Symbolic bins (per row, from product count and column span).
enum SymBin : int {
    SB_EMPTY = 0,
    SB_H64, SB_H128, SB_H256, SB_H512, SB_H1024, SB_H2048, SB_H4096,
    SB_BMS,  // one-pass bitmap over a span of <= 16384 columns
    SB_BIG,  // multi-pass 2^19-column bitmap, plus window cursors
    SB_COUNT
};

// Numeric bins: 0 empty; NB_HASH0 + 4*(log2(H)-6) + (log2(G)-3) for LDS
// hash tables of H slots worked by G lanes; NB_DENSE0 + 4*(log2(W)-9) +
// (log2(G)-3) for dense windows of W columns over the row's product span;
// NB_RANK0 + 2*(log2(CAP)-9) + (G == 64) for rank windows (a bitmap of C_i's
// columns over a span of at most kRankMaxCols, its popcount prefix, and an
// accumulator of CAP >= nnz(C_i) output positions; G = 32 or 64); NB_WIN
// for big rows (rvsp_bigrow.hip).
constexpr int NB_EMPTY = 0;
constexpr int NB_HASH0 = 1;
constexpr int NB_DENSE0 = NB_HASH0 + 7 * 4;  // 29
constexpr int NB_RANK0 = NB_DENSE0 + 6 * 4;  // 53
constexpr int NB_WIN = NB_RANK0 + 4 * 2;     // 61
constexpr int NB_COUNT = NB_WIN + 1;         // 62
// bin scatter keeps per-bin offsets and cursors in 64-entry regions
static_assert(NB_COUNT <= 64 && SB_COUNT <= 64, "bin tables hold 64 bins");

constexpr int kBmsWords = 512;            // 16384 columns
constexpr int kBmsSpillOps = 2 * kBmsWords;  // products above which a bitmap row spills
constexpr int kBigPassWords = 16384;      // big-row bitmap: 2^19 columns per pass
constexpr int kChunkCols = 1024;          // big-row chunk (window granularity)
constexpr int kChunkWords = kChunkCols / 32;
constexpr int kSegProducts = 16384;       // default products per big-row segment (one wave)
constexpr int kDenseMaxCols = 16384;      // widest dense window (64 KB acc)
constexpr int kHashMaxNnz = 2730;         // numeric hash rows: H = 4096 at load <= 2/3
constexpr int kRankMaxCols = 16384;       // rank windows: span (512 bitmap words)
constexpr int kRankWords = kRankMaxCols / 32;
constexpr int kRankMaxNnz = 4096;         // rank windows: outputs
constexpr int kSymHashMaxU = 2048;        // largest symbolic hash limit (tables of 4096 keys)
// A numeric pass runs on the order of 2048 rows at once (256 CUs, a few
// LDS-sized rows each): a row with at least 1/kCriticalShare of the pass's
// products can take longer alone than the rest of the pass together.
constexpr int64_t kCriticalShare = 2048;

__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

// Adds x into an LDS accumulator slot that no other lane of the same
// instruction touches (the lanes of one k-step hold distinct columns), as a
// plain read-modify-write. A wave's LDS operations take effect in issue
// order, and the compiler keeps program order between accesses it cannot
// tell apart (k-steps are also separated by compiler fences), so the next
// k-step's read of the slot sees this write. The value is the one rounding
// fl(slot + x) that the LDS atomic add gives (integers wrap). On gfx950 the
// read-modify-write is much the faster: the 27-point stencil's ranked
// reuse pass takes 0.50 ms with it and 1.04 ms with ds_add_f32, R-MAT
// 2^18's big-row numeric pass 19.7 ms against 23.9 ms.
template <class T>
__device__ __forceinline__ void lds_add(T *slot, T x)
{
    if constexpr (std::is_integral<T>::value)
        *slot = (T)((std::make_unsigned_t<T>)*slot + (std::make_unsigned_t<T>)x);
    else
        *slot = *slot + x;
}

// A big row's bitmap starts at its lowest product column rounded down to a
// chunk, so its chunks are aligned to global 1024-column boundaries.
__host__ __device__ __forceinline__ int32_t big_base(int32_t lo) { return lo & ~(kChunkCols - 1); }
__host__ __device__ __forceinline__ int big_nchunks(int32_t lo, int32_t hi)
{
    return (int)(((int64_t)hi - big_base(lo)) / kChunkCols) + 1;
}

// One A entry of a row walk, as the lanes of a group share it: its value,
// its B row [bs, be), and (rank walks) the products of earlier entries of
// its chunk. Read by every lane of the group with one 16-byte LDS load.
struct __align__(16) KMeta {
    float av;
    int bs, be, pre;
};

// Value of lane `src` (the same for every lane of the wave) broadcast to all
// lanes through a scalar register; __shfl would use an LDS permute.
__device__ __forceinline__ int lane_bcast(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
// Lane `dst` (the same for every lane) of a and b set to the wave-uniform
// va and vb: one v_writelane_b32 each, which write that lane whatever EXEC
// holds (instead of a compare and two selects). The lane select goes
// through M0 (a VOP3 reads one SGPR on the constant bus), written by a SALU
// move, which the lane selects may read after one wait state.
__device__ __forceinline__ void lane_set2(int &a, int &b, int va, int vb, int dst)
{
    asm("s_mov_b32 m0, %4\n\ts_nop 0\n\tv_writelane_b32 %0, %2, m0\n\t"
        "v_writelane_b32 %1, %3, m0"
        : "+v"(a), "+v"(b)
        : "s"(va), "s"(vb), "s"(dst)
        : "m0");
}
__device__ __forceinline__ float lane_bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

__device__ __forceinline__ double lane_bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- value types of the numeric pass ----------------------------------------
// FP32 (rv-sparse's contract) and the legacy dtypes: FP64 in and out, INT8 in
// with INT32 accumulation. B is read as (column, value) pairs: int2 for FP32
// (value bits) and INT8 (value widened to int), PairD for FP64.
struct __align__(16) PairD {
    int32_t col, pad;
    double val;
};

__device__ __forceinline__ int32_t pair_col(int2 p) { return p.x; }
__device__ __forceinline__ int32_t pair_col(const PairD &p) { return p.col; }
template <class Pair>
__device__ __forceinline__ Pair pair_none();
template <>
__device__ __forceinline__ int2 pair_none<int2>() { return make_int2(INT32_MAX, 0); }
template <>
__device__ __forceinline__ PairD pair_none<PairD>() { return PairD{INT32_MAX, 0, 0.0}; }
__device__ __forceinline__ void pair_value(int2 p, float &v) { v = __int_as_float(p.y); }
__device__ __forceinline__ void pair_value(int2 p, int32_t &v) { v = p.y; }
__device__ __forceinline__ void pair_value(const PairD &p, double &v) { v = p.val; }

// One product, rounded once (no fused multiply-add); INT32 wraps.
__device__ __forceinline__ float vmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double vmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ int32_t vmul(int32_t a, int32_t b)
{
    return (int32_t)((uint32_t)a * (uint32_t)b);
}

// Orders LDS traffic of one wave: everything issued before is visible to
// every lane of the wave after. Same sequence rocPRIM uses for wave-level
// cooperation through LDS.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

__device__ __forceinline__ uint32_t hash_slot(int32_t col, int lg)
{
    return ((uint32_t)col * 0x9E3779B1u) >> (32 - lg);
}

__device__ __forceinline__ int next_pow2(int v)
{
    return v <= 1 ? 1 : 1 << (32 - __clz(v - 1));
}

// Inserts col into an LDS key table of H slots (empty = -1). Returns 1 when
// this call placed it, 0 when it was present. Linear probing, bounded.
template <int H>
__device__ __forceinline__ int hash_insert(int32_t *T, int32_t col, int *err)
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
    atomicOr(err, kErrTableFull);
    return 0;
}

// Inserts col and returns its slot (-1 if the table is full).
template <int H>
__device__ __forceinline__ int hash_insert_slot(int32_t *T, int32_t col, int *err)
{
    constexpr int LG = ilog2(H);
    uint32_t s = hash_slot(col, LG);
    for (int probe = 0; probe < H; ++probe) {
        int32_t cur = __hip_atomic_load(&T[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == col) return (int)s;
        if (cur == -1) {
            int32_t old = atomicCAS(&T[s], -1, col);
            if (old == -1 || old == col) return (int)s;
        }
        s = (s + 1) & (H - 1);
    }
    atomicOr(err, kErrTableFull);
    return -1;
}

// Slot of a key known to be present (-1 and an error bit if absent).
template <int H>
__device__ __forceinline__ int hash_find(const int32_t *T, int32_t col, int *err)
{
    constexpr int LG = ilog2(H);
    uint32_t s = hash_slot(col, LG);
    for (int probe = 0; probe < H; ++probe) {
        int32_t cur = T[s];
        if (cur == col) return (int)s;
        if (cur == -1) break;
        s = (s + 1) & (H - 1);
    }
    atomicOr(err, kErrProbeMiss);
    return -1;
}

template <int G>
__device__ __forceinline__ int group_sum(int v)
{
    static_assert(G <= kWave, "group_sum is wave-level");
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

// Cross-lane steps on DPP (data-parallel primitives: a VALU operand read
// from another lane of the same row of 16, or broadcast across rows), for
// code where every lane of the wave is active. __shfl compiles to an LDS
// permute (ds_bpermute) and a wait per step instead.
namespace dpp {
constexpr int kQuadSwap1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kRowMirror = 0x140;    // lane i <- lane 15 - i of its row
constexpr int kRowHalfMirror = 0x141;  // lane i <- lane 7 - i of its half row
constexpr int kRowBcast15 = 0x142;   // lane 15 of each row to the next row
constexpr int kRowBcast31 = 0x143;   // lane 31 to the rows above it
constexpr int kRowShr = 0x110;       // + k: lane i <- lane i - k of its row

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int read(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xF, false);
}
} // namespace dpp

// Minimum over the 64 lanes (all active), as a wave-uniform value.
__device__ __forceinline__ int wave_min_all(int v)
{
    v = min(v, dpp::read<dpp::kQuadSwap1>(INT32_MAX, v));
    v = min(v, dpp::read<dpp::kQuadSwap2>(INT32_MAX, v));
    v = min(v, dpp::read<dpp::kRowHalfMirror>(INT32_MAX, v));
    v = min(v, dpp::read<dpp::kRowMirror>(INT32_MAX, v));      // every lane: its row's minimum
    v = min(v, dpp::read<dpp::kRowBcast15, 0xA>(INT32_MAX, v));  // rows 1, 3: two rows'
    v = min(v, dpp::read<dpp::kRowBcast31, 0xC>(INT32_MAX, v));  // row 3: all four rows'
    return __builtin_amdgcn_readlane(v, 63);
}

// Inclusive prefix sum over the 64 lanes (all active).
__device__ __forceinline__ int wave_incl_scan_all(int v)
{
    v += dpp::read<dpp::kRowShr + 1>(0, v);
    v += dpp::read<dpp::kRowShr + 2>(0, v);
    v += dpp::read<dpp::kRowShr + 4>(0, v);
    v += dpp::read<dpp::kRowShr + 8>(0, v);                      // within each row of 16
    v += dpp::read<dpp::kRowBcast15, 0xA>(0, v);                 // rows 1, 3 add the row before
    v += dpp::read<dpp::kRowBcast31, 0xC>(0, v);                 // rows 2, 3 add rows 0-1
    return v;
}

// Exclusive scan over a block of NWV waves (256 threads by default); *total
// gets the sum. Every thread of the block must call it.
template <int NWV = 4>
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *total)
{
    __shared__ int64_t wsum[NWV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NWV; ++w) {
        if (w < wid) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// First index p in [lo, hi) of a sorted device array with a[p] >= key,
// found by the whole wave (uniform result). Every lane must call it.
__device__ __forceinline__ int wave_lower_bound(const int32_t *__restrict__ a, int lo, int hi,
                                                int32_t key)
{
    const int lane = threadIdx.x & (kWave - 1);
    while (hi - lo > kWave) {
        const int len = hi - lo;
        // lane j samples the last element of the j-th of 64 slices
        const int pos = lo + (int)(((int64_t)(lane + 1) * len) / kWave) - 1;
        const bool below = a[pos] < key;
        const uint64_t m = __ballot(below);
        const int c = __popcll(m);  // slices entirely below key
        const int nlo = lo + (int)(((int64_t)c * len) / kWave);
        const int nhi = c == kWave ? hi : lo + (int)(((int64_t)(c + 1) * len) / kWave);
        lo = nlo;
        hi = nhi;
    }
    const int p = lo + lane;
    const bool below = p < hi && a[p] < key;
    return lo + __popcll(__ballot(below));
}

} // namespace rvsp
