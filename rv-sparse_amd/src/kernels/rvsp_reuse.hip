// rvsp_reuse.hip -- numeric passes under a reused structure (descriptor
// compute() called again on the structure it already filled).
//
// While the structure stays the same, every product's place in C is fixed:
// product (k, j) of row i lands at the rank of column j among C_i's sorted
// columns. For rows on the LDS paths (hash and dense bins: at most 16384
// outputs, so a rank fits 16 bits) the plan computes that rank once per
// product, in the row's walk order (A entries in stored order, each B row in
// order), and keeps it in HBM: 1 byte per product for rows of at most 256
// outputs, 2 bytes otherwise (each row's map starts at an even byte offset).
// The rank stream is the pass's only traffic without reuse, so the byte
// ranks cut what a pass reads by up to a third. A numeric pass then adds
// fl(a*b) into acc[rank] -- a dense LDS accumulator over the row's OUTPUT
// positions -- with no column read and no probing, in the same ascending-k
// order as every other kernel here (one group of lanes per row, one LDS add
// per product, a compiler fence between k-steps), and drains acc[0..nnz)
// straight into C's values: bit-identical results. Big rows keep their
// windowed kernel (rvsp_bigrow.hip).
#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

__device__ __forceinline__ bool small_bin(int b) { return b >= NB_HASH0 && b < NB_WIN; }

// Rows of at most this many outputs keep 1-byte ranks.
constexpr int kByteRankMax = 256;

// Bytes of each row's rank map (0 for empty and big rows): its products at
// 1 byte (rounded up to even, so 2-byte maps stay aligned) or 2 bytes each.
__global__ __launch_bounds__(256) void k_pm_count(int32_t m, const int64_t *__restrict__ ops,
                                                  const uint8_t *__restrict__ nbin,
                                                  const int32_t *__restrict__ crp,
                                                  int64_t *__restrict__ cnt)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    if (!small_bin(nbin[i])) {
        cnt[i] = 0;
        return;
    }
    const bool narrow = crp[i + 1] - crp[i] <= kByteRankMax;
    cnt[i] = narrow ? (ops[i] + 1) & ~(int64_t)1 : 2 * ops[i];
}

// Exclusive int64 scan of cnt into off (three launches, tiles of 2048).
constexpr int kPmItems = 8;
constexpr int kPmTile = 256 * kPmItems;

__global__ __launch_bounds__(256) void k_pm_scan_tiles(int32_t m, const int64_t *__restrict__ cnt,
                                                       int64_t *__restrict__ tsum)
{
    const int64_t base = (int64_t)blockIdx.x * kPmTile + threadIdx.x * kPmItems;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < kPmItems; ++j)
        if (base + j < m) s += cnt[base + j];
    int64_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_pm_scan_top(int32_t ntiles, int64_t *__restrict__ tsum,
                                                     int64_t *__restrict__ total)
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
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_pm_scan_apply(int32_t m, const int64_t *__restrict__ cnt,
                                                       const int64_t *__restrict__ tsum,
                                                       int64_t *__restrict__ off)
{
    const int64_t base = (int64_t)blockIdx.x * kPmTile + threadIdx.x * kPmItems;
    int64_t v[kPmItems], s = 0;
#pragma unroll
    for (int j = 0; j < kPmItems; ++j) {
        v[j] = base + j < m ? cnt[base + j] : 0;
        s += v[j];
    }
    int64_t tot;
    int64_t run = tsum[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
    for (int j = 0; j < kPmItems; ++j)
        if (base + j < m) {
            off[base + j] = run;
            run += v[j];
        }
}

// Ranks of every product of the rows on the LDS paths: one wave per row,
// C_i's columns staged in LDS when they fit (else searched in HBM).
constexpr int kPmStage = 2048;

__global__ __launch_bounds__(256) void k_pm_build(int32_t m, const uint8_t *__restrict__ nbin,
                                                  const int32_t *__restrict__ arp,
                                                  const int32_t *__restrict__ aci,
                                                  const int32_t *__restrict__ brp,
                                                  const int32_t *__restrict__ bci,
                                                  const int32_t *__restrict__ crp,
                                                  const int32_t *__restrict__ cci,
                                                  const int64_t *__restrict__ prow,
                                                  uint8_t *__restrict__ pmap, int *err)
{
    __shared__ int32_t stage[4][kPmStage];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wv;
    if (row >= m || !small_bin(nbin[row])) return;  // whole wave; no block barriers below
    const int cs = crp[row], nz = crp[row + 1] - cs;
    if (nz > 65535) {  // the LDS paths never hold such a row
        if (lane == 0) atomicOr(err, kErrWindow);
        return;
    }
    const bool staged = nz <= kPmStage;
    const int32_t *cols = staged ? stage[wv] : cci + cs;
    if (staged)
        for (int q = lane; q < nz; q += 64) stage[wv][q] = cci[cs + q];
    wave_sync();
    const bool narrow = nz <= kByteRankMax;
    uint8_t *m8 = pmap + prow[row];
    uint16_t *m16 = reinterpret_cast<uint16_t *>(m8);  // even offset for 2-byte maps
    int64_t p = 0;
    const int as = arp[row], ae = arp[row + 1];
    for (int ap = as; ap < ae; ++ap) {
        const int k = aci[ap];
        const int bs = brp[k], be = brp[k + 1];
        for (int bp = bs + lane; bp < be; bp += 64) {
            const int32_t col = bci[bp];
            int lo = 0, hi = nz;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cols[mid] < col)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            if (lo >= nz || cols[lo] != col) atomicOr(err, kErrProbeMiss);
            const int r = lo < nz ? lo : 0;
            if (narrow)
                m8[p + (bp - bs)] = (uint8_t)r;
            else
                m16[p + (bp - bs)] = (uint16_t)r;
        }
        p += be - bs;
    }
}

// Walk of one row with its ranks: G lanes, A entries in stored order, each
// k-step's adds separated from the next by a compiler fence; loads for four
// k-steps are issued before their adds. A k-step's entry reaches the lanes
// through a scalar register (G = 64) or the group's LDS table M.
template <int G, typename Add>
__device__ __forceinline__ void walk_row_ranked(const NumArgs &a, int as, int ae,
                                                const uint8_t *__restrict__ rmap, bool narrow,
                                                KMeta *M, Add add)
{
    const int lane = threadIdx.x % G;
    const uint16_t *__restrict__ rmap16 = reinterpret_cast<const uint16_t *>(rmap);
    auto rank = [&](int64_t q) { return narrow ? (int)rmap[q] : (int)rmap16[q]; };
    int64_t run = 0;  // first product (row-relative) of the current chunk of G A entries
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
        // products of earlier entries of the chunk (inclusive scan over G lanes)
        const int len = be - bs;
        int incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int y = __shfl_up(incl, o, G);
            if (lane >= o) incl += y;
        }
        const int pre = incl - len;
        if constexpr (G < 64) {
            M[lane] = KMeta{av, bs, be, pre};
            wave_sync();
        }
        const int nch = min(G, ae - base);
        for (int j0 = 0; j0 < nch; j0 += 4) {
            float au[4], vu[4];
            int su[4], eu[4], pu[4], ru[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = min(j0 + u, G - 1);
                if constexpr (G == 64) {
                    au[u] = lane_bcast(av, j);
                    su[u] = lane_bcast(bs, j);
                    eu[u] = lane_bcast(be, j);
                    pu[u] = lane_bcast(pre, j);
                } else {
                    const KMeta e = M[j];
                    au[u] = e.av;
                    su[u] = e.bs;
                    eu[u] = e.be;
                    pu[u] = e.pre;
                }
                const int bp = su[u] + lane;
                const bool ok = j0 + u < nch && bp < eu[u];
                ru[u] = ok ? rank(run + pu[u] + lane) : -1;
                vu[u] = ok ? a.bval[bp] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (j0 + u < nch) {  // uniform within the group
                    if (ru[u] >= 0) add(ru[u], __fmul_rn(au[u], vu[u]));
                    for (int bp = su[u] + G + lane; bp < eu[u]; bp += G)
                        add(rank(run + pu[u] + (bp - su[u])), __fmul_rn(au[u], a.bval[bp]));
                    compiler_fence();
                }
            }
        }
        if constexpr (G == 64)
            run += lane_bcast(incl, G - 1);  // the chunk's total products
        else {
            const KMeta last = M[G - 1];
            run += last.pre + (last.be - last.bs);
            wave_sync();  // the table is rewritten for the next chunk
        }
    }
}

// Numeric pass of one hash or dense bin with ranks: R rows of G lanes per
// block, an LDS accumulator of CAP >= nnz(C_i) floats per row.
template <int G, int CAP, int R>
__global__ __launch_bounds__(R *G) void k_num_ranked(const int32_t *__restrict__ rows,
                                                     int32_t nrows, NumArgs a,
                                                     const int64_t *__restrict__ prow,
                                                     const uint8_t *__restrict__ pmap)
{
    __shared__ float accs[R * CAP];
    __shared__ KMeta meta[G < 64 ? R * G : 1];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t ridx = (int64_t)blockIdx.x * R + g;
    const bool valid = ridx < nrows;
    const int row = valid ? rows[ridx] : 0;
    float *S = accs + g * CAP;
    const int cs = valid ? a.crp[row] : 0;
    const int nz = valid ? min(a.crp[row + 1] - cs, CAP) : 0;
    for (int i = lane; i < nz; i += G) S[i] = 0.0f;
    __syncthreads();
    if (valid) {
        walk_row_ranked<G>(a, a.arp[row], a.arp[row + 1], pmap + prow[row],
                           a.crp[row + 1] - cs <= kByteRankMax, meta + (G < 64 ? g * G : 0),
                           [&](int pos, float prod) {
                               if (pos < nz) lds_add(&S[pos], prod);
                           });
    }
    __syncthreads();
    for (int q = lane; q < nz; q += G) a.cval[cs + q] = S[q];
}

template <int G, int CAP>
static void launch_ranked(hipStream_t st, const NumArgs &a, const int32_t *rows, int nrows,
                          const int64_t *prow, const uint8_t *pmap)
{
    constexpr int R = (256 / G) < (16384 / CAP) ? (256 / G) : (16384 / CAP);
    static_assert(R * G >= 64, "block must hold at least one full wave");
    k_num_ranked<G, CAP, R><<<(nrows + R - 1) / R, R * G, 0, st>>>(rows, nrows, a, prow, pmap);
}

// The hash bin (H, G), the dense bin (W, G) and the rank-window bin (CAP, G)
// hold rows of at most H, W and CAP outputs: the rank accumulator takes that
// many floats.
#define RVSP_RANK_HASH(HI, GI, H, G)                                                            \
    case NB_HASH0 + 4 * (HI) + (GI):                                                            \
        launch_ranked<G, H>(st, a, rows, nr, prow, pmap);                                       \
        return true;
#define RVSP_RANK_WIN(CI, CAP, G)                                                               \
    case NB_RANK0 + 2 * (CI) + ((G) == 64 ? 1 : 0):                                             \
        launch_ranked<G, CAP>(st, a, rows, nr, prow, pmap);                                     \
        return true;
#define RVSP_RANK_DENSE(WI, GI, W, G)                                                           \
    case NB_DENSE0 + 4 * (WI) + (GI):                                                           \
        launch_ranked<G, W>(st, a, rows, nr, prow, pmap);                                       \
        return true;

static bool launch_ranked_bin(hipStream_t st, const NumArgs &a, int bin, const int32_t *rows,
                              int nr, const int64_t *prow, const uint8_t *pmap)
{
    switch (bin) {
        RVSP_RANK_HASH(0, 0, 64, 8)
        RVSP_RANK_HASH(0, 1, 64, 16)
        RVSP_RANK_HASH(0, 2, 64, 32)
        RVSP_RANK_HASH(0, 3, 64, 64)
        RVSP_RANK_HASH(1, 0, 128, 8)
        RVSP_RANK_HASH(1, 1, 128, 16)
        RVSP_RANK_HASH(1, 2, 128, 32)
        RVSP_RANK_HASH(1, 3, 128, 64)
        RVSP_RANK_HASH(2, 0, 256, 8)
        RVSP_RANK_HASH(2, 1, 256, 16)
        RVSP_RANK_HASH(2, 2, 256, 32)
        RVSP_RANK_HASH(2, 3, 256, 64)
        RVSP_RANK_HASH(3, 0, 512, 8)
        RVSP_RANK_HASH(3, 1, 512, 16)
        RVSP_RANK_HASH(3, 2, 512, 32)
        RVSP_RANK_HASH(3, 3, 512, 64)
        RVSP_RANK_HASH(4, 0, 1024, 8)
        RVSP_RANK_HASH(4, 1, 1024, 16)
        RVSP_RANK_HASH(4, 2, 1024, 32)
        RVSP_RANK_HASH(4, 3, 1024, 64)
        RVSP_RANK_HASH(5, 1, 2048, 16)
        RVSP_RANK_HASH(5, 2, 2048, 32)
        RVSP_RANK_HASH(5, 3, 2048, 64)
        RVSP_RANK_HASH(6, 2, 4096, 32)
        RVSP_RANK_HASH(6, 3, 4096, 64)
        RVSP_RANK_DENSE(0, 0, 512, 8)
        RVSP_RANK_DENSE(0, 1, 512, 16)
        RVSP_RANK_DENSE(0, 2, 512, 32)
        RVSP_RANK_DENSE(0, 3, 512, 64)
        RVSP_RANK_DENSE(1, 0, 1024, 8)
        RVSP_RANK_DENSE(1, 1, 1024, 16)
        RVSP_RANK_DENSE(1, 2, 1024, 32)
        RVSP_RANK_DENSE(1, 3, 1024, 64)
        RVSP_RANK_DENSE(2, 0, 2048, 8)
        RVSP_RANK_DENSE(2, 1, 2048, 16)
        RVSP_RANK_DENSE(2, 2, 2048, 32)
        RVSP_RANK_DENSE(2, 3, 2048, 64)
        RVSP_RANK_DENSE(3, 1, 4096, 16)
        RVSP_RANK_DENSE(3, 2, 4096, 32)
        RVSP_RANK_DENSE(3, 3, 4096, 64)
        RVSP_RANK_DENSE(4, 2, 8192, 32)
        RVSP_RANK_DENSE(4, 3, 8192, 64)
        RVSP_RANK_DENSE(5, 3, 16384, 64)
        RVSP_RANK_WIN(0, 512, 32)
        RVSP_RANK_WIN(0, 512, 64)
        RVSP_RANK_WIN(1, 1024, 32)
        RVSP_RANK_WIN(1, 1024, 64)
        RVSP_RANK_WIN(2, 2048, 32)
        RVSP_RANK_WIN(2, 2048, 64)
        RVSP_RANK_WIN(3, 4096, 32)
        RVSP_RANK_WIN(3, 4096, 64)
    default:
        return false;
    }
}

void launch_pm_build(hipStream_t st, int32_t m, const uint8_t *nbin, const int64_t *ops,
                     const int32_t *arp, const int32_t *aci, const int32_t *brp,
                     const int32_t *bci, const int32_t *crp, const int32_t *cci, int64_t *cnt,
                     int64_t *tsum, int64_t *total, int64_t *prow, uint8_t *pmap, int *err)
{
    const int grid = (int)(((int64_t)m + 255) / 256);
    const int ntiles = (int)(((int64_t)m + kPmTile - 1) / kPmTile);
    k_pm_count<<<grid, 256, 0, st>>>(m, ops, nbin, crp, cnt);
    k_pm_scan_tiles<<<ntiles, 256, 0, st>>>(m, cnt, tsum);
    k_pm_scan_top<<<1, 256, 0, st>>>(ntiles, tsum, total);
    k_pm_scan_apply<<<ntiles, 256, 0, st>>>(m, cnt, tsum, prow);
    if (pmap)
        k_pm_build<<<(int)(((int64_t)m + 3) / 4), 256, 0, st>>>(m, nbin, arp, aci, brp, bci, crp,
                                                                 cci, prow, pmap, err);
}

int pm_scan_tiles(int32_t m) { return (int)(((int64_t)m + kPmTile - 1) / kPmTile); }

void launch_numeric_ranked(hipStream_t st, const NumArgs &a, const int32_t *perm,
                           const int32_t *count, const int32_t *off, const BigArgs &big,
                           int nsegs, const int64_t *prow, const uint8_t *pmap, bool fan,
                           const unsigned long long *work)
{
    // bins touch disjoint rows: big rows first on the caller's stream, the
    // others over the streams largest first (none under graph capture)
    StreamFan f(st, fan);
    launch_num_big(st, nsegs, a, big);
    if (work) f.preload(work[NB_WIN]);
    int order[NB_COUNT];
    hipStream_t sts[NB_COUNT];
    const int n = f.assign(NB_HASH0, NB_WIN, count, work, order, sts);
    for (int i = 0; i < n; ++i)
        (void)launch_ranked_bin(sts[i], a, order[i], perm + off[order[i]], count[order[i]], prow,
                                pmap);
}

} // namespace rvsp
