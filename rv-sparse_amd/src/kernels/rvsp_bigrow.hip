// rvsp_bigrow.hip -- rows whose products span more than 16384 columns and
// exceed 2048 distinct-column candidates ("big rows").
//
// A big row's structure is a bitmap over its product column span, built
// once by the symbolic count and kept in HBM (one bit per column, plus a
// popcount per 1024-column chunk). Everything after the count reads that
// bitmap instead of walking products again:
//   fill     ordered compaction of the bitmap into C's column indices;
//   plan     the row is cut into windows of <= g.win_outs outputs and at
//            most g.win_chunks chunks, and windows are grouped into segments;
//   numeric  one wave per segment. For each window the wave loads the
//            window's bitmap words into LDS with their exclusive popcount
//            prefix (a rank dictionary: column -> position in C_i), zeroes a
//            dense accumulator over the window's OUTPUT positions, walks the
//            row's A entries in stored order with one cursor per entry
//            (carried from window to window; the first cursor of each group
//            of 64 also holds the group's smallest next column, so groups
//            with nothing in the window are skipped unread), and adds each product into
//            acc[rank(col)] with one LDS read-modify-write. One wave, entries in
//            ascending k, columns of one B row distinct: the same ordered,
//            separately rounded accumulation as rv-sparse's scalar kernel
//            (gustavson_core_f32.inc:31-63, accum_scalar_f32.c:67-70).
#include <stdlib.h>

#include <type_traits>

#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Symbolic count of big rows: one block per row, an LDS bitmap of
// g.pass_words words (sized at launch from B's column count, at most 2^19
// columns) per pass over the row's span. Product order does not matter for
// the structure, so B's rows are read as (word, bits) pairs, one per
// distinct 32-column word of a row (built once per analysis by
// launch_big_words): each pair is one LDS atomic OR of a whole word, and
// the pairs of one B row name distinct words, so the lanes of one atomic
// never share a word. R-MAT 2^18's big rows read 1.53e9 pairs for 2.90e9
// products (hub rows are dense in their low columns). The A entries of a
// chunk of NT are split by their B row's pair count: a thread walks a short
// row alone (four loads in flight); long ones are cut into pieces of
// kSymPiece pairs shared out over the waves, eight loads per lane.
// ---------------------------------------------------------------------------
// Big rows of at most kRegGroups * 64 A entries keep their per-entry cursors
// in registers in the numeric pass (and need no cursor storage).
#ifndef RVSP_BIG_REG_GROUPS
#define RVSP_BIG_REG_GROUPS 2
#endif
constexpr int kRegGroups = RVSP_BIG_REG_GROUPS;
static_assert(kRegGroups >= 1 && kRegGroups <= 4, "register cursors: one round of the search");

constexpr int kSymLong = 48;    // B rows of more pairs than this are shared by a wave
constexpr int kSymPiece = 512;  // pairs per wave piece (8 per lane)

// LDS slot of bitmap word w: its low six bits XORed with the next six.
// Skewed column distributions (R-MAT's bits are each 0 with probability
// a + c) crowd the words that share their low bits, and so one LDS bank;
// the field spreads them about as well as a multiplicative hash (64-word
// batches of R-MAT 2^18's B rows: 4.45 lanes on the fullest of 32 banks
// against 4.35, and 6.07 unhashed) with a bit-field extract and an XOR
// instead of a quarter-rate 32-bit multiply. A bijection on every 64-word
// block, so the bitmap needs its word count rounded up to 64.
static_assert(kBigPassWords % 64 == 0, "a full bitmap pass is whole 64-word blocks");

__device__ __forceinline__ int lds_word(int w)
{
    return w ^ (int)__builtin_amdgcn_ubfe((uint32_t)w, 6u, 6u);
}

template <int NT>
__global__ __launch_bounds__(NT) void k_sym_big(const int32_t *__restrict__ rows, int32_t nrows,
                                                BigArgs g)
{
    extern __shared__ __align__(16) uint32_t bits[];
    __shared__ long long s_off;
    __shared__ int s_part[NT / 64];
    __shared__ int s_pfx[NT + 1], s_bs[NT], s_be[NT];
    constexpr int NWV = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row = rows[blockIdx.x];
    const int32_t blo = big_base(g.lo[row]);
    const int nwords = big_nchunks(g.lo[row], g.hi[row]) * kChunkWords;
    const int passw = g.pass_words;
    if (tid == 0) {
        s_off = (long long)atomicAdd(g.bits_fill, (unsigned long long)nwords);
        g.big_off[row] = s_off;
    }
    __syncthreads();
    const int64_t off = s_off;
    const int as = g.arp[row], ae = g.arp[row + 1];
    int total = 0;
    for (int pw0 = 0; pw0 < nwords; pw0 += passw) {
        const int pwn = min(passw, nwords - pw0);
        // lds_word permutes within whole 64-word blocks, so the words of
        // [0, pwn) land in [0, pwn rounded up to 64): zeroed 16 bytes a store
        for (int i = tid; i < ((pwn + 63) & ~63) / 4; i += NT)
            reinterpret_cast<uint4 *>(bits)[i] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        // pair words are absolute: this pass holds words [wbase, wbase + pwn)
        const int32_t wbase = (blo >> 5) + pw0;
        for (int cb = as; cb < ae; cb += NT) {
            const int myap = cb + tid;
            int mbs = 0, mbe = 0;
            if (myap < ae) {
                const int k = g.aci[myap];
                mbs = g.bwp[k];
                mbe = g.bwp[k + 1];
            }
            const bool lng = mbe - mbs > kSymLong;
            int64_t npieces;
            const int pfx = (int)block_excl_scan<NWV>(
                lng ? (mbe - mbs + kSymPiece - 1) / kSymPiece : 0, &npieces);
            s_pfx[tid] = pfx;
            s_bs[tid] = mbs;
            s_be[tid] = mbe;
            if (tid == 0) s_pfx[NT] = (int)npieces;
            __syncthreads();
            if (!lng) {  // this thread's own short B row
                for (int bp = mbs; bp < mbe; bp += 4) {
                    int2 c[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) c[u] = bp + u < mbe ? g.bwm[bp + u] : make_int2(-1, 0);
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t rel = (uint32_t)(c[u].x - wbase);  // a sentinel is out
                        if (rel < (uint32_t)pwn) atomicOr(&bits[lds_word((int)rel)], (uint32_t)c[u].y);
                    }
                }
            }
            for (int pc = wv; pc < (int)npieces; pc += NWV) {
                // entry e holding piece pc: the last e with s_pfx[e] <= pc
                // (s_pfx is non-decreasing). pc is wave-uniform, so the
                // lanes search together: every S-th prefix, then the S
                // prefixes of the block found, each a ballot whose set
                // lanes are a prefix of the wave (one LDS read a level
                // instead of a binary search's ten dependent rounds)
                constexpr int S = NT / 64;
                static_assert(NT % 64 == 0 && S <= 64, "two ballot levels cover the entries");
                const int b0 = (__popcll(__ballot(s_pfx[lane * S] <= pc)) - 1) * S;
                const int lo = b0 - 1 + __popcll(__ballot(lane < S && s_pfx[b0 + min(lane, S - 1)] <= pc));
                // wave-uniform (pc is): held in scalar registers
                const int bs =
                    __builtin_amdgcn_readfirstlane(s_bs[lo] + (pc - s_pfx[lo]) * kSymPiece);
                const int be = __builtin_amdgcn_readfirstlane(min(bs + kSymPiece, s_be[lo]));
                // wave-uniform base and 32-bit lane offsets: the loads take
                // the scalar-base address form
                const int2 *__restrict__ at = g.bwm + bs;
                const int left = be - bs;
                int2 c[kSymPiece / 64];
#pragma unroll
                for (int u = 0; u < kSymPiece / 64; ++u)
                    c[u] = u * 64 + lane < left ? at[u * 64 + lane] : make_int2(-1, 0);
#pragma unroll
                for (int u = 0; u < kSymPiece / 64; ++u) {
                    const uint32_t rel = (uint32_t)(c[u].x - wbase);
                    if (rel < (uint32_t)pwn) atomicOr(&bits[lds_word((int)rel)], (uint32_t)c[u].y);
                }
            }
            __syncthreads();
        }
        for (int w = tid; w < pwn; w += NT) g.gbits[off + pw0 + w] = bits[lds_word(w)];
        const int nch = pwn / kChunkWords;
        int part = 0;
        for (int ch = tid; ch < nch; ch += NT) {
            int sc = 0;
#pragma unroll 8
            for (int w = 0; w < kChunkWords; ++w) sc += __popc(bits[lds_word(ch * kChunkWords + w)]);
            g.gchunk[off / kChunkWords + pw0 / kChunkWords + ch] = sc;
            part += sc;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if (lane == 0) s_part[wv] = part;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NWV; ++i) total += s_part[i];
        __syncthreads();
    }
    if (tid == 0) g.cnt[row] = total;
}

// ---------------------------------------------------------------------------
// B's rows as (word, bits) pairs: an entry starts a pair when its word
// differs from the previous entry's or it opens its row; a scan of those
// flags places the pairs, and each pair's first entry ORs its run (at most
// 32 entries: a row's columns are distinct).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bw_flags(int32_t nnz, const int32_t *__restrict__ bci,
                                                  int32_t *__restrict__ flag)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nnz) flag[i] = i == 0 || (bci[i] >> 5) != (bci[i - 1] >> 5);
}

__global__ __launch_bounds__(256) void k_bw_row_starts(int32_t k, const int32_t *__restrict__ brp,
                                                       int32_t *__restrict__ flag)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < k && brp[r + 1] > brp[r]) flag[brp[r]] = 1;
}

__global__ __launch_bounds__(256) void k_bw_write(int32_t nnz, const int32_t *__restrict__ bci,
                                                  const int32_t *__restrict__ flag,
                                                  const int32_t *__restrict__ pidx,
                                                  int2 *__restrict__ bwm)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz || !flag[i]) return;
    uint32_t bits = 1u << (bci[i] & 31);
    for (int64_t j = i + 1; j < nnz && j < i + 32 && !flag[j]; ++j) bits |= 1u << (bci[j] & 31);
    bwm[pidx[i]] = make_int2(bci[i] >> 5, (int)bits);
}

__global__ __launch_bounds__(256) void k_bw_rows(int32_t k, const int32_t *__restrict__ brp,
                                                 const int32_t *__restrict__ pidx,
                                                 int32_t *__restrict__ bwp)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r <= k) bwp[r] = pidx[brp[r]];  // pidx[nnz] is the pair total
}

void launch_big_words(hipStream_t st, int32_t k, int32_t nnz, const int32_t *brp,
                      const int32_t *bci, int32_t *flag, int32_t *pidx, int64_t *tsum,
                      int64_t *total, int2 *bwm, int32_t *bwp)
{
    const int ge = (int)(((int64_t)nnz + 255) / 256), gr = (int)(((int64_t)k + 256) / 256);
    if (nnz > 0) {
        k_bw_flags<<<ge, 256, 0, st>>>(nnz, bci, flag);
        k_bw_row_starts<<<gr, 256, 0, st>>>(k, brp, flag);
    }
    launch_scan(st, nnz, flag, tsum, total, pidx);
    if (nnz > 0) k_bw_write<<<ge, 256, 0, st>>>(nnz, bci, flag, pidx, bwm);
    k_bw_rows<<<gr, 256, 0, st>>>(k, brp, pidx, bwp);
}

// ---------------------------------------------------------------------------
// Big rows largest first. The count runs one block per row, about four
// blocks per CU, and product counts are heavy-tailed (R-MAT 2^18: 90 359 big
// rows, the largest 0.9 of the mean load of one of ~1024 concurrent
// blocks), so a row started late stretches the whole launch. Rows are
// bucketed by floor(4 log2 products) (two mantissa bits), descending: a
// greedy schedule of that order is within 0.1 % of the exact sort's.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int big_order_key(int64_t ops)
{
    if (ops < 4) return 255 - (int)(ops < 0 ? 0 : ops);
    const int e = 63 - __clzll((unsigned long long)ops);
    return 255 - (4 * e + (int)((ops >> (e - 2)) & 3));  // 4e + f <= 251
}

__global__ __launch_bounds__(256) void k_big_order_hist(const int32_t *__restrict__ rows, int32_t n,
                                                        const int64_t *__restrict__ ops,
                                                        int32_t *__restrict__ hist)
{
    __shared__ int lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) atomicAdd(&lh[big_order_key(ops[rows[j]])], 1);
    __syncthreads();
    if (lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

// One block: hist (256) -> exclusive offsets in work[256..511]; hist is
// cleared to serve as the scatter's cursors.
__global__ __launch_bounds__(256) void k_big_order_scan(int32_t *__restrict__ work)
{
    const int h = work[threadIdx.x];
    int64_t total;
    const int64_t x = block_excl_scan(h, &total);
    work[256 + threadIdx.x] = (int)x;
    work[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_big_order_scatter(const int32_t *__restrict__ rows,
                                                           int32_t n,
                                                           const int64_t *__restrict__ ops,
                                                           int32_t *__restrict__ work,
                                                           int32_t *__restrict__ out)
{
    __shared__ int lcount[256], lbase[256];
    lcount[threadIdx.x] = 0;
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int row = j < n ? rows[j] : 0;
    const int b = j < n ? big_order_key(ops[row]) : -1;
    const int lpos = b >= 0 ? atomicAdd(&lcount[b], 1) : 0;
    __syncthreads();
    if (lcount[threadIdx.x] > 0) lbase[threadIdx.x] = atomicAdd(&work[threadIdx.x], lcount[threadIdx.x]);
    __syncthreads();
    if (b >= 0) out[work[256 + b] + lbase[b] + lpos] = row;
}

hipError_t launch_big_order(hipStream_t st, int32_t *rows, int nrows, const int64_t *ops,
                            int32_t *work, int32_t *tmp)
{
    if (nrows < 2) return hipSuccess;
    const int grid = (nrows + 255) / 256;
    hipError_t e = hipMemsetAsync(work, 0, 256 * 4, st);
    if (e != hipSuccess) return e;
    k_big_order_hist<<<grid, 256, 0, st>>>(rows, nrows, ops, work);
    k_big_order_scan<<<1, 256, 0, st>>>(work);
    k_big_order_scatter<<<grid, 256, 0, st>>>(rows, nrows, ops, work, tmp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(rows, tmp, (size_t)nrows * 4, hipMemcpyDeviceToDevice, st);
}

// ---------------------------------------------------------------------------
// Structure fill of big rows from their bitmaps: one wave per segment of the
// numeric plan, window by window (at most PW outputs, MAXCH chunks). The
// wave loads the window's bitmap words 64 at a time (one per lane), ranks
// each round with a wave scan, expands the set bits into an LDS stage at
// their ordered positions and copies the stage out in coalesced runs.
// ---------------------------------------------------------------------------
template <int NWV, int PW, int MAXCH>
__global__ __launch_bounds__(NWV * 64) void k_big_fill(int32_t nsegs, BigArgs g,
                                                       const int32_t *__restrict__ crp,
                                                       int32_t *__restrict__ cci)
{
    constexpr int PER = MAXCH * kChunkWords / 64;
    __shared__ int32_t stage_s[NWV][PW];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t it = (int64_t)blockIdx.x * NWV + wv;
    if (it >= nsegs) return;  // whole wave leaves; no block barriers below
    int32_t *stage = stage_s[wv];
    const int4 sg = g.segs[it];
    const int row = sg.x, w0 = sg.y, w1 = sg.z;
    const int32_t blo = big_base(g.lo[row]);
    const int nchunks = big_nchunks(g.lo[row], g.hi[row]);
    const int64_t boff = g.big_off[row];
    const int64_t c32 = boff / kChunkWords;
    const int nwin = g.big_nwin[row];
    const int rowbase = crp[row];
    const int nz = crp[row + 1] - rowbase;
    for (int w = w0; w < w1; ++w) {
        const int ch0 = g.win_chunk0[c32 + w];
        const int ch1 = w + 1 < nwin ? g.win_chunk0[c32 + w + 1] : nchunks;
        const int o0 = g.win_out[c32 + w];
        const int outs = (w + 1 < nwin ? g.win_out[c32 + w + 1] : nz) - o0;
        const int nw = (ch1 - ch0) * kChunkWords;  // <= MAXCH * kChunkWords
        // round r: lane l holds word r * 64 + l (coalesced), so no lane
        // expands more than one word's 32 bits per round
        uint32_t wl[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int i = r * 64 + lane;
            wl[r] = i < nw ? g.gbits[boff + (int64_t)ch0 * kChunkWords + i] : 0u;
        }
        const int32_t cb0 = blo + ch0 * kChunkCols + lane * 32;
        int run = 0;
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int cnt = __popc(wl[r]);
            const int incl = wave_incl_scan_all(cnt);
            int pos = run + incl - cnt;
            // one bound check per word (never failing for the plan's
            // windows) instead of one per bit
            uint32_t x = pos + cnt <= PW ? wl[r] : 0u;
            const int32_t cw = cb0 + r * 64 * 32;
            while (x) {
                const int b = __ffs(x) - 1;
                x &= x - 1;
                stage[pos++] = cw + b;
            }
            run += lane_bcast(incl, 63);
        }
        wave_sync();
        for (int i = lane; i < outs && i < PW; i += 64) cci[rowbase + o0 + i] = stage[i];
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Window and segment plan: one thread per big row. Windows are runs of
// whole chunks with at most g.win_outs outputs and g.win_chunks chunks.
// Segments are runs of windows holding about g.seg_products products
// (estimated from the row's products per output), each processed by one
// wave, so the widest rows spread over many waves; their cursor blocks
// (alen entries each) are carved from a running counter.
// ---------------------------------------------------------------------------
// Four ints read as one 16-byte access at 4-byte alignment.
struct __attribute__((aligned(4))) I4 {
    int x, y, z, w;
};

// The segment list holds the rows' segments one row after another in the
// rows' list order, largest first, so the numeric pass starts the longest
// rows first. (Reserving each block's range with an atomic left the order
// to the blocks' finishing times: R-MAT 2^18's numeric pass took 15.8 to
// 16.8 ms depending on it, 18.7 ms with the smallest rows first, 15.7 ms
// largest first; generated 32768^2 6.0 to 7.5 ms, 5.45 ms largest first.)
//
// EMIT false: windows, then each row's segment count (segcnt[j]); EMIT
// true: each row's segments at its place in the list, from the scanned
// counts (segoff, nrows + 1 entries).
template <bool EMIT>
__global__ __launch_bounds__(256) void k_big_plan(const int32_t *__restrict__ rows, int32_t nrows,
                                                  BigArgs g, const int32_t *__restrict__ crp,
                                                  AnalysisCounters *__restrict__ ctr,
                                                  int32_t *__restrict__ segcnt,
                                                  const int32_t *__restrict__ segoff)
{
    __shared__ long long s_cur;      // block's first cursor entry
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = j < nrows;    // every thread joins the block scans below
    const int row = valid ? rows[j] : 0;
    const int nchunks = big_nchunks(g.lo[row], g.hi[row]);
    const int64_t c32 = valid ? g.big_off[row] / kChunkWords : 0;
    const int alen = g.arp[row + 1] - g.arp[row];
    const int nz = crp[row + 1] - crp[row];
    // products per output of this row: segments aim at ~g.seg_products each
    const double ratio = nz > 0 ? (double)g.ops[row] / (double)nz : 1.0;
    int nwin = 0;
    if (EMIT) {
        if (valid) nwin = g.big_nwin[row];
    } else if (valid) {
        int out = 0, acc = 0, start = 0;
        // pass 1: windows (chunk counts read 16 bytes at a time, eight per
        // round trip: a thread walks its own row, so each load is a line of
        // its own; the target allows 16-byte loads at 4-byte alignment, and
        // a load that starts inside the row's counts ends at most 12 bytes
        // past them, inside the big-row allocation, where the window arrays
        // follow the counts and the segments follow the window offsets;
        // entries past the row are not used)
        for (int c0 = 0; c0 < nchunks; c0 += 32) {
            int cc[32];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const I4 v = c0 + 4 * q < nchunks
                                 ? *reinterpret_cast<const I4 *>(g.gchunk + c32 + c0 + 4 * q)
                                 : I4{0, 0, 0, 0};
                cc[4 * q] = v.x;
                cc[4 * q + 1] = v.y;
                cc[4 * q + 2] = v.z;
                cc[4 * q + 3] = v.w;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int c = c0 + u;
                if (c >= nchunks) break;
                const int cnt_c = cc[u];
                if (c > start && (acc + cnt_c > g.win_outs || c - start >= g.win_chunks)) {
                    g.win_chunk0[c32 + nwin] = start;
                    g.win_out[c32 + nwin] = out;
                    ++nwin;
                    out += acc;
                    start = c;
                    acc = 0;
                }
                acc += cnt_c;
            }
        }
        g.win_chunk0[c32 + nwin] = start;
        g.win_out[c32 + nwin] = out;
        ++nwin;
        g.big_nwin[row] = nwin;
    }
    // pass 2: segments, counted (EMIT false), or emitted at the row's place
    // in the list with cursor ranges reserved for the whole block by one
    // atomic (EMIT true)
    {
        int sbase = 0;
        long long cbase = 0;
        if (EMIT) {
            const int nseg = valid ? segoff[j + 1] - segoff[j] : 0;
            int64_t tcur;
            // rows short enough for register cursors reserve none
            const int64_t need = alen > kRegGroups * 64 ? (int64_t)alen : 0;
            const int64_t xcur = block_excl_scan(need * nseg, &tcur);
            if (threadIdx.x == 0)
                s_cur = tcur ? (long long)atomicAdd(&ctr->cur_total, (unsigned long long)tcur) : 0;
            __syncthreads();
            if (valid) sbase = segoff[j];
            cbase = s_cur + xcur;
        }
        int s = 0, w0 = 0;
        double est = 0.0;
        // window output offsets, 16 bytes a load, 32 per round trip (the
        // next one, the first of the following batch, read alone)
        for (int wb = 0; wb < nwin; wb += 32) {
            int ow[33];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const I4 v = wb + 4 * q < nwin
                                 ? *reinterpret_cast<const I4 *>(g.win_out + c32 + wb + 4 * q)
                                 : I4{nz, nz, nz, nz};
                ow[4 * q] = v.x;
                ow[4 * q + 1] = v.y;
                ow[4 * q + 2] = v.z;
                ow[4 * q + 3] = v.w;
            }
            ow[32] = wb + 32 < nwin ? g.win_out[c32 + wb + 32] : nz;
#pragma unroll
            for (int u = 0; u < 33; ++u)
                if (wb + u >= nwin) ow[u] = nz;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int w = wb + u;
                if (w >= nwin) break;
                est += (ow[u + 1] - ow[u]) * ratio;
                if (est >= (double)g.seg_products || w == nwin - 1) {
                    if (EMIT) {
                        g.segs[sbase + s] = make_int4(row, w0, w + 1, 0);
                        g.seg_cur[sbase + s] = cbase + (long long)s * (alen > kRegGroups * 64 ? alen : 0);
                    }
                    ++s;
                    w0 = w + 1;
                    est = 0.0;
                }
            }
        }
        if (!EMIT && valid) segcnt[j] = s;
    }
}

// ---------------------------------------------------------------------------
// Numeric pass of big rows: one wave per segment (see the file comment).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_lower_bound(const int32_t *__restrict__ a, int lo, int hi,
                                                int32_t key)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Adds one batch of a piece (lanes hold consecutive entries of B_k, so
// distinct columns) and returns how many of its entries lie in the window.
// Branch-free apart from the add itself: every lane reads the dictionary
// (lanes outside the window at whatever slot their column masks to). A
// piece starts at its cursor, the first entry of B_k at or past the
// window's first column c0, so `col < c1` is the whole window test, and the
// dictionary holds word w of the columns at slot w mod MAXW (c0 is a
// multiple of 1024 columns and a window spans at most MAXW words, so its
// words take distinct slots and a column's bit index is its low five bits).
// The window's dictionary was built from the same B columns this pass
// reads, and B's rows reach the kernels canonical (checked on the host or
// the device before the analysis; the legacy dtypes' B rows are sorted on
// the host and repeated columns refused), so an in-window column is always
// in it. (Testing each product's dictionary bit and reporting misses cost
// 4 % of R-MAT 2^18's numeric pass.) Prefixes are clamped to PW when the
// dictionary is built and a word adds at most 31, so `pos` stays below
// PW + kAccPad, inside the window's accumulator, whatever the data.
constexpr int kAccPad = 32;

template <int PW, int MAXW, class Acc>
__device__ __forceinline__ int big_batch(Acc aj, int32_t col, Acc bv, int32_t c1,
                                         const uint32_t *W, const uint16_t *P, Acc *acc)
{
    static_assert((MAXW & (MAXW - 1)) == 0, "dictionary slots are word offsets mod MAXW");
    // the sentinel INT32_MAX past the end of B_k is out; the batch's
    // in-window entries are a prefix of it
    const bool inw = col < c1;
    const int nin = __popcll(__ballot(inw));  // taken while every lane is on
    const int wi = (int)(((uint32_t)col >> 5) & (uint32_t)(MAXW - 1));
    const uint32_t word = W[wi];
    // bits of the word below the column: v_bfe_u32 takes the width modulo 32
    int pos = (int)P[wi] + __popc(__builtin_amdgcn_ubfe(word, 0u, (uint32_t)col));
    // kept whole (v_bcnt_u32_b32 adds the prefix), so the slot address is
    // one shift-add instead of the compiler's redistributed shifts
    asm volatile("" : "+v"(pos));
    if (inw) lds_add(&acc[pos], vmul(aj, bv));
    return nin;
}

// Adds the products of one A entry's piece in [c0, c1) that start at `cj`
// (the first 64 entries already loaded in col/val); returns the new
// cursor and the column found there. All adds of a piece have the same k and distinct columns, so
// their order does not matter: a piece longer than 64 entries has its next
// kPieceBatches batches loaded at once before their adds.
constexpr int kPieceBatches = 4;

// Entries of a piece taken by its first load (lanes past it load nothing):
// pieces average about 30 entries, so a narrower first load fetches fewer
// lines past the window; a piece that fills it continues 64 at a time.
#ifndef RVSP_BIG_FIRST
#define RVSP_BIG_FIRST 64
#endif
constexpr int kFirstWidth = RVSP_BIG_FIRST;

template <int PW, int MAXW, class Acc, class Pair>
__device__ __forceinline__ int2 big_piece(int cj, int ej, Acc aj, int32_t col, Acc bv,
                                          int32_t c1, const uint32_t *W, const uint16_t *P,
                                          Acc *acc,
                                          const Pair *__restrict__ pairs)
{
    const int lane = threadIdx.x & 63;
    int nin = big_batch<PW, MAXW>(aj, col, bv, c1, W, P, acc);
    cj += nin;
    int32_t last = col;  // the batch whose first out-of-window entry ends the piece
    bool more = nin == kFirstWidth;  // uniform: the batch lay wholly in the window
    while (more) {
        int32_t cb[kPieceBatches];
        Acc vb[kPieceBatches];
#pragma unroll
        for (int b = 0; b < kPieceBatches; ++b) {
            const int bp = cj + b * 64 + lane;
            const Pair cv = bp < ej ? pairs[bp] : pair_none<Pair>();
            cb[b] = pair_col(cv);
            pair_value(cv, vb[b]);
        }
#pragma unroll
        for (int b = 0; b < kPieceBatches; ++b) {
            if (more) {
                nin = big_batch<PW, MAXW>(aj, cb[b], vb[b], c1, W, P, acc);
                cj += nin;
                last = cb[b];
                more = nin == 64;
            }
        }
    }
    // new cursor, and the column there (INT32_MAX past the end of B_k)
    return make_int2(cj, lane_bcast(last, nin));
}

// Live A entries of a window taken per memory round trip (R-MAT 2^18
// numeric pass: 4 / 8 / 16 give 26.3 / 25.8 / 25.7 ms; the pass is bound by
// the rate its scattered piece loads miss L2, not by loads in flight).
#ifndef RVSP_BIG_LIVE
#define RVSP_BIG_LIVE 8
#endif
constexpr int kLiveGroup = RVSP_BIG_LIVE;


// Up to kLiveGroup live A entries of a window, taken from a lane mask in lane
// (= ascending k) order, with the first 64 entries of each piece loaded
// from `pairs` (B's pairs, or their column-blocked copy).
template <class Acc, class Pair>
struct PieceGroup {
    int n;
    int j[kLiveGroup], c[kLiveGroup], e[kLiveGroup];
    Acc av[kLiveGroup], bv[kLiveGroup];
    int32_t col[kLiveGroup];

    __device__ __forceinline__ void take(uint64_t &mask, int4 ce, Acc a_val,
                                         const Pair *__restrict__ pairs)
    {
        const int lane = threadIdx.x & 63;
        n = 0;
#pragma unroll
        for (int u = 0; u < kLiveGroup; ++u) {
            j[u] = 0;
            if (mask) {
                j[u] = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                n = u + 1;
            }
        }
#pragma unroll
        for (int u = 0; u < kLiveGroup; ++u) {
            c[u] = lane_bcast(ce.x, j[u]);
            e[u] = lane_bcast(ce.y, j[u]);
            av[u] = lane_bcast(a_val, j[u]);
            const int bp = c[u] + lane;
            const bool ok = u < n && bp < e[u] && lane < kFirstWidth;
            const Pair cv = ok ? pairs[bp] : pair_none<Pair>();
            col[u] = pair_col(cv);
            pair_value(cv, bv[u]);
        }
    }
};

// Copies n elements from a 16-byte aligned LDS array to global memory
// aligned only to the element: whole 16-byte groups one store per lane (the
// target allows 16-byte global accesses at 4-byte alignment), the last one
// to three elements one at a time, nothing past n.
template <class T>
__device__ __forceinline__ void copy_out(T *__restrict__ dst, const T *src, int n)
{
    struct __attribute__((aligned(4))) Q {
        uint32_t x, y, z, w;
    };
    constexpr int E = 16 / sizeof(T);
    const int lane = threadIdx.x & 63;
    const int nq = n / E;
    for (int i = lane; i < nq; i += 64) {
        const uint4 q = reinterpret_cast<const uint4 *>(src)[i];
        reinterpret_cast<Q *>(dst)[i] = Q{q.x, q.y, q.z, q.w};
    }
    for (int i = nq * E + lane; i < n; i += 64) dst[i] = src[i];
}

// COLS: the pass also writes each window's column indices (g.cols_out),
// expanded from the rank dictionary into the accumulator's LDS once the
// values are out, so a one-shot needs no separate big-row structure fill.
template <int NW, int PW, int MAXCH, bool COLS, class Args>
__global__ __launch_bounds__(NW * 64) void k_num_big(int32_t nsegs, Args a, BigArgs g)
{
    using Acc = typename Args::acc_t;
    using Pair = typename Args::pair_t;
    constexpr int MAXW = MAXCH * kChunkWords;
    static_assert(PW * sizeof(Acc) % 16 == 0, "window accumulators are zeroed 16 bytes a store");
    __shared__ __align__(16) Acc acc_s[NW][PW + kAccPad];
    static_assert(MAXW % 256 == 0, "a lane's dictionary words are whole runs of four");
    __shared__ __align__(16) uint32_t W_s[NW][MAXW];
    __shared__ __align__(16) uint16_t P_s[NW][MAXW];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t it = (int64_t)blockIdx.x * NW + wv;
    if (it >= nsegs) return;  // whole wave leaves; no block barriers below
    Acc *acc = acc_s[wv];
    uint32_t *W = W_s[wv];
    uint16_t *P = P_s[wv];
    const int4 sg = g.segs[it];
    const int row = sg.x, w0 = sg.y, w1 = sg.z;
    int4 *cur_a = g.cursors + g.seg_cur[it];
    const int as = a.arp[row], ae = a.arp[row + 1];
    const int32_t rlo = big_base(g.lo[row]);
    const int nchunks = big_nchunks(g.lo[row], g.hi[row]);
    const Pair *__restrict__ pairs = a.bcv;
    const int64_t boff = g.big_off[row];
    const int64_t c32 = boff / kChunkWords;
    const int nwin = g.big_nwin[row];
    const int rowbase = a.crp[row];
    const int nz = a.crp[row + 1] - rowbase;

    constexpr int PER = MAXW / 64;
    // Rows of at most kRegGroups * 64 A entries keep each entry's state in
    // registers across the segment's windows (lane l holds entries l,
    // l + 64, ...): no cursor traffic per window. Longer rows keep it in HBM
    // (cur_a), read and written back per window.
    const bool regs = ae - as <= kRegGroups * 64;  // wave-uniform
    int rcur[kRegGroups], rend[kRegGroups], rcol[kRegGroups];
    Acc rav[kRegGroups];

    // per-entry state at the segment's first column: (cursor, end of B_k,
    // column at the cursor); four binary searches per lane run interleaved
    const int32_t cstart = rlo + g.win_chunk0[c32 + w0] * kChunkCols;
    for (int base = as; base < ae; base += 4 * 64) {
        int lo4[4], hi4[4], e4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ap = base + u * 64 + lane;
            lo4[u] = hi4[u] = e4[u] = 0;
            if (ap < ae) {
                const int k = a.aci[ap];
                lo4[u] = a.brp[k];
                e4[u] = a.brp[k + 1];
                hi4[u] = w0 > 0 ? e4[u] : lo4[u];
            }
        }
        for (;;) {
            bool any = false;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (lo4[u] < hi4[u]) {
                    any = true;
                    const int mid = (lo4[u] + hi4[u]) >> 1;
                    if (a.bci[mid] < cstart)
                        lo4[u] = mid + 1;
                    else
                        hi4[u] = mid;
                }
            }
            if (!any) break;
        }
        if (regs) {  // one round: at most kRegGroups * 64 <= 4 * 64 entries
#pragma unroll
            for (int u = 0; u < kRegGroups; ++u) {
                const int ap = base + u * 64 + lane;
                rcur[u] = lo4[u];
                rend[u] = e4[u];
                rcol[u] = ap < ae && lo4[u] < e4[u] ? a.bci[lo4[u]] : INT32_MAX;
                rav[u] = ap < ae ? Acc(a.aval[ap]) : Acc(0);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int ap = base + u * 64 + lane;
                const int32_t nc = ap < ae && lo4[u] < e4[u] ? a.bci[lo4[u]] : INT32_MAX;
                // the group's first cursor also holds the group's smallest column
                const int gmin = wave_min_all(nc);
                if (ap < ae) cur_a[ap - as] = make_int4(lo4[u], e4[u], nc, lane == 0 ? gmin : 0);
            }
        }
    }

    // window w: chunks [wchunk(w), wchunk(w + 1)), outputs [wout(w), wout(w + 1))
    auto wchunk = [&](int w) { return w < nwin ? g.win_chunk0[c32 + w] : nchunks; };
    auto wout = [&](int w) { return w < nwin ? g.win_out[c32 + w] : nz; };
    // up to MAXW bitmap words from chunk ch (lane l: words l * PER ..), past
    // the row's bitmap zero; words past the window are cleared at use
    // (a row's words start at a multiple of 32 in gbits and PER is a
    // multiple of 4, so each run of four is 16-byte aligned and wholly in or
    // out of the row)
    auto load_words = [&](uint32_t *dst, int ch) {
#pragma unroll
        for (int j = 0; j < PER; j += 4) {
            const int64_t i = (int64_t)ch * kChunkWords + lane * PER + j;
            const uint4 q = i < (int64_t)nchunks * kChunkWords
                                ? *reinterpret_cast<const uint4 *>(g.gbits + boff + i)
                                : make_uint4(0u, 0u, 0u, 0u);
            dst[j] = q.x;
            dst[j + 1] = q.y;
            dst[j + 2] = q.z;
            dst[j + 3] = q.w;
        }
    };
    int ch0 = wchunk(w0), o0 = wout(w0), ch1 = wchunk(w0 + 1), o1 = wout(w0 + 1);
    uint32_t wl[PER];
    load_words(wl, ch0);

    for (int w = w0; w < w1; ++w) {
        // the next window's bounds and bitmap words load while this one runs
        const int ch2 = wchunk(w + 2), o2 = wout(w + 2);
        uint32_t wn[PER];
        if (w + 1 < w1) load_words(wn, ch1);
        const int outs = o1 - o0;
        const int32_t c0 = rlo + ch0 * kChunkCols;
        const int64_t c1l = (int64_t)rlo + (int64_t)ch1 * kChunkCols;
        const int32_t c1 = c1l > INT32_MAX ? INT32_MAX : (int32_t)c1l;
        const int nw = (ch1 - ch0) * kChunkWords;  // <= MAXW
        // rank dictionary: words and their exclusive popcount prefix (clamped
        // to PW, which the window's outputs never exceed)
        int lsum = 0;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (lane * PER + j >= nw) wl[j] = 0u;
            lsum += __popc(wl[j]);
        }
        int run = wave_incl_scan_all(lsum) - lsum;
        // the lane's PER words and their prefixes, 16 (or 8) bytes a store
        uint32_t pp[PER / 2];
#pragma unroll
        for (int j = 0; j < PER; j += 2) {
            const uint32_t p0 = (uint32_t)min(run, PW);
            run += __popc(wl[j]);
            pp[j / 2] = p0 | ((uint32_t)min(run, PW) << 16);
            run += __popc(wl[j + 1]);
        }
        // word i of the window sits at slot (c0 / 32 + i) mod MAXW; c0 / 32 is
        // a multiple of 32, so a lane's PER words stay one aligned run
        const int slot0 = ((c0 >> 5) + lane * PER) & (MAXW - 1);
#pragma unroll
        for (int j = 0; j < PER; j += 4)
            *reinterpret_cast<uint4 *>(W + slot0 + j) = make_uint4(wl[j], wl[j + 1], wl[j + 2], wl[j + 3]);
        if constexpr (PER % 8 == 0) {
#pragma unroll
            for (int j = 0; j < PER / 2; j += 4)
                *reinterpret_cast<uint4 *>(P + slot0 + 2 * j) =
                    make_uint4(pp[j], pp[j + 1], pp[j + 2], pp[j + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < PER / 2; j += 2)
                *reinterpret_cast<uint2 *>(P + slot0 + 2 * j) = make_uint2(pp[j], pp[j + 1]);
        }
        // zero the window's outputs 16 bytes a store (PW * sizeof(Acc) is a
        // multiple of 16, so the rounded-up range stays in the accumulator)
        for (int i = lane; i < (int)((outs * sizeof(Acc) + 15) / 16); i += 64)
            reinterpret_cast<uint4 *>(acc)[i] = make_uint4(0u, 0u, 0u, 0u);
        wave_sync();

        if (regs) {
            // entry groups in A order, live entries of a group in lane order
#pragma unroll
            for (int u = 0; u < kRegGroups; ++u) {
                if (as + u * 64 >= ae) break;  // uniform
                uint64_t mask = __ballot(rcur[u] < rend[u] && rcol[u] < c1);
                while (mask) {
                    PieceGroup<Acc, Pair> g0;
                    g0.take(mask, make_int4(rcur[u], rend[u], 0, 0), rav[u], pairs);
#pragma unroll
                    for (int v = 0; v < kLiveGroup; ++v) {
                        if (v < g0.n) {
                            const int2 nc = big_piece<PW, MAXW>(g0.c[v], g0.e[v], g0.av[v],
                                                                g0.col[v], g0.bv[v], c1, W,
                                                                P, acc, pairs);
                            compiler_fence();
                            lane_set2(rcur[u], rcol[u], nc.x, nc.y, g0.j[v]);
                        }
                    }
                }
            }
        } else {
        // groups of 64 entries whose smallest next column lies before the
        // window's end, found 64 groups at a time; the others are skipped
        for (int gbase = as; gbase < ae; gbase += 64 * 64) {
            const int gp = gbase + lane * 64;
            const int gmin0 = gp < ae ? cur_a[gp - as].w : INT32_MAX;
            uint64_t groups = __ballot(gmin0 < c1);
            while (groups) {
                const int base = gbase + 64 * (__ffsll((unsigned long long)groups) - 1);
                groups &= groups - 1;
                const int ap = base + lane;
                int4 ce = make_int4(0, 0, INT32_MAX, 0);
                Acc av = Acc(0);
                if (ap < ae) {
                    ce = cur_a[ap - as];
                    av = Acc(a.aval[ap]);
                }
                // the cursor carries the column it points at: no load to test it
                const bool live = ce.x < ce.y && ce.z < c1;
                int cur = ce.x, nxt = ce.z;
                uint64_t mask = __ballot(live);
                while (mask) {
                    // up to kLiveGroup live entries: issue all their loads first
                    PieceGroup<Acc, Pair> g0;
                    g0.take(mask, ce, av, pairs);
#pragma unroll
                    for (int u = 0; u < kLiveGroup; ++u) {
                        if (u < g0.n) {
                            const int2 nc = big_piece<PW, MAXW>(g0.c[u], g0.e[u], g0.av[u],
                                                                g0.col[u], g0.bv[u], c1, W,
                                                                P, acc, pairs);
                            compiler_fence();
                            lane_set2(cur, nxt, nc.x, nc.y, g0.j[u]);
                        }
                    }
                }
                const int gmin = wave_min_all(nxt);  // lanes past the row hold INT32_MAX
                if (live || lane == 0) cur_a[ap - as] = make_int4(cur, ce.y, nxt, lane == 0 ? gmin : 0);
            }
        }
        }
        wave_sync();
        copy_out(a.cval + rowbase + o0, acc, outs);
        wave_sync();
        if (COLS) {
            // round j expands words j * 64 + lane, so a dense run of words
            // (the low columns of a skewed row) is shared by all lanes
            // instead of falling to the few lanes that hold it in the
            // dictionary's layout
            int32_t *stage = reinterpret_cast<int32_t *>(acc);
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int i = j * 64 + lane;
                const int si = ((c0 >> 5) + i) & (MAXW - 1);
                uint32_t x = W[si];  // zero past the window's words
                int pos = P[si];
                // a word's outputs lie in [0, outs), outs <= PW, for every
                // window the plan makes; one check per word keeps the writes
                // in the stage whatever the data
                if (pos + __popc(x) > PW) x = 0u;
                const int32_t cw = c0 + i * 32;
                while (x) {
                    const int b = __ffs(x) - 1;
                    x &= x - 1;
                    stage[pos++] = cw + b;
                }
            }
            wave_sync();
            copy_out(g.cols_out + rowbase + o0, stage, min(outs, PW));
            wave_sync();
        }
        ch0 = ch1;
        o0 = o1;
        ch1 = ch2;
        o1 = o2;
#pragma unroll
        for (int j = 0; j < PER; ++j) wl[j] = wn[j];
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int NT>
static void launch_sym_big_nt(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g)
{
    static bool raised = false;  // allow the largest pass once per process
    if (!raised) {
        (void)hipFuncSetAttribute((const void *)&k_sym_big<NT>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kBigPassWords * 4);
        raised = true;
    }
    // whole 64-word blocks for the LDS slot hash (kBigPassWords is one)
    const size_t lds_words = ((size_t)g.pass_words + 63) & ~(size_t)63;
    k_sym_big<NT><<<nrows, NT, lds_words * 4, st>>>(rows, nrows, g);
}

void launch_sym_big(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g)
{
    if (nrows <= 0) return;
    // the tuned block size (g.sym_threads); RVSP_SYM_BIG_THREADS overrides it
    const char *e = getenv("RVSP_SYM_BIG_THREADS");
    const int nt = e ? atoi(e) : g.sym_threads;
    if (nt == 1024)
        launch_sym_big_nt<1024>(st, rows, nrows, g);
    else if (nt == 256)
        launch_sym_big_nt<256>(st, rows, nrows, g);
    else
        launch_sym_big_nt<512>(st, rows, nrows, g);
}

template <int PW, int MAXCH>
static void launch_big_fill_shape(hipStream_t st, const BigArgs &g, const int32_t *crp,
                                  int32_t *cci)
{
    constexpr int NWV = 4;
    k_big_fill<NWV, PW, MAXCH><<<(g.nsegs + NWV - 1) / NWV, NWV * 64, 0, st>>>(g.nsegs, g, crp, cci);
}

void launch_big_fill(hipStream_t st, const BigArgs &g, const int32_t *crp, int32_t *cci)
{
    if (g.nsegs <= 0) return;
    // the shape must be one of kBigShapes (the plan only sets those)
    if (g.win_outs == 4096 && g.win_chunks == 32)
        launch_big_fill_shape<4096, 32>(st, g, crp, cci);
    else if (g.win_chunks == 32)
        launch_big_fill_shape<2048, 32>(st, g, crp, cci);
    else if (g.win_outs == 1024 && g.win_chunks == 8)
        launch_big_fill_shape<1024, 8>(st, g, crp, cci);
    else  // 2048 x 16 and 1024 x 16 (a 1024-column chunk never exceeds 2048 outputs)
        launch_big_fill_shape<2048, 16>(st, g, crp, cci);
}

void launch_big_plan(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g,
                     const int32_t *crp, AnalysisCounters *ctr, int32_t *segcnt, int32_t *segoff,
                     int64_t *tsum)
{
    if (nrows <= 0) return;
    const int grid = (nrows + 255) / 256;
    k_big_plan<false><<<grid, 256, 0, st>>>(rows, nrows, g, crp, ctr, segcnt, segoff);
    launch_scan(st, nrows, segcnt, tsum, (int64_t *)&ctr->seg_total, segoff);
    k_big_plan<true><<<grid, 256, 0, st>>>(rows, nrows, g, crp, ctr, segcnt, segoff);
}

template <int PW, int MAXCH, class Args>
static void launch_num_big_shape(hipStream_t st, int nsegs, const Args &a, const BigArgs &g)
{
    constexpr int NW = 2;
    if constexpr (std::is_same<Args, NumArgs>::value) {
        if (g.cols_out) {
            k_num_big<NW, PW, MAXCH, true><<<(nsegs + NW - 1) / NW, NW * 64, 0, st>>>(nsegs, a, g);
            return;
        }
    }
    k_num_big<NW, PW, MAXCH, false><<<(nsegs + NW - 1) / NW, NW * 64, 0, st>>>(nsegs, a, g);
}

template <class Args>
static void launch_num_big_any(hipStream_t st, int nsegs, const Args &a, const BigArgs &g)
{
    if (nsegs <= 0) return;
    // the shape must be one of kBigShapes (the plan only sets those)
    if (g.win_outs == 4096 && g.win_chunks == 32)
        launch_num_big_shape<4096, 32>(st, nsegs, a, g);
    else if (g.win_outs == 2048 && g.win_chunks == 32)
        launch_num_big_shape<2048, 32>(st, nsegs, a, g);
    else if (g.win_outs == 1024 && g.win_chunks == 16)
        launch_num_big_shape<1024, 16>(st, nsegs, a, g);
    else if (g.win_outs == 1024 && g.win_chunks == 8)
        launch_num_big_shape<1024, 8>(st, nsegs, a, g);
    else
        launch_num_big_shape<2048, 16>(st, nsegs, a, g);
}

void launch_num_big(hipStream_t st, int nsegs, const NumArgs &a, const BigArgs &g)
{
    launch_num_big_any(st, nsegs, a, g);
}

void launch_num_big(hipStream_t st, int nsegs, const NumArgsD &a, const BigArgs &g)
{
    launch_num_big_any(st, nsegs, a, g);
}

void launch_num_big(hipStream_t st, int nsegs, const NumArgsI &a, const BigArgs &g)
{
    launch_num_big_any(st, nsegs, a, g);
}

} // namespace rvsp
