// rvsp_bigrow.hip -- rows whose products span more than 16384 columns and
// exceed 2048 distinct-column candidates ("big rows").
//
// A big row's structure is a bitmap over its product column span, built
// once by the symbolic count and kept in HBM (one bit per column, plus a
// popcount per 1024-column chunk). Everything after the count reads that
// bitmap instead of walking products again:
//   fill     ordered compaction of the bitmap into C's column indices;
//   plan     the row is cut into windows of <= kBigPW outputs and at most
//            kBigMaxChunks chunks, and windows are grouped into segments;
//   numeric  one wave per segment. For each window the wave loads the
//            window's bitmap words into LDS with their exclusive popcount
//            prefix (a rank dictionary: column -> position in C_i), zeroes a
//            dense accumulator over the window's OUTPUT positions, walks the
//            row's A entries in stored order with one cursor per entry
//            (carried from window to window), and adds each product into
//            acc[rank(col)] with one LDS float add. One wave, entries in
//            ascending k, columns of one B row distinct: the same ordered,
//            separately rounded accumulation as rv-sparse's scalar kernel
//            (gustavson_core_f32.inc:31-63, accum_scalar_f32.c:67-70).
#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

// ---------------------------------------------------------------------------
// Symbolic count of big rows: one block per row, LDS bitmap of 2^19
// columns per pass. B_k's entries are sorted, so a wave's 64 consecutive
// entries fall in runs of equal bitmap words: each run is OR-combined with
// shuffles and written with one LDS atomic.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(NT) void k_sym_big(const int32_t *__restrict__ rows, int32_t nrows,
                                                BigArgs g)
{
    __shared__ uint32_t bits[kBigPassWords];
    __shared__ long long s_off;
    __shared__ int s_part[NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row = rows[blockIdx.x];
    const int32_t rlo = g.lo[row], rhi = g.hi[row];
    const int nchunks = (int)(((int64_t)rhi - rlo + kChunkCols) / kChunkCols);
    const int nwords = nchunks * kChunkWords;
    if (tid == 0) {
        s_off = (long long)atomicAdd(g.bits_fill, (unsigned long long)nwords);
        g.big_off[row] = s_off;
    }
    __syncthreads();
    const int64_t off = s_off;
    const int as = g.arp[row], ae = g.arp[row + 1];
    int total = 0;
    for (int pw0 = 0; pw0 < nwords; pw0 += kBigPassWords) {
        const int pwn = min(kBigPassWords, nwords - pw0);
        const int32_t base_col = rlo + pw0 * 32;
        const int32_t pcols = pwn * 32;
        for (int w = tid; w < pwn; w += NT) bits[w] = 0u;
        __syncthreads();
        // A entries in chunks of NT: with short B rows each thread walks its
        // own B row (order does not matter for the structure); with long
        // ones the block covers one B row at a time.
        for (int cb = as; cb < ae; cb += NT) {
            const int myap = cb + tid;
            int mbs = 0, mbe = 0;
            if (myap < ae) {
                const int k = g.aci[myap];
                mbs = g.brp[k];
                mbe = g.brp[k + 1];
            }
            const int cnt = min(NT, ae - cb);
            int len = mbe - mbs;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) len += __shfl_xor(len, o, 64);
            if (lane == 0) s_part[wv] = len;
            __syncthreads();
            int tot = 0;
#pragma unroll
            for (int i = 0; i < NT / 64; ++i) tot += s_part[i];
            __syncthreads();
            if (tot < 64 * cnt) {
                for (int bp = mbs; bp < mbe; ++bp) {
                    const int32_t c = g.bci[bp] - base_col;
                    if (c >= 0 && c < pcols) atomicOr(&bits[c >> 5], 1u << (c & 31));
                }
                continue;
            }
            for (int ap = cb; ap < cb + cnt; ++ap) {
                const int k = g.aci[ap];
                const int bs = g.brp[k], be = g.brp[k + 1];
                // 4 entries per thread in flight, then OR-combine each slot
                for (int base = bs; base < be; base += 4 * NT) {
                    int32_t col[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int bp = base + u * NT + tid;
                        col[u] = bp < be ? g.bci[bp] : -1;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        if (base + u * NT >= be) break;  // block-uniform
                        const bool valid = base + u * NT + tid < be;
                        const int32_t c = col[u] - base_col;
                        const bool in = valid && c >= 0 && c < pcols;
                        int wd = in ? (c >> 5) : -1 - lane;
                        uint32_t wb = in ? (1u << (c & 31)) : 0u;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const uint32_t ob = __shfl_up(wb, o, 64);
                            const int ow = __shfl_up(wd, o, 64);
                            if (lane >= o && ow == wd) wb |= ob;
                        }
                        const int nwd = __shfl_down(wd, 1, 64);
                        const bool last = lane == 63 || nwd != wd;
                        if (in && last) atomicOr(&bits[wd], wb);
                    }
                }
            }
        }
        __syncthreads();
        for (int w = tid; w < pwn; w += NT) g.gbits[off + pw0 + w] = bits[w];
        const int nch = pwn / kChunkWords;
        int part = 0;
        for (int ch = tid; ch < nch; ch += NT) {
            int s = 0;
#pragma unroll 8
            for (int w = 0; w < kChunkWords; ++w) s += __popc(bits[ch * kChunkWords + w]);
            g.gchunk[off / kChunkWords + pw0 / kChunkWords + ch] = s;
            part += s;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if (lane == 0) s_part[wv] = part;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) total += s_part[i];
        __syncthreads();
    }
    if (tid == 0) g.cnt[row] = total;
}

// ---------------------------------------------------------------------------
// Structure fill of big rows from their bitmaps: one block per row, each
// thread owns 8 consecutive words of a 2048-word tile.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_big_fill(const int32_t *__restrict__ rows, int32_t nrows,
                                                  BigArgs g, const int32_t *__restrict__ crp,
                                                  int32_t *__restrict__ cci)
{
    constexpr int PER = 8;
    const int row = rows[blockIdx.x];
    const int32_t rlo = g.lo[row], rhi = g.hi[row];
    const int nwords = (int)(((int64_t)rhi - rlo + kChunkCols) / kChunkCols) * kChunkWords;
    const int64_t off = g.big_off[row];
    int out = crp[row];
    for (int t0 = 0; t0 < nwords; t0 += 256 * PER) {
        uint32_t w[PER];
        int local = 0;
        const int wbase = t0 + threadIdx.x * PER;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            w[j] = wbase + j < nwords ? g.gbits[off + wbase + j] : 0u;
            local += __popc(w[j]);
        }
        int64_t tot;
        int pos = out + (int)block_excl_scan(local, &tot);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            uint32_t x = w[j];
            const int32_t cb = rlo + (wbase + j) * 32;
            while (x) {
                const int b = __ffs(x) - 1;
                x &= x - 1;
                cci[pos++] = cb + b;
            }
        }
        out += (int)tot;
    }
}

// ---------------------------------------------------------------------------
// Window and segment plan: one thread per big row. Windows are runs of
// whole chunks with at most kBigPW outputs and kBigMaxChunks chunks.
// Segments are runs of windows holding about kSegProducts products
// (estimated from the row's products per output), each processed by one
// wave, so the widest rows spread over many waves; their cursor blocks
// (alen entries each) are carved from a running counter.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_big_plan(const int32_t *__restrict__ rows, int32_t nrows,
                                                  BigArgs g, const int32_t *__restrict__ crp,
                                                  AnalysisCounters *__restrict__ ctr)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= nrows) return;
    const int row = rows[j];
    const int32_t rlo = g.lo[row], rhi = g.hi[row];
    const int nchunks = (int)(((int64_t)rhi - rlo + kChunkCols) / kChunkCols);
    const int64_t c32 = g.big_off[row] / kChunkWords;
    const int alen = g.arp[row + 1] - g.arp[row];
    const int nz = crp[row + 1] - crp[row];
    // products per output of this row: segments aim at ~kSegProducts each
    const double ratio = nz > 0 ? (double)g.ops[row] / (double)nz : 1.0;
    int nwin = 0, out = 0, acc = 0, start = 0;
    // pass 1: windows
    for (int c = 0; c < nchunks; ++c) {
        const int cc = g.gchunk[c32 + c];
        if (c > start && (acc + cc > kBigPW || c - start >= kBigMaxChunks)) {
            g.win_chunk0[c32 + nwin] = start;
            g.win_out[c32 + nwin] = out;
            ++nwin;
            out += acc;
            start = c;
            acc = 0;
        }
        acc += cc;
    }
    g.win_chunk0[c32 + nwin] = start;
    g.win_out[c32 + nwin] = out;
    ++nwin;
    g.big_nwin[row] = nwin;
    // pass 2: segments (count, then emit)
    int nseg = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int sbase = 0;
        long long cbase = 0;
        if (pass == 1) {
            sbase = atomicAdd(&ctr->nseg, nseg);
            cbase = (long long)atomicAdd(&ctr->cur_total,
                                         (unsigned long long)alen * (unsigned long long)nseg);
        }
        int s = 0, w0 = 0;
        double est = 0.0;
        for (int w = 0; w < nwin; ++w) {
            const int o1 = w + 1 < nwin ? g.win_out[c32 + w + 1] : nz;
            est += (o1 - g.win_out[c32 + w]) * ratio;
            if (est >= (double)kSegProducts || w == nwin - 1) {
                if (pass == 1) {
                    g.segs[sbase + s] = make_int4(row, w0, w + 1, 0);
                    g.seg_cur[sbase + s] = cbase + (long long)s * alen;
                }
                ++s;
                w0 = w + 1;
                est = 0.0;
            }
        }
        nseg = s;
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

// Adds the products of one A entry's piece in [c0, c1) that start at `cj`
// (loaded values already in col/val for the first 64 entries); returns the
// new cursor. Lanes cover consecutive entries of B_k, so columns are
// distinct within an instruction.
__device__ __forceinline__ int big_piece(int cj, int ej, float aj, int32_t col, float bv,
                                         int32_t c0, int32_t c1, int outs, const uint32_t *W,
                                         const uint16_t *P, float *acc, const NumArgs &a)
{
    const int lane = threadIdx.x & 63;
    for (;;) {
        const bool in = col < c1;
        if (in) {
            const int rel = col - c0;
            if (rel >= 0) {
                const int wi = rel >> 5;
                const uint32_t word = W[wi];
                const int pos = (int)P[wi] + __popc(word & ((1u << (rel & 31)) - 1u));
                if (((word >> (rel & 31)) & 1u) && pos < outs)
                    atomicAdd(&acc[pos], __fmul_rn(aj, bv));
                else
                    atomicOr(a.err, kErrProbeMiss);
            } else {
                atomicOr(a.err, kErrWindow);
            }
        }
        const int nin = __popcll(__ballot(in));
        cj += nin;
        if (nin < 64) return cj;
        const int bp = cj + lane;  // piece continues past 64 entries
        col = bp < ej ? a.bci[bp] : INT32_MAX;
        bv = bp < ej ? a.bval[bp] : 0.0f;
    }
}

template <int NW>
__global__ __launch_bounds__(NW * 64) void k_num_big(int32_t nsegs, NumArgs a, BigArgs g)
{
    __shared__ float acc_s[NW][kBigPW];
    __shared__ uint32_t W_s[NW][kBigMaxWords];
    __shared__ uint16_t P_s[NW][kBigMaxWords];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t it = (int64_t)blockIdx.x * NW + wv;
    if (it >= nsegs) return;  // whole wave leaves; no block barriers below
    float *acc = acc_s[wv];
    uint32_t *W = W_s[wv];
    uint16_t *P = P_s[wv];
    const int4 sg = g.segs[it];
    const int row = sg.x, w0 = sg.y, w1 = sg.z;
    int2 *cur_a = g.cursors + g.seg_cur[it];
    const int as = a.arp[row], ae = a.arp[row + 1];
    const int32_t rlo = g.lo[row], rhi = g.hi[row];
    const int nchunks = (int)(((int64_t)rhi - rlo + kChunkCols) / kChunkCols);
    const int64_t boff = g.big_off[row];
    const int64_t c32 = boff / kChunkWords;
    const int nwin = g.big_nwin[row];
    const int rowbase = a.crp[row];
    const int nz = a.crp[row + 1] - rowbase;

    // per-entry state at the segment's first column: (cursor, end of B_k);
    // four binary searches per lane run interleaved
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
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int ap = base + u * 64 + lane;
            if (ap < ae) cur_a[ap - as] = make_int2(lo4[u], e4[u]);
        }
    }

    for (int w = w0; w < w1; ++w) {
        const int ch0 = g.win_chunk0[c32 + w];
        const int ch1 = w + 1 < nwin ? g.win_chunk0[c32 + w + 1] : nchunks;
        const int o0 = g.win_out[c32 + w];
        const int o1 = w + 1 < nwin ? g.win_out[c32 + w + 1] : nz;
        const int outs = o1 - o0;
        const int32_t c0 = rlo + ch0 * kChunkCols;
        const int64_t c1l = (int64_t)rlo + (int64_t)ch1 * kChunkCols;
        const int32_t c1 = c1l > INT32_MAX ? INT32_MAX : (int32_t)c1l;
        const int nw = (ch1 - ch0) * kChunkWords;  // <= kBigMaxWords
        // rank dictionary: words and their exclusive popcount prefix
        constexpr int PER = kBigMaxWords / 64;
        int lsum = 0;
        uint32_t wl[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = lane * PER + j;
            wl[j] = i < nw ? g.gbits[boff + (int64_t)ch0 * kChunkWords + i] : 0u;
            lsum += __popc(wl[j]);
        }
        int incl = lsum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        int run = incl - lsum;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = lane * PER + j;
            W[i] = wl[j];
            P[i] = (uint16_t)run;
            run += __popc(wl[j]);
        }
        for (int i = lane; i < outs; i += 64) acc[i] = 0.0f;
        wave_sync();

        for (int base = as; base < ae; base += 64) {
            const int ap = base + lane;
            int2 ce = make_int2(0, 0);
            float av = 0.0f;
            bool live = false;
            if (ap < ae) {
                ce = cur_a[ap - as];
                av = a.aval[ap];
                live = ce.x < ce.y && a.bci[ce.x] < c1;
            }
            int cur = ce.x;
            uint64_t mask = __ballot(live);
            while (mask) {
                // up to four live entries: issue all their loads first
                int js[4], cs[4], es[4];
                float as4[4], bvs[4];
                int32_t cols[4];
                int nj = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    js[u] = -1;
                    if (mask) {
                        js[u] = __ffsll((unsigned long long)mask) - 1;
                        mask &= mask - 1;
                        nj = u + 1;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int ju = js[u] < 0 ? 0 : js[u];
                    cs[u] = __shfl(ce.x, ju, 64);
                    es[u] = __shfl(ce.y, ju, 64);
                    as4[u] = __shfl(av, ju, 64);
                    const int bp = cs[u] + lane;
                    const bool ok = u < nj && bp < es[u];
                    cols[u] = ok ? a.bci[bp] : INT32_MAX;
                    bvs[u] = ok ? a.bval[bp] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (u < nj) {
                        const int nc = big_piece(cs[u], es[u], as4[u], cols[u], bvs[u], c0, c1, outs,
                                                 W, P, acc, a);
                        compiler_fence();
                        if (lane == js[u]) cur = nc;
                    }
                }
            }
            if (ap < ae) cur_a[ap - as].x = cur;
        }
        wave_sync();
        for (int i = lane; i < outs; i += 64) a.cval[rowbase + o0 + i] = acc[i];
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_sym_big(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g)
{
    if (nrows > 0) k_sym_big<256><<<nrows, 256, 0, st>>>(rows, nrows, g);
}

void launch_big_fill(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g,
                     const int32_t *crp, int32_t *cci)
{
    if (nrows > 0) k_big_fill<<<nrows, 256, 0, st>>>(rows, nrows, g, crp, cci);
}

void launch_big_plan(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g,
                     const int32_t *crp, AnalysisCounters *ctr)
{
    if (nrows > 0) k_big_plan<<<(nrows + 255) / 256, 256, 0, st>>>(rows, nrows, g, crp, ctr);
}

void launch_num_big(hipStream_t st, int nsegs, const NumArgs &a, const BigArgs &g)
{
    constexpr int NW = 2;
    if (nsegs > 0) k_num_big<NW><<<(nsegs + NW - 1) / NW, NW * 64, 0, st>>>(nsegs, a, g);
}

} // namespace rvsp
