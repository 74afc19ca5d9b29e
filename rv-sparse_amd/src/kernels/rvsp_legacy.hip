// rvsp_legacy.hip -- numeric pass and output compaction for rv-sparse's
// legacy dtypes: FP64 in and out (src/kernels/spgemm/legacy/
// csr_scalar_f64.c:176-230) and INT8 in with INT32 accumulation
// (csr_scalar_i8.c:138-186). Both keep the FP32 path's structure (the same
// analysis and fill) and its ascending-k accumulation; then, as those
// kernels do, entries whose value is exactly zero are dropped from C.
//
// Big rows run the FP32 big-row kernel instantiated for the value type
// (rvsp_bigrow.hip). Every other row runs k_num_generic: one wave per row,
// C_i's columns staged in LDS window by window (WIN outputs at a time), a
// product's output position found by binary search among them, one LDS add
// per product in k order (products rounded once, FP64 sums rounded once).
#include <type_traits>

#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

template <class Args, int WIN>
__global__ __launch_bounds__(256) void k_num_generic(const int32_t *__restrict__ rows,
                                                     int32_t nrows, Args a)
{
    using Acc = typename Args::acc_t;
    using Pair = typename Args::pair_t;
    constexpr int NW = 4;
    __shared__ int32_t cols_s[NW][WIN];
    __shared__ Acc acc_s[NW][WIN];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t it = (int64_t)blockIdx.x * NW + wv;
    if (it >= nrows) return;  // whole wave leaves; no block barriers below
    const int row = rows[it];
    const int as = a.arp[row], ae = a.arp[row + 1];
    const int cs = a.crp[row], nz = a.crp[row + 1] - cs;
    int32_t *Wc = cols_s[wv];
    Acc *acc = acc_s[wv];
    for (int o0 = 0; o0 < nz; o0 += WIN) {
        const int n = min(WIN, nz - o0);
        for (int i = lane; i < n; i += 64) {
            Wc[i] = a.cci[cs + o0 + i];
            acc[i] = Acc(0);
        }
        wave_sync();
        const int32_t clo = Wc[0], chi = Wc[n - 1];
        for (int base = as; base < ae; base += 64) {
            int bs = 0, be = 0;
            Acc av = Acc(0);
            if (base + lane < ae) {
                const int k = a.aci[base + lane];
                bs = a.brp[k];
                be = a.brp[k + 1];
                av = Acc(a.aval[base + lane]);
            }
            const int nch = min(64, ae - base);
            for (int j = 0; j < nch; ++j) {
                const int sj = lane_bcast(bs, j), ej = lane_bcast(be, j);
                const Acc aj = lane_bcast(av, j);
                for (int bp = sj + lane; bp < ej; bp += 64) {
                    const Pair pr = a.bcv[bp];
                    const int32_t col = pair_col(pr);
                    if (col < clo || col > chi) continue;  // another window's output
                    Acc bv;
                    pair_value(pr, bv);
                    int lo = 0, hi = n;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (Wc[mid] < col)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    if (lo < n && Wc[lo] == col)
                        atomicAdd(&acc[lo], vmul(aj, bv));
                    else
                        atomicOr(a.err, kErrProbeMiss);
                }
                compiler_fence();  // k-steps reach the accumulator in order
            }
        }
        wave_sync();
        for (int i = lane; i < n; i += 64) a.cval[cs + o0 + i] = acc[i];
        wave_sync();
    }
}

template <class In, class Pair>
__global__ __launch_bounds__(256) void k_pairs_typed(int32_t nnz, const int32_t *__restrict__ bci,
                                                     const In *__restrict__ bval,
                                                     Pair *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz) return;
    if constexpr (std::is_same<Pair, PairD>::value)
        out[i] = PairD{bci[i], 0, (double)bval[i]};
    else
        out[i] = make_int2(bci[i], (int)bval[i]);
}

// Kept entries (value != 0) per row, one wave per row.
template <class Acc>
__global__ __launch_bounds__(256) void k_nz_count(int32_t m, const int32_t *__restrict__ crp,
                                                  const Acc *__restrict__ cval,
                                                  int32_t *__restrict__ cnt)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wv;
    if (row >= m) return;
    const int s = crp[row], e = crp[row + 1];
    int c = 0;
    for (int i = s; i < e; i += 64) {
        const bool keep = i + lane < e && cval[i + lane] != Acc(0);
        c += __popcll(__ballot(keep));
    }
    if (lane == 0) cnt[row] = c;
}

// Stable compaction of the kept entries into the new row pointer's places.
template <class Acc>
__global__ __launch_bounds__(256) void k_nz_compact(int32_t m, const int32_t *__restrict__ crp,
                                                    const int32_t *__restrict__ cci,
                                                    const Acc *__restrict__ cval,
                                                    const int32_t *__restrict__ rp_out,
                                                    int32_t *__restrict__ ci_out,
                                                    Acc *__restrict__ cv_out)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wv;
    if (row >= m) return;
    const int s = crp[row], e = crp[row + 1];
    int dst = rp_out[row];
    const uint64_t below = (1ull << lane) - 1ull;
    for (int i = s; i < e; i += 64) {
        const bool in = i + lane < e;
        const Acc v = in ? cval[i + lane] : Acc(0);
        const bool keep = in && v != Acc(0);
        const uint64_t mask = __ballot(keep);
        if (keep) {
            const int p = dst + __popcll(mask & below);
            ci_out[p] = cci[i + lane];
            cv_out[p] = v;
        }
        dst += __popcll(mask);
    }
}

template <class Args>
static void launch_generic(hipStream_t st, const int32_t *rows, int nrows, const Args &a)
{
    if (nrows > 0)
        k_num_generic<Args, 1024><<<(nrows + 3) / 4, 256, 0, st>>>(rows, nrows, a);
}

void launch_numeric_typed(hipStream_t st, const NumArgsD &a, const int32_t *rows, int nrows,
                          const BigArgs &big, int nsegs)
{
    launch_generic(st, rows, nrows, a);
    launch_num_big(st, nsegs, a, big);
}

void launch_numeric_typed(hipStream_t st, const NumArgsI &a, const int32_t *rows, int nrows,
                          const BigArgs &big, int nsegs)
{
    launch_generic(st, rows, nrows, a);
    launch_num_big(st, nsegs, a, big);
}

void launch_pairs_typed(hipStream_t st, int32_t nnz, const int32_t *bci, const double *bval,
                        PairD *out)
{
    if (nnz > 0)
        k_pairs_typed<<<(int)(((int64_t)nnz + 255) / 256), 256, 0, st>>>(nnz, bci, bval, out);
}

void launch_pairs_typed(hipStream_t st, int32_t nnz, const int32_t *bci, const int8_t *bval,
                        int2 *out)
{
    if (nnz > 0)
        k_pairs_typed<<<(int)(((int64_t)nnz + 255) / 256), 256, 0, st>>>(nnz, bci, bval, out);
}

template <class Acc>
static void launch_drop_zeros_t(hipStream_t st, int32_t m, const int32_t *crp, const int32_t *cci,
                                const Acc *cval, int32_t *cnt, int64_t *tsum, int64_t *total,
                                int32_t *rp_out, int32_t *ci_out, Acc *cv_out)
{
    const int g = (int)(((int64_t)m + 3) / 4);
    k_nz_count<Acc><<<g, 256, 0, st>>>(m, crp, cval, cnt);
    launch_scan(st, m, cnt, tsum, total, rp_out);
    k_nz_compact<Acc><<<g, 256, 0, st>>>(m, crp, cci, cval, rp_out, ci_out, cv_out);
}

void launch_drop_zeros(hipStream_t st, int32_t m, const int32_t *crp, const int32_t *cci,
                       const double *cval, int32_t *cnt, int64_t *tsum, int64_t *total,
                       int32_t *rp_out, int32_t *ci_out, double *cv_out)
{
    launch_drop_zeros_t(st, m, crp, cci, cval, cnt, tsum, total, rp_out, ci_out, cv_out);
}

void launch_drop_zeros(hipStream_t st, int32_t m, const int32_t *crp, const int32_t *cci,
                       const int32_t *cval, int32_t *cnt, int64_t *tsum, int64_t *total,
                       int32_t *rp_out, int32_t *ci_out, int32_t *cv_out)
{
    launch_drop_zeros_t(st, m, crp, cci, cval, cnt, tsum, total, rp_out, ci_out, cv_out);
}

} // namespace rvsp
