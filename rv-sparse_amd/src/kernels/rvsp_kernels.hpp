// rvsp_kernels.hpp -- host-side declarations shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rvsp_device.hpp"

namespace rvsp {

// Counters the analysis kernels accumulate; copied to the host once per
// analysis phase.
struct AnalysisCounters {
    int sym_hist[16];
    int fb_hist[16];               // learned sizing: rows recounted, by heuristic bin
    int num_hist[64];
    unsigned long long products;
    unsigned long long bits_cap;   // sum over big rows of bitmap words
    unsigned long long bits_fill;  // running allocation of bitmap words
    unsigned long long cur_total;  // cursors over all big-row segments
    int moved;                     // learned sizing: rows in a smaller table
    long long nnz_total;           // scan total
    int n_long;                    // row statistics: rows left to the long-row pass
    long long seg_total;           // big-row segments (scan of the per-row counts)
    unsigned long long sym_ops[16];  // products per symbolic bin (stream balance)
    unsigned long long num_ops[64];  // products per numeric bin (stream balance)
};

// One block's totals of the row-statistics passes (summed into
// AnalysisCounters by a single block, instead of global atomics per block).
struct StatPart {
    int hist[16];
    unsigned long long prod, cap;
    unsigned long long ops[16];  // products per bin
};
constexpr int kStatsParts = 4096;  // partials a launch_row_stats call may write

// Big-row state: bitmaps, chunk counts, windows, segments, cursors.
struct BigArgs {
    const int32_t *arp, *aci, *brp, *bci;
    const int32_t *lo, *hi;
    const int64_t *ops;
    int32_t *cnt;
    int64_t *big_off;     // per row: offset of its bitmap words in gbits
    uint32_t *gbits;
    int32_t *gchunk;      // popcount per chunk (index big_off/32 + chunk)
    unsigned long long *bits_fill;
    int32_t *win_chunk0;  // first chunk of each window (same indexing)
    int32_t *win_out;     // first output (row-relative) of each window
    int32_t *big_nwin;    // per row: number of windows
    int4 *segs;           // (row, first window, end window, 0)
    int64_t *seg_cur;     // per segment: offset of its cursor block
    int4 *cursors;        // per (segment, A entry): (next B position, end of B row,
                          // column at that position or INT32_MAX, 0); the first cursor
                          // of a group of 64 keeps the group's smallest column in .w
    int pass_words;       // symbolic bitmap words per pass (LDS), multiple of kChunkWords
    int win_outs;         // numeric window: at most this many outputs ...
    int win_chunks;       // ... and this many chunks (one of the launchable shapes)
    int32_t nsegs;        // segments planned (set once the analysis has read the count)
    int seg_products;     // products per segment the plan aims at
    int sym_threads;      // count block size (256, 512 or 1024)
    int32_t *cols_out;    // numeric pass also writes the big rows' columns here (or null)
    // B's rows as (bitmap word, bits of the row in that word) pairs, one per
    // distinct 32-column word of a row, ascending: the big-row count ORs
    // whole words; bwp is their row pointer (k + 1 entries)
    const int2 *bwm;
    const int32_t *bwp;
};

// Launch parameters of one analysis, from the tuning table
// (rvsp_tuning.hip): they change how rows are binned and scheduled, never
// the result.
struct TuneParams {
    int big_window;           // index into kBigShapes
    int sym_big_threads;      // big-row count block: 256, 512 or 1024
    int seg_products;         // products per big-row segment
    int hash_max_u;           // symbolic hash rows: min(ops, span) <= this
    int64_t fan_min_products; // row bins on side streams from this many products
    int64_t fan_min_fill;     // ... and for the structure fill, outputs
};
// may_load = false: no file is read (the table if already loaded, else the
// defaults), for callers that must not allocate
TuneParams tune_for(int64_t nnz_a, int64_t nnz_b, int32_t k, int32_t n, bool may_load = true);

// Canonical-CSR check of a device operand (rvsp_validate.hip): sets
// kErrOperand in *err on a violation; `starts` holds nnz/32 + 1 words.
hipError_t launch_csr_validate(hipStream_t st, int32_t rows, int32_t cols, int32_t nnz,
                               const int32_t *rp, const int32_t *ci, uint32_t *starts, int *err);

// Launchable big-row window shapes (outputs, chunks); the first is the default.
constexpr int kBigShapes[][2] = {{2048, 16}, {4096, 32}, {2048, 32}, {1024, 16}, {1024, 8}};
constexpr int kBigShapeCount = sizeof(kBigShapes) / sizeof(kBigShapes[0]);

struct SymArgs {
    const int32_t *arp, *aci, *brp, *bci;
    const int32_t *lo, *hi;
    int32_t *cnt;
    const int32_t *crp;
    int32_t *cci;
    BigArgs big;
    int *err;
    const int64_t *ops;
    uint32_t *bms_spill;  // kBmsWords words per one-pass-bitmap row (its slot in the bin), or null
};

// Operands of a numeric pass: In = A's and B's value type, Acc = the
// accumulation and C value type, Pair = B's (column, value) pair.
template <class In, class Acc, class Pair>
struct NumArgsT {
    using in_t = In;
    using acc_t = Acc;
    using pair_t = Pair;
    const int32_t *arp, *aci;
    const In *aval;
    const int32_t *brp, *bci;
    const In *bval;
    const Pair *bcv;      // B as (column, value) pairs: one load per entry
    const int32_t *crp, *cci;
    Acc *cval;
    const int32_t *lo;
    int *err;
};
using NumArgs = NumArgsT<float, float, int2>;      // rv-sparse's FP32 contract
using NumArgsD = NumArgsT<double, double, PairD>;  // legacy FP64
using NumArgsI = NumArgsT<int8_t, int32_t, int2>;  // legacy INT8 -> INT32

// Independent bin launches spread over the caller's stream and up to three
// side streams of the device (default priority): forked from the caller's
// stream on construction, joined back to it on destruction. One scope per
// device at a time (a lock is held from fork to join, so host threads that
// share the side streams cannot cross their events); no fan-out while the
// caller's stream is being captured into a graph. RVSP_NUM_STREAMS (1-4,
// default 4) sets the streams in use.
class StreamFan {
public:
    StreamFan(hipStream_t st, bool want);
    ~StreamFan();
    StreamFan(const StreamFan &) = delete;
    StreamFan &operator=(const StreamFan &) = delete;
    // stream for the i-th independent launch (the side streams in turn)
    hipStream_t pick(int i) const { return n_ <= 1 ? main_ : side_[i % (n_ - 1)]; }
    // stream for a launch of `work` units: the one with the least work
    // queued so far (the caller's stream included), which then holds it
    hipStream_t take(unsigned long long work);
    // big-row work already queued on the caller's stream
    void preload(unsigned long long work) { load_[0] += work; }
    // streams for the bins [lo, hi) with rows, in launch order: with work
    // figures and no big rows queued, largest first by take(); else bin
    // order over the side streams in turn. Returns how many.
    int assign(int lo, int hi, const int32_t *count, const unsigned long long *work, int *order,
               hipStream_t *sts);

private:
    hipStream_t main_;
    int n_ = 1, dev_ = -1;
    hipStream_t side_[3] = {};
    hipEvent_t join_[3] = {};
    unsigned long long load_[4] = {};
};

// Bins [lo, hi) with rows, largest `work` first (launch order for
// StreamFan::take; with no work figures, bin order). Returns how many.
int bins_by_work(int lo, int hi, const int32_t *count, const unsigned long long *work, int *order);
// Passes with fewer products run on one stream: the fork and join (about
// 50 us) would cost more than the launches' overlap saves.
constexpr int64_t kFanMinProducts = int64_t(1) << 21;
// The structure fill writes nnz(C) columns; below this it runs on one stream.
constexpr int64_t kFanMinFill = int64_t(1) << 24;

// bst: k records of scratch, the B rows' (length, first column, last column)
// long_rows: scratch of m ints (rows of more than 512 A entries, reduced by
// a block each after the per-row pass); parts: kStatsParts partial totals
// Every index the statistics read through is clamped, so they may run on
// device operands whose check the host has not read yet.
void launch_row_stats(hipStream_t st, int32_t m, int32_t nnz_a, const int32_t *arp,
                      const int32_t *aci, int32_t k, int32_t nnz_b, const int32_t *brp,
                      const int32_t *bci, int4 *bst, int64_t *ops, int32_t *lo, int32_t *hi,
                      uint8_t *sbin, AnalysisCounters *ctr, int hash_max_u, int32_t *long_rows,
                      StatPart *parts);
void launch_bin_scatter(hipStream_t st, int32_t m, const uint8_t *bin, const int32_t *off,
                        int32_t *cursor, int32_t *perm, int nb);
// learned: hash bins use the overflow-checking count (a row whose count
// exceeds H/2 reports H + 1 instead of raising an error).
// work: products per symbolic bin (sym_ops), as for launch_numeric.
void launch_sym_count(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                      const int32_t *off, bool learned, bool fan = false,
                      const unsigned long long *work = nullptr);
// skip_big: the big rows' columns are left to a numeric pass with cols_out set.
void launch_sym_fill(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                     const int32_t *off, bool skip_big = false, bool fan = false,
                     const unsigned long long *work = nullptr);
void launch_scan(hipStream_t st, int32_t m, const int32_t *cnt, int64_t *tsum, int64_t *total,
                 int32_t *crp);
int scan_tiles(int32_t m);
// chooser: device parameters of the learned accumulator chooser, or null
// for the heuristic; lg_margin: log2 of the predicted speed-up the chooser
// needs to leave the heuristic's accumulator.
void launch_num_bins(hipStream_t st, int32_t m, const int32_t *crp, const int64_t *ops,
                     const int32_t *arp, const uint8_t *sbin, const int32_t *lo,
                     const int32_t *hi, const float *chooser, float lg_margin,
                     int64_t products, uint8_t *nbin, AnalysisCounters *ctr);
// work: products per numeric bin (the analysis counters' num_ops), which
// spread the bins over the streams largest first; null: bins in turn.
void launch_numeric(hipStream_t st, const NumArgs &a, const int32_t *perm, const int32_t *count,
                    const int32_t *off, const BigArgs &big, int nsegs, bool fan = false,
                    const unsigned long long *work = nullptr);
// Writes B's (column, value) pairs for a numeric pass: bcv[i] = (bci[i], bval[i]).
void launch_interleave(hipStream_t st, int32_t nnz, const int32_t *bci, const float *bval,
                       int2 *bcv);
// One hash or dense numeric bin over the given rows (false: not such a bin).
bool launch_num_bin(hipStream_t st, const NumArgs &a, int bin, const int32_t *rows, int nrows);
void launch_sym_big(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g);
// Reorders the big-row list largest first (buckets of a quarter octave of
// products); work holds kBigOrderWork ints, tmp nrows ints.
constexpr int kBigOrderWork = 512;
hipError_t launch_big_order(hipStream_t st, int32_t *rows, int nrows, const int64_t *ops,
                            int32_t *work, int32_t *tmp);
// B's (word, bits) pairs for the big-row count: flag (nnz words) and pidx
// (nnz + 1) are scratch, tsum holds scan_tiles(nnz) and total one int64;
// writes bwm (at most nnz pairs) and bwp (k + 1).
void launch_big_words(hipStream_t st, int32_t k, int32_t nnz, const int32_t *brp,
                      const int32_t *bci, int32_t *flag, int32_t *pidx, int64_t *tsum,
                      int64_t *total, int2 *bwm, int32_t *bwp);
// Big-row structure fill over every planned segment (g.nsegs).
void launch_big_fill(hipStream_t st, const BigArgs &g, const int32_t *crp, int32_t *cci);
// segcnt: nrows ints; segoff: nrows + 1 ints; tsum: scan tiles of nrows
void launch_big_plan(hipStream_t st, const int32_t *rows, int nrows, const BigArgs &g,
                     const int32_t *crp, AnalysisCounters *ctr, int32_t *segcnt, int32_t *segoff,
                     int64_t *tsum);
void launch_num_big(hipStream_t st, int nsegs, const NumArgs &a, const BigArgs &g);
void launch_num_big(hipStream_t st, int nsegs, const NumArgsD &a, const BigArgs &g);
void launch_num_big(hipStream_t st, int nsegs, const NumArgsI &a, const BigArgs &g);

// legacy dtypes (rvsp_legacy.hip): the numeric pass over the non-big rows
// (`rows`) and the big-row segments, B's pairs for the pass, and the
// compaction that drops entries equal to zero (counts, scan scratch and
// the new row pointer in rp_out, columns and values in ci_out / cv_out)
void launch_numeric_typed(hipStream_t st, const NumArgsD &a, const int32_t *rows, int nrows,
                          const BigArgs &big, int nsegs);
void launch_numeric_typed(hipStream_t st, const NumArgsI &a, const int32_t *rows, int nrows,
                          const BigArgs &big, int nsegs);
void launch_pairs_typed(hipStream_t st, int32_t nnz, const int32_t *bci, const double *bval,
                        PairD *out);
void launch_pairs_typed(hipStream_t st, int32_t nnz, const int32_t *bci, const int8_t *bval,
                        int2 *out);
void launch_drop_zeros(hipStream_t st, int32_t m, const int32_t *crp, const int32_t *cci,
                       const double *cval, int32_t *cnt, int64_t *tsum, int64_t *total,
                       int32_t *rp_out, int32_t *ci_out, double *cv_out);
void launch_drop_zeros(hipStream_t st, int32_t m, const int32_t *crp, const int32_t *cci,
                       const int32_t *cval, int32_t *cnt, int64_t *tsum, int64_t *total,
                       int32_t *rp_out, int32_t *ci_out, int32_t *cv_out);

// reused structure (rvsp_reuse.hip): per-product output ranks of the rows on
// the LDS paths (a byte map: 1- or 2-byte ranks per row); with pmap null only
// the per-row byte offsets and their total are made
void launch_pm_build(hipStream_t st, int32_t m, const uint8_t *nbin, const int64_t *ops,
                     const int32_t *arp, const int32_t *aci, const int32_t *brp,
                     const int32_t *bci, const int32_t *crp, const int32_t *cci, int64_t *cnt,
                     int64_t *tsum, int64_t *total, int64_t *prow, uint8_t *pmap, int *err);
int pm_scan_tiles(int32_t m);
void launch_numeric_ranked(hipStream_t st, const NumArgs &a, const int32_t *perm,
                           const int32_t *count, const int32_t *off, const BigArgs &big,
                           int nsegs, const int64_t *prow, const uint8_t *pmap,
                           bool fan = false, const unsigned long long *work = nullptr);

// learned selectors (rvsp_selector.hip)
const float *selector_device_params(int which, float *margin);
void launch_sel_bins(hipStream_t st, int32_t m, const float *params, float margin,
                     const int64_t *ops, const int32_t *arp, const int32_t *lo, const int32_t *hi,
                     uint8_t *sbin, uint8_t *hbin, AnalysisCounters *ctr);
void launch_sel_fallback(hipStream_t st, int32_t m, const int32_t *cnt, uint8_t *sbin,
                         const uint8_t *hbin, uint8_t *fbin, AnalysisCounters *ctr);
void launch_sel_features(hipStream_t st, int32_t m, const int64_t *ops, const int32_t *arp,
                         const int32_t *lo, const int32_t *hi, const int32_t *crp, float *feats,
                         float *ratio);

} // namespace rvsp
