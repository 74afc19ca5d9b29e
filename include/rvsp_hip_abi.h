/*
 * rvsp_hip_abi.h -- thin C ABI between the C host core (descriptor state,
 * dispatch, host-pointer API) and the HIP kernels for gfx950.
 *
 * Internal to the library: no counterpart in rv-sparse. It takes device
 * pointers and plain sizes only; HIP types cross it as void *. A "plan" is
 * the device half of one SpGEMM structure: per-row product counts, the row
 * bins the kernels are launched over, C's row pointer, and the scratch the
 * numeric pass needs. It borrows the operand structure pointers given to
 * rvsp_hip_plan_analyse until the next analyse or destroy.
 */
#ifndef RVSP_HIP_ABI_H
#define RVSP_HIP_ABI_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rvsp_hip_plan rvsp_hip_plan_t;

/* Row-bin selector used by the analysis. */
typedef enum
{
    RVSP_SELECT_HEURISTIC = 0,
    RVSP_SELECT_LEARNED = 1
} rvsp_select_t;

/* stream: an existing hipStream_t, or NULL to create (and own) one. */
rvsp_status_t rvsp_hip_plan_create(rvsp_hip_plan_t **plan, void *stream);
void rvsp_hip_plan_destroy(rvsp_hip_plan_t *plan);
void *rvsp_hip_plan_stream(const rvsp_hip_plan_t *plan);

/*
 * Symbolic analysis of C = A * B (A m x k, B k x n) on device-resident
 * structure. Fills per-row product counts, per-row nnz(C), C's row pointer
 * (device) and the numeric bins. Blocks until nnz(C) is known.
 * Returns RVSP_ERROR_ALLOCATION_FAILED when nnz(C) > INT32_MAX, and
 * RVSP_ERROR_INVALID_CSR, before any kernel reads through their indices,
 * when A or B is not canonical CSR (checked on the device unless
 * rvsp_hip_plan_operands_checked was called since the last analysis).
 */
/* The next rvsp_hip_plan_analyse skips its device canonical-CSR check: the
 * caller has checked the operands on the host (rvsp_csr_check). */
void rvsp_hip_plan_operands_checked(rvsp_hip_plan_t *plan);

rvsp_status_t rvsp_hip_plan_analyse(rvsp_hip_plan_t *plan, int32_t m,
                                    int32_t k, int32_t n, int32_t nnz_a,
                                    const int32_t *a_row_ptr,
                                    const int32_t *a_col_idx, int32_t nnz_b,
                                    const int32_t *b_row_ptr,
                                    const int32_t *b_col_idx,
                                    rvsp_select_t select, int64_t *c_nnz,
                                    int64_t *products);

/* Device pointer to C's row pointer (m+1 entries) of the last analysis. */
const int32_t *rvsp_hip_plan_row_ptr(const rvsp_hip_plan_t *plan);
/* Device pointer to the per-row product counts (m entries). */
const int64_t *rvsp_hip_plan_op_counts(const rvsp_hip_plan_t *plan);

/* Writes C's ascending column indices (device). Queued on the plan stream. */
rvsp_status_t rvsp_hip_plan_fill(rvsp_hip_plan_t *plan, int32_t *c_col_idx);

/* Numeric pass: C's values (device) from the stored structure in
 * c_col_idx. Queued on the plan stream. */
rvsp_status_t rvsp_hip_plan_numeric(rvsp_hip_plan_t *plan,
                                    const float *a_values,
                                    const float *b_values,
                                    const int32_t *c_col_idx,
                                    float *c_values);

/* Structure fill and numeric pass together, for a C whose column indices
 * are not yet written: the big rows' columns come out of the numeric pass,
 * which already holds each window's bitmap, instead of a separate fill.
 * Queued on the plan stream; same results as fill then numeric. */
rvsp_status_t rvsp_hip_plan_fill_numeric(rvsp_hip_plan_t *plan,
                                         const float *a_values,
                                         const float *b_values,
                                         int32_t *c_col_idx,
                                         float *c_values);

/* Numeric pass in one of rv-sparse's legacy dtypes, after the structure
 * fill into c_col_idx: in_dtype RVSP_DTYPE_FP64 (double values in and out)
 * or RVSP_DTYPE_INT8 (int8 values in, int32 accumulation and values out).
 * Same ascending-k accumulation as the FP32 pass (FP64 products and sums
 * each rounded once; INT32 wraps). Queued on the plan stream. */
rvsp_status_t rvsp_hip_plan_numeric_typed(rvsp_hip_plan_t *plan, rvsp_dtype_t in_dtype,
                                          const void *a_values, const void *b_values,
                                          const int32_t *c_col_idx, void *c_values);

/* Drops the entries of C whose value is exactly zero (value_dtype FP64 or
 * INT32), as rv-sparse's legacy kernels store only nonzero results: writes
 * the new row pointer (m + 1), columns and values into the device arrays
 * rp_out / ci_out / cv_out (sized like C) and the kept count into *nnz_out
 * (host; the call synchronises the plan stream). */
rvsp_status_t rvsp_hip_plan_drop_zeros(rvsp_hip_plan_t *plan, rvsp_dtype_t value_dtype,
                                       const int32_t *c_col_idx, const void *c_values,
                                       int32_t *rp_out, int32_t *ci_out, void *cv_out,
                                       int32_t *nnz_out);

/* Same numeric pass, captured once into a hipGraph and replayed while the
 * four pointers stay the same (re-captured when they change). */
rvsp_status_t rvsp_hip_plan_numeric_graph(rvsp_hip_plan_t *plan,
                                          const float *a_values,
                                          const float *b_values,
                                          const int32_t *c_col_idx,
                                          float *c_values);

/* Prepares numeric passes under a reused structure: computes, once per
 * analysis, the output rank of every product of the rows on the LDS paths
 * (1 byte per product for rows of at most 256 outputs, else 2, from C's
 * column indices in c_col_idx) so later
 * numeric passes need no column reads or probing. Skipped (the plain
 * kernels stay in use) when it would take more than a quarter of the free
 * device memory or when RVSP_REUSE_RANKS=0. Queued on the plan stream
 * after one synchronisation. Drops a captured numeric graph. */
rvsp_status_t rvsp_hip_plan_prepare_reuse(rvsp_hip_plan_t *plan, const int32_t *c_col_idx);

/* Checks the device-side overflow/consistency flag the kernels raise
 * (table overflow, probe exhaustion). RVSP_ERROR_INVALID_CSR if raised;
 * synchronises the plan stream. */
rvsp_status_t rvsp_hip_plan_check(rvsp_hip_plan_t *plan);

/* Host copies (m bytes each, either may be NULL) of the per-row symbolic
 * and numeric bins of the last analysis. Synchronous. */
rvsp_status_t rvsp_hip_plan_get_bins(rvsp_hip_plan_t *plan, uint8_t *sym_bins,
                                     uint8_t *num_bins);

/* Replaces the numeric bins (host array, m bytes). Empty and big-row bins
 * must stay as they are; every other row takes a hash or dense bin, which
 * must hold the row (else the numeric pass raises the consistency flag).
 * Drops a captured numeric graph. Synchronous. */
rvsp_status_t rvsp_hip_plan_set_num_bins(rvsp_hip_plan_t *plan, const uint8_t *num_bins);

/* Selector features of every row (device, m x 5 floats) and the sizer's
 * target log2(nnz(C_i) / min(ops_i, span_i)) (device, m floats, NaN for
 * rows without products). Synchronous. */
rvsp_status_t rvsp_hip_plan_features(rvsp_hip_plan_t *plan, float *feats, float *ratio);

/* Learned sizing of the last analysis: rows kept in a smaller table than
 * the heuristic's, and rows recounted after an overflow. */
rvsp_status_t rvsp_hip_plan_selector_stats(const rvsp_hip_plan_t *plan, int64_t *sized_rows,
                                           int64_t *fallback_rows);

/* Average milliseconds of one numeric launch of hash/dense bin `bin` over
 * the given device row list (rows may repeat), after one warm-up launch;
 * writes those rows' values. RVSP_ERROR_INVALID_ARGUMENT if the bin does
 * not hold a row. Synchronous. */
rvsp_status_t rvsp_hip_plan_time_bin(rvsp_hip_plan_t *plan, int32_t bin, const int32_t *rows,
                                     int32_t nrows, const float *a_values,
                                     const float *b_values, const int32_t *c_col_idx,
                                     float *c_values, int32_t reps, float *ms);

/* The stream the calling host thread's one-shot and batched calls run on:
 * one per thread, created on its first use and kept for the thread's
 * lifetime, so the library's block cache hands each call the memory the
 * thread's previous call freed (a fresh stream per call can make every call
 * map new memory). Threads do not queue behind each other's products. NULL
 * if no stream could be created. */
void *rvsp_hip_oneshot_stream(void);

/* Stream helpers used by the host core. */
rvsp_status_t rvsp_hip_stream_sync(void *stream);
rvsp_status_t rvsp_hip_copy_async(void *dst, const void *src, size_t bytes,
                                  void *stream);
rvsp_status_t rvsp_hip_alloc_async(void **ptr, size_t bytes, void *stream);
rvsp_status_t rvsp_hip_free_async(void *ptr, void *stream);

/* Batch packing (see rvsp_batch.hip). One item per stacked matrix: its
 * device arrays, its first row and first entry in the packed arrays, the
 * shift added to its column indices, and its own column count (its column
 * indices must lie in [0, cols)). */
typedef struct
{
    const int32_t *row_ptr;
    const int32_t *col_idx;
    const float *values;
    int32_t rows;
    int32_t nnz;
    int64_t row_off;
    int64_t nnz_off;
    int32_t col_shift;
    int32_t cols;
} rvsp_hip_pack_item_t;

/* Writes the stacked row pointer (rows+1 entries in total), column indices
 * and values of `count` matrices into out_*. Synchronises the stream.
 * Returns RVSP_ERROR_INVALID_CSR when an item's row pointer is not a valid
 * CSR row pointer over its nnz or a column lies outside [0, cols) (order
 * within rows is left to the analysis of the stacked product). */
rvsp_status_t rvsp_hip_pack(void *stream, int32_t count, const rvsp_hip_pack_item_t *items,
                            int32_t *out_rp, int32_t *out_ci, float *out_va);

/* Splits a stacked row pointer back into per-problem local row pointers,
 * written consecutively to out_rp (rows[i]+1 entries each), and returns
 * each problem's first entry in the stacked arrays in bases_host. */
rvsp_status_t rvsp_hip_unpack_row_ptr(void *stream, int32_t count, const int64_t *row_off,
                                      const int32_t *rows, const int32_t *big_rp,
                                      int32_t *out_rp, int64_t *bases_host);

/* Test hook: makes the next n device allocations fail, so cleanup paths
 * can be exercised without exhausting memory. */
void rvsp_hip_fail_next_allocs(int n);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_HIP_ABI_H */
