/*
 * oracle.h -- CPU restatement of rv-sparse's FP32 scalar SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this code. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker or as the timed CPU baseline.
 *
 * Each function follows the rv-sparse source cited on it. Build flags:
 * -std=c11 -O3 -ffp-contract=off, so acc += a * b rounds the product and the
 * sum separately, as rv-sparse's -std=c11 build does (Makefile:74,82).
 */
#ifndef RVSP_ORACLE_H
#define RVSP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Canonical CSR check; 0 ok, 1 bad row_ptr, 2 column out of range,
 * 3 unsorted, 4 duplicate (src/kernels/spgemm/rvsp_support.c:36-126). */
int orc_csr_check(int32_t rows, int32_t cols, const int32_t *row_ptr,
                  const int32_t *col_idx, int32_t *bad_row_out);

/* Symbolic count: distinct columns per C row, per-row product counts,
 * exclusive scan into c_row_ptr[m+1]. Returns 0, or -6 when nnz(C) exceeds
 * INT32_MAX (src/kernels/spgemm/rvsp_symbolic.c:13-88). ops may be NULL. */
int orc_symbolic_count(int32_t m, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const int32_t *b_rp,
                       const int32_t *b_ci, int32_t *c_rp, int64_t *ops,
                       int64_t *total_out);

/* Symbolic fill: C's columns per row, ascending
 * (rvsp_symbolic.c:91-138, rvsp_sort.h:105-124). */
void orc_symbolic_fill(int32_t m, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const int32_t *b_rp,
                       const int32_t *b_ci, const int32_t *c_rp,
                       int32_t *c_ci);

/* FP32 numeric pass: acc starts at +0.0f; A entries in stored (ascending k)
 * order; acc[col] = acc[col] + a*b, product and sum rounded separately;
 * drained in C's column order (gustavson_core_f32.inc:18-64,
 * accum_scalar_f32.c:19-71, spgemm_descr.c:343). */
void orc_numeric_f32(int32_t m, int32_t n, const int32_t *a_rp,
                     const int32_t *a_ci, const float *a_val,
                     const int32_t *b_rp, const int32_t *b_ci,
                     const float *b_val, const int32_t *c_rp,
                     const int32_t *c_ci, float *c_val);

/* FP64 Gustavson on the same structure (products and sums in double), the
 * error yardstick; not part of the reference. */
void orc_numeric_f64(int32_t m, int32_t n, const int32_t *a_rp,
                     const int32_t *a_ci, const float *a_val,
                     const int32_t *b_rp, const int32_t *b_ci,
                     const float *b_val, const int32_t *c_rp,
                     const int32_t *c_ci, double *c_val);

/* Row-parallel orc_symbolic_count over `threads` OpenMP threads: same
 * outputs and return codes. */
int orc_symbolic_count_par(int32_t m, int32_t n, const int32_t *a_rp,
                           const int32_t *a_ci, const int32_t *b_rp,
                           const int32_t *b_ci, int32_t *c_rp, int64_t *ops,
                           int64_t *total_out, int threads);

/* Row-parallel orc_symbolic_fill + orc_numeric_f32 (+ orc_numeric_f64 into
 * c_val64 when it is not NULL) on a counted c_rp: same outputs, rows shared
 * over `threads` OpenMP threads. Returns 0, or -6 when scratch is short. */
int orc_fill_numeric_par(int32_t m, int32_t n, const int32_t *a_rp,
                         const int32_t *a_ci, const float *a_val,
                         const int32_t *b_rp, const int32_t *b_ci,
                         const float *b_val, const int32_t *c_rp,
                         int32_t *c_ci, float *c_val, double *c_val64,
                         int threads);

/* Full one-shot like rvsp_spgemm_scalar_f32 (gustavson_core_f32.inc:66-194):
 * count, malloc C, fill + sort, numeric. Outputs are malloc'd; release with
 * orc_free. Returns 0, -2 on bad sizes, -6 on allocation failure/overflow. */
int orc_spgemm_f32(int32_t m, int32_t k, int32_t n, const int32_t *a_rp,
                   const int32_t *a_ci, const float *a_val,
                   const int32_t *b_rp, const int32_t *b_ci,
                   const float *b_val, int32_t **c_rp_out,
                   int32_t **c_ci_out, float **c_val_out,
                   int32_t *c_nnz_out);

/* Row-parallel variant of orc_spgemm_f32 over `threads` OpenMP threads
 * (not the reference: the reference is single-threaded). Same outputs. */
int orc_spgemm_f32_omp(int32_t m, int32_t k, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const float *a_val,
                       const int32_t *b_rp, const int32_t *b_ci,
                       const float *b_val, int32_t **c_rp_out,
                       int32_t **c_ci_out, float **c_val_out,
                       int32_t *c_nnz_out, int threads);

void orc_free(void *p);

/* Wall-clock seconds (CLOCK_MONOTONIC), for the CPU baseline timer. */
double orc_now(void);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_ORACLE_H */
