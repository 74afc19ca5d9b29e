/*
 * rvsp_spgemm.h -- semi-internal helpers of rv_sparse, MI355X build.
 *
 * rv-sparse exposes these from src/kernels/spgemm/rvsp_spgemm.h:34-121 and its
 * own bench (bench/bench.c) and matrix checker (bench/csr_check.c) call them.
 * The canonical check and the buffer-size query keep their signatures. The
 * raw one-shot kernels keep the RVSP_SPGEMM_PARAMS signature; this build
 * provides the HIP one, rvsp_spgemm_hip_f32.
 */
#ifndef RVSP_SPGEMM_H
#define RVSP_SPGEMM_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Replaces rvsp_spgemm_buffer_size (rvsp_spgemm.h:34,
 * src/kernels/spgemm/rvsp_support.c:11-29). The GPU kernels keep their
 * scratch in device memory owned by the call or the descriptor, so the host
 * workspace the caller passes is only checked for presence; the size
 * reported is a small constant, independent of b_cols, and positive.
 */
rvsp_status_t rvsp_spgemm_buffer_size(int32_t b_cols, size_t *bytes_out);

/* Canonical-CSR check results (rvsp_spgemm.h:38-45). */
typedef enum
{
    RVSP_CSR_OK = 0,
    RVSP_CSR_BAD_ROW_PTR,      /* row_ptr[0] != 0 or decreasing            */
    RVSP_CSR_COL_OUT_OF_RANGE, /* a column index outside [0, cols)         */
    RVSP_CSR_UNSORTED,         /* columns of a row not ascending           */
    RVSP_CSR_DUPLICATE         /* the same column twice in a row           */
} rvsp_csr_status_t;

/*
 * Replaces rvsp_csr_check (rvsp_spgemm.h:48-51,
 * src/kernels/spgemm/rvsp_support.c:36-126). O(nnz) host scan, no
 * allocation; the first violation found in row order is reported, and
 * bad_row_out (optional) receives its row, or -1 when not row specific.
 */
rvsp_csr_status_t rvsp_csr_check(int32_t rows,
                                 int32_t cols,
                                 const int32_t *row_ptr,
                                 const int32_t *col_idx,
                                 int32_t *bad_row_out);

/* Replaces rvsp_csr_status_string (rvsp_spgemm.h:53,
 * rvsp_support.c:128-151). */
const char *rvsp_csr_status_string(rvsp_csr_status_t status);

/*
 * Raw one-shot kernel signature (rvsp_spgemm.h:68-77). Output arrays are
 * malloc'd host memory owned by the caller; op_counts_out is optional
 * (a_rows entries, products per output row).
 */
#define RVSP_SPGEMM_PARAMS                                                   \
    int32_t a_rows, int32_t a_cols, int32_t b_cols,                          \
        const int32_t *a_row_ptr, const int32_t *a_col_idx,                  \
        const float *a_values, const int32_t *b_row_ptr,                     \
        const int32_t *b_col_idx, const float *b_values, void *workspace,    \
        size_t workspace_bytes, int32_t **c_row_ptr_out,                     \
        int32_t **c_col_idx_out, float **c_values_out, int32_t *c_nnz_out,   \
        int64_t *op_counts_out

/*
 * HIP counterpart of rvsp_spgemm_scalar_f32 (rvsp_spgemm.h:79,
 * gustavson_core_f32.inc:66-194): same arguments, same checks, same outputs
 * (row_ptr, ascending col_idx, values bit-identical to the scalar kernel).
 * A caller of the scalar raw kernel switches by name only.
 */
rvsp_status_t rvsp_spgemm_hip_f32(RVSP_SPGEMM_PARAMS);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_SPGEMM_H */
