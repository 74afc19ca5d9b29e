/*
 * spgemm.c -- one-shot SpGEMM entry points and backend dispatch.
 *
 * rvsp_spgemm_csr keeps the dispatch rules of rv-sparse
 * src/core/spgemm.c:135-211 (options NULL = SCALAR/FP32/FP32; SCALAR and
 * SCALAR_UNROLL4 alias; RVV routes on the operands' dtype) and the one-shot
 * sequence of spgemm.c:26-133 (dtype and shape checks, canonical check,
 * descriptor estimate, library-allocated C, compute, cleanup on failure).
 * FP32 SCALAR/SCALAR_UNROLL4/HIP run on the GPU. RVV is unsupported here,
 * as in any non-vector rv-sparse build. The legacy dtypes (FP64 in and out,
 * INT8 in with INT32 out; rv-sparse spgemm.c:182-194) run on the GPU too,
 * through spgemm_legacy.c; their RVV variants stay unsupported.
 */
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"
#include "rv_sparse_hip.h"
#include "rvsp_hip_abi.h"
#include "rvsp_spgemm.h"

/* spgemm_descr.c */
void rvsp_descr_operands_checked(rvsp_spgemm_descr_t d);

/* spgemm_legacy.c */
rvsp_status_t rvsp_spgemm_legacy_hip(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                     rvsp_csr_matrix_t *C, rvsp_dtype_t in);

static void reset_c(rvsp_csr_matrix_t *C)
{
    C->row_ptr = NULL;
    C->col_idx = NULL;
    C->values = NULL;
    C->nnz = 0;
    C->owns_data = 0;
}

static rvsp_status_t oneshot_host(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                  rvsp_csr_matrix_t *C, rvsp_spgemm_algo_t algo)
{
    if (A->dtype != RVSP_DTYPE_FP32 || B->dtype != RVSP_DTYPE_FP32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (A->cols != B->rows)
        return RVSP_ERROR_INVALID_ARGUMENT;
    /* the kernels trust canonical CSR, so check once up front */
    if (rvsp_csr_check(A->rows, A->cols, A->row_ptr, A->col_idx, NULL) != RVSP_CSR_OK ||
        rvsp_csr_check(B->rows, B->cols, B->row_ptr, B->col_idx, NULL) != RVSP_CSR_OK)
        return RVSP_ERROR_INVALID_CSR;

    rvsp_spgemm_descr_t d = NULL;
    rvsp_status_t s = rvsp_spgemm_descr_create(&d);
    if (s != RVSP_SUCCESS)
        return s;
    s = rvsp_spgemm_set_algo(d, algo);
    size_t ws_bytes = 0;
    int32_t c_nnz = 0;
    void *ost = rvsp_hip_oneshot_stream();
    if (s == RVSP_SUCCESS && ost)
        s = rvsp_spgemm_set_stream(d, ost);
    rvsp_descr_operands_checked(d);
    if (s == RVSP_SUCCESS)
        s = rvsp_spgemm_work_estimation(d, A, B, &ws_bytes, &c_nnz);
    if (s != RVSP_SUCCESS)
    {
        rvsp_spgemm_descr_destroy(d);
        return s;
    }

    const size_t cap = c_nnz > 0 ? (size_t)c_nnz : 1;
    int32_t *rp = (int32_t *)malloc(((size_t)A->rows + 1) * sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc(cap * sizeof(int32_t));
    float *val = (float *)malloc(cap * sizeof(float));
    void *ws = malloc(ws_bytes);
    if (!rp || !ci || !val || !ws)
    {
        free(rp);
        free(ci);
        free(val);
        free(ws);
        rvsp_spgemm_descr_destroy(d);
        return RVSP_ERROR_ALLOCATION_FAILED;
    }

    C->rows = A->rows;
    C->cols = B->cols;
    C->nnz = c_nnz;
    C->row_ptr = rp;
    C->col_idx = ci;
    C->values = val;
    C->dtype = RVSP_DTYPE_FP32;
    C->format = RVSP_FORMAT_CSR;
    C->owns_data = 1;

    s = rvsp_spgemm_compute(d, A, B, C, ws);
    free(ws);
    rvsp_spgemm_descr_destroy(d);
    if (s != RVSP_SUCCESS)
    {
        free(rp);
        free(ci);
        free(val);
        reset_c(C);
    }
    return s;
}

rvsp_status_t rvsp_spgemm_csr(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                              rvsp_csr_matrix_t *C, const rvsp_spgemm_options_t *options)
{
    if (!A || !B || !C)
        return RVSP_ERROR_NULL_POINTER;
    rvsp_backend_t backend = RVSP_BACKEND_SCALAR;
    rvsp_dtype_t in = RVSP_DTYPE_FP32, out = RVSP_DTYPE_FP32;
    if (options)
    {
        backend = options->backend;
        in = options->input_dtype;
        out = options->output_dtype;
    }
    if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_SCALAR_UNROLL4 ||
         backend == RVSP_BACKEND_HIP) &&
        in == RVSP_DTYPE_FP32 && out == RVSP_DTYPE_FP32)
        return oneshot_host(A, B, C, RVSP_SPGEMM_ALGO_DEFAULT);
    if (backend == RVSP_BACKEND_RVV_INTRINSICS && A->dtype == RVSP_DTYPE_FP32 &&
        B->dtype == RVSP_DTYPE_FP32)
        return oneshot_host(A, B, C, RVSP_SPGEMM_ALGO_RVV); /* unsupported here */
    /* legacy dtypes (spgemm.c:182-194): the scalar backend, or HIP by name */
    if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_HIP) &&
        in == RVSP_DTYPE_INT8 && out == RVSP_DTYPE_INT32)
        return rvsp_spgemm_legacy_hip(A, B, C, RVSP_DTYPE_INT8);
    if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_HIP) &&
        in == RVSP_DTYPE_FP64 && out == RVSP_DTYPE_FP64)
        return rvsp_spgemm_legacy_hip(A, B, C, RVSP_DTYPE_FP64);
    return RVSP_ERROR_UNSUPPORTED_BACKEND;
}

rvsp_status_t rvsp_spgemm_hip_f32(int32_t a_rows, int32_t a_cols, int32_t b_cols,
                                  const int32_t *a_row_ptr, const int32_t *a_col_idx,
                                  const float *a_values, const int32_t *b_row_ptr,
                                  const int32_t *b_col_idx, const float *b_values,
                                  void *workspace, size_t workspace_bytes,
                                  int32_t **c_row_ptr_out, int32_t **c_col_idx_out,
                                  float **c_values_out, int32_t *c_nnz_out,
                                  int64_t *op_counts_out)
{
    /* checks of rv-sparse gustavson_core_f32.inc:76-101 */
    if (!a_row_ptr || !a_col_idx || !a_values || !b_row_ptr || !b_col_idx || !b_values ||
        !c_row_ptr_out || !c_col_idx_out || !c_values_out || !c_nnz_out || !workspace)
        return RVSP_ERROR_NULL_POINTER;
    if (a_rows <= 0 || a_cols <= 0 || b_cols <= 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    size_t need = 0;
    rvsp_spgemm_buffer_size(b_cols, &need);
    if (workspace_bytes < need)
        return RVSP_ERROR_INVALID_ARGUMENT;
    *c_row_ptr_out = NULL;
    *c_col_idx_out = NULL;
    *c_values_out = NULL;
    *c_nnz_out = 0;

    rvsp_csr_matrix_t A = {a_rows, a_cols, a_row_ptr[a_rows], (int32_t *)a_row_ptr,
                           (int32_t *)a_col_idx, (void *)a_values, RVSP_DTYPE_FP32,
                           RVSP_FORMAT_CSR, 0};
    rvsp_csr_matrix_t B = {a_cols, b_cols, b_row_ptr[a_cols], (int32_t *)b_row_ptr,
                           (int32_t *)b_col_idx, (void *)b_values, RVSP_DTYPE_FP32,
                           RVSP_FORMAT_CSR, 0};
    rvsp_spgemm_descr_t d = NULL;
    rvsp_status_t s = rvsp_spgemm_descr_create(&d);
    if (s != RVSP_SUCCESS)
        return s;
    size_t ws_bytes = 0;
    int32_t c_nnz = 0;
    s = rvsp_spgemm_work_estimation(d, &A, &B, &ws_bytes, &c_nnz);
    if (s == RVSP_SUCCESS && op_counts_out != NULL)
        s = rvsp_spgemm_get_op_counts(d, op_counts_out, a_rows);
    if (s != RVSP_SUCCESS)
    {
        rvsp_spgemm_descr_destroy(d);
        return s;
    }
    const size_t cap = c_nnz > 0 ? (size_t)c_nnz : 1;
    rvsp_csr_matrix_t C = {a_rows, b_cols, c_nnz, NULL, NULL, NULL, RVSP_DTYPE_FP32,
                           RVSP_FORMAT_CSR, 1};
    C.row_ptr = (int32_t *)malloc(((size_t)a_rows + 1) * sizeof(int32_t));
    C.col_idx = (int32_t *)malloc(cap * sizeof(int32_t));
    C.values = malloc(cap * sizeof(float));
    if (!C.row_ptr || !C.col_idx || !C.values)
        s = RVSP_ERROR_ALLOCATION_FAILED;
    else
        s = rvsp_spgemm_compute(d, &A, &B, &C, workspace);
    rvsp_spgemm_descr_destroy(d);
    if (s != RVSP_SUCCESS)
    {
        free(C.row_ptr);
        free(C.col_idx);
        free(C.values);
        return s;
    }
    *c_row_ptr_out = C.row_ptr;
    *c_col_idx_out = C.col_idx;
    *c_values_out = (float *)C.values;
    *c_nnz_out = c_nnz;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_csr_device_algo(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                          rvsp_csr_matrix_t *C, rvsp_spgemm_algo_t algo)
{
    if (!A || !B || !C)
        return RVSP_ERROR_NULL_POINTER;
    if (A->dtype != RVSP_DTYPE_FP32 || B->dtype != RVSP_DTYPE_FP32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (A->cols != B->rows)
        return RVSP_ERROR_INVALID_ARGUMENT;

    rvsp_spgemm_descr_t d = NULL;
    rvsp_status_t s = rvsp_spgemm_descr_create(&d);
    if (s != RVSP_SUCCESS)
        return s;
    size_t ws_bytes = 0;
    int32_t c_nnz = 0;
    void *st = NULL;
    void *ost = rvsp_hip_oneshot_stream();
    if ((s = rvsp_spgemm_set_algo(d, algo)) || (s = rvsp_spgemm_set_memory(d, RVSP_MEMORY_DEVICE)) ||
        (ost && (s = rvsp_spgemm_set_stream(d, ost))) ||
        (s = rvsp_spgemm_work_estimation(d, A, B, &ws_bytes, &c_nnz)) ||
        (s = rvsp_spgemm_get_stream(d, &st)))
    {
        rvsp_spgemm_descr_destroy(d);
        return s;
    }
    const size_t cap = c_nnz > 0 ? (size_t)c_nnz : 1;
    rvsp_csr_matrix_t out = {A->rows, B->cols, c_nnz, NULL, NULL, NULL, RVSP_DTYPE_FP32,
                             RVSP_FORMAT_CSR, RVSP_OWNS_DEVICE};
    /* stream-ordered pool allocations: repeated products reuse memory */
    if ((s = rvsp_hip_alloc_async((void **)&out.row_ptr, ((size_t)A->rows + 1) * 4, st)) ||
        (s = rvsp_hip_alloc_async((void **)&out.col_idx, cap * 4, st)) ||
        (s = rvsp_hip_alloc_async(&out.values, cap * 4, st)))
    {
        rvsp_hip_free_async(out.row_ptr, st);
        rvsp_hip_free_async(out.col_idx, st);
        rvsp_hip_free_async(out.values, st);
        rvsp_spgemm_descr_destroy(d);
        return s;
    }
    char ws[8];
    s = rvsp_spgemm_compute(d, A, B, &out, ws);
    if (s == RVSP_SUCCESS)
        s = rvsp_spgemm_wait(d);
    rvsp_spgemm_descr_destroy(d);
    if (s != RVSP_SUCCESS)
    {
        rvsp_hip_free_async(out.row_ptr, st);
        rvsp_hip_free_async(out.col_idx, st);
        rvsp_hip_free_async(out.values, st);
        rvsp_hip_stream_sync(st);
        return s;
    }
    *C = out;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_csr_device(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                     rvsp_csr_matrix_t *C, const rvsp_spgemm_options_t *options)
{
    if (!A || !B || !C)
        return RVSP_ERROR_NULL_POINTER;
    rvsp_backend_t backend = options ? options->backend : RVSP_BACKEND_HIP;
    rvsp_dtype_t in = options ? options->input_dtype : RVSP_DTYPE_FP32;
    rvsp_dtype_t out = options ? options->output_dtype : RVSP_DTYPE_FP32;
    if ((backend == RVSP_BACKEND_SCALAR || backend == RVSP_BACKEND_SCALAR_UNROLL4 ||
         backend == RVSP_BACKEND_HIP) &&
        in == RVSP_DTYPE_FP32 && out == RVSP_DTYPE_FP32)
        return rvsp_spgemm_csr_device_algo(A, B, C, RVSP_SPGEMM_ALGO_DEFAULT);
    return RVSP_ERROR_UNSUPPORTED_BACKEND;
}
