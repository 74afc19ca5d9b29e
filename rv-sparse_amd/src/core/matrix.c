/*
 * matrix.c -- CSR matrix lifecycle (borrowing create, basic validation,
 * owner-aware destroy). Behaviour of rv-sparse src/core/matrix.c:18-115,
 * extended so a library-allocated DEVICE C (owns_data == RVSP_OWNS_DEVICE)
 * is released on the GPU.
 */
#include <stdlib.h>

#include "rv_sparse.h"
#include "rvsp_hip_abi.h"

rvsp_status_t rvsp_csr_create(rvsp_csr_matrix_t *A, int32_t rows, int32_t cols, int32_t nnz,
                              int32_t *row_ptr, int32_t *col_idx, void *values,
                              rvsp_dtype_t dtype)
{
    if (A == NULL || row_ptr == NULL || col_idx == NULL || values == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (rows <= 0 || cols <= 0 || nnz < 0)
        return RVSP_ERROR_INVALID_ARGUMENT;

    A->rows = rows;
    A->cols = cols;
    A->nnz = nnz;
    A->row_ptr = row_ptr;
    A->col_idx = col_idx;
    A->values = values;
    A->dtype = dtype;
    A->format = RVSP_FORMAT_CSR;
    A->owns_data = 0; /* borrowed: destroy never frees these */

    return rvsp_csr_validate(A);
}

rvsp_status_t rvsp_csr_validate(const rvsp_csr_matrix_t *A)
{
    if (A == NULL || A->row_ptr == NULL || A->col_idx == NULL || A->values == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (A->rows <= 0 || A->cols <= 0 || A->nnz < 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (A->row_ptr[0] != 0 || A->row_ptr[A->rows] != A->nnz)
        return RVSP_ERROR_INVALID_CSR;
    for (int32_t r = 0; r < A->rows; r++)
        if (A->row_ptr[r + 1] < A->row_ptr[r])
            return RVSP_ERROR_INVALID_CSR;
    for (int32_t p = 0; p < A->nnz; p++)
        if (A->col_idx[p] < 0 || A->col_idx[p] >= A->cols)
            return RVSP_ERROR_INVALID_CSR;
    return RVSP_SUCCESS;
}

void rvsp_csr_destroy(rvsp_csr_matrix_t *A)
{
    if (A == NULL)
        return;

    if (A->owns_data == 1)
    {
        free(A->row_ptr);
        free(A->col_idx);
        free(A->values);
    }
    else if (A->owns_data == RVSP_OWNS_DEVICE)
    {
        /* stream-ordered pool allocations of the device one-shot, returned on
           the stream they came from so the next call reuses them */
        void *st = rvsp_hip_oneshot_stream();
        rvsp_hip_free_async(A->row_ptr, st);
        rvsp_hip_free_async(A->col_idx, st);
        rvsp_hip_free_async(A->values, st);
        rvsp_hip_stream_sync(st);
    }

    A->rows = 0;
    A->cols = 0;
    A->nnz = 0;
    A->row_ptr = NULL;
    A->col_idx = NULL;
    A->values = NULL;
    A->dtype = RVSP_DTYPE_FP32;
    A->format = RVSP_FORMAT_CSR;
    A->owns_data = 0;
}
