/*
 * support.c -- canonical-CSR check and workspace size query.
 * Behaviour of rv-sparse src/kernels/spgemm/rvsp_support.c:11-151.
 */
#include "rvsp_spgemm.h"

/* The GPU path keeps its scratch in device memory; the host workspace the
 * API requires is only checked for presence. */
#define RVSP_HOST_WORKSPACE_BYTES ((size_t)256)

rvsp_status_t rvsp_spgemm_buffer_size(int32_t b_cols, size_t *bytes_out)
{
    if (bytes_out == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (b_cols <= 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    *bytes_out = RVSP_HOST_WORKSPACE_BYTES;
    return RVSP_SUCCESS;
}

static rvsp_csr_status_t fail_at(int32_t *bad_row_out, int32_t row, rvsp_csr_status_t why)
{
    if (bad_row_out != NULL)
        *bad_row_out = row;
    return why;
}

rvsp_csr_status_t rvsp_csr_check(int32_t rows, int32_t cols, const int32_t *row_ptr,
                                 const int32_t *col_idx, int32_t *bad_row_out)
{
    if (bad_row_out != NULL)
        *bad_row_out = -1;
    if (rows < 0 || cols < 0 || row_ptr == NULL || row_ptr[0] != 0)
        return RVSP_CSR_BAD_ROW_PTR;

    for (int32_t r = 0; r < rows; r++)
    {
        const int32_t start = row_ptr[r];
        const int32_t end = row_ptr[r + 1];
        if (end < start)
            return fail_at(bad_row_out, r, RVSP_CSR_BAD_ROW_PTR);
        if (end == start)
            continue;
        if (col_idx == NULL)
            return RVSP_CSR_BAD_ROW_PTR;
        int32_t last = -1; /* every valid column exceeds it */
        for (int32_t p = start; p < end; p++)
        {
            const int32_t c = col_idx[p];
            if (c < 0 || c >= cols)
                return fail_at(bad_row_out, r, RVSP_CSR_COL_OUT_OF_RANGE);
            if (c == last)
                return fail_at(bad_row_out, r, RVSP_CSR_DUPLICATE);
            if (c < last)
                return fail_at(bad_row_out, r, RVSP_CSR_UNSORTED);
            last = c;
        }
    }
    return RVSP_CSR_OK;
}

const char *rvsp_csr_status_string(rvsp_csr_status_t status)
{
    switch (status)
    {
    case RVSP_CSR_OK:
        return "canonical CSR";
    case RVSP_CSR_BAD_ROW_PTR:
        return "malformed row pointer";
    case RVSP_CSR_COL_OUT_OF_RANGE:
        return "column index out of range";
    case RVSP_CSR_UNSORTED:
        return "column indices not ascending";
    case RVSP_CSR_DUPLICATE:
        return "duplicate column in row";
    }
    return "unknown";
}
