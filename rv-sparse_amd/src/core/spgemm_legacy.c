/*
 * spgemm_legacy.c -- the legacy dtypes of rvsp_spgemm_csr on the GPU: FP64
 * in and out, and INT8 in with INT32 out. rv-sparse runs them on its
 * legacy scalar kernels (src/core/spgemm.c:182-194, wrappers in
 * src/kernels/spgemm/legacy/csr_spgemm_wrappers.c:15-120 and :352-400);
 * here they share the FP32 path's analysis and structure fill, run the
 * typed numeric pass, and drop entries that cancel to exactly zero, as the
 * legacy kernels store only nonzero results. Argument checks and their
 * status codes follow the wrappers; unlike the legacy kernels, rows must
 * be canonical (ascending, no repeated column): other input reports
 * RVSP_ERROR_INVALID_CSR.
 */
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"
#include "rvsp_hip_abi.h"
#include "rvsp_spgemm.h"

static size_t value_bytes(rvsp_dtype_t t)
{
    return t == RVSP_DTYPE_FP64 ? 8 : t == RVSP_DTYPE_INT8 ? 1 : 4;
}

/* the wrappers' checks: FP64 (csr_spgemm_wrappers.c:357-371), INT8 (:24-49) */
static rvsp_status_t legacy_checks(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                   rvsp_dtype_t in)
{
    rvsp_status_t s;
    if (in == RVSP_DTYPE_INT8)
    {
        if ((s = rvsp_csr_validate(A)) != RVSP_SUCCESS || (s = rvsp_csr_validate(B)) != RVSP_SUCCESS)
            return s;
        if (A->dtype != RVSP_DTYPE_INT8 || B->dtype != RVSP_DTYPE_INT8)
            return RVSP_ERROR_UNSUPPORTED_DTYPE;
    }
    else if (A->dtype != RVSP_DTYPE_FP64 || B->dtype != RVSP_DTYPE_FP64)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (A->cols != B->rows)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (!A->row_ptr || !A->col_idx || !A->values || !B->row_ptr || !B->col_idx || !B->values ||
        A->rows < 0 || A->cols < 0 || B->cols < 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    const rvsp_csr_matrix_t *M[2] = {A, B};
    for (int i = 0; i < 2; i++)
    {
        int32_t bad = -1;
        int r = rvsp_csr_check(M[i]->rows, M[i]->cols, M[i]->row_ptr, M[i]->col_idx, &bad);
        if (r == RVSP_CSR_BAD_ROW_PTR || r == RVSP_CSR_COL_OUT_OF_RANGE)
            return RVSP_ERROR_INVALID_ARGUMENT; /* the legacy kernels' own range checks */
        if (r != RVSP_CSR_OK)
            return RVSP_ERROR_INVALID_CSR;
    }
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_legacy_hip(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                     rvsp_csr_matrix_t *C, rvsp_dtype_t in)
{
    rvsp_status_t s = legacy_checks(A, B, in);
    if (s != RVSP_SUCCESS)
        return s;
    const rvsp_dtype_t outt = in == RVSP_DTYPE_FP64 ? RVSP_DTYPE_FP64 : RVSP_DTYPE_INT32;
    const size_t vin = value_bytes(in), vout = value_bytes(outt);
    const int32_t m = A->rows, k = A->cols, n = B->cols;
    if (m == 0 || k == 0 || n == 0 || A->nnz == 0 || B->nnz == 0)
    {
        /* nothing to multiply: an empty C, as the legacy kernels return */
        C->row_ptr = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
        C->col_idx = (int32_t *)malloc(sizeof(int32_t));
        C->values = malloc(vout);
        if (!C->row_ptr || !C->col_idx || !C->values)
        {
            free(C->row_ptr);
            free(C->col_idx);
            free(C->values);
            C->row_ptr = NULL;
            C->col_idx = NULL;
            C->values = NULL;
            return RVSP_ERROR_ALLOCATION_FAILED;
        }
        C->rows = m;
        C->cols = n;
        C->nnz = 0;
        C->dtype = outt;
        C->format = RVSP_FORMAT_CSR;
        C->owns_data = 1;
        return RVSP_SUCCESS;
    }

    void *st = rvsp_hip_oneshot_stream();
    rvsp_hip_plan_t *plan = NULL;
    void *d[12] = {0};
    enum { ARP, ACI, AVA, BRP, BCI, BVA, CCI, CVA, RP2, CI2, CV2, NBUF };
    int32_t *hrp = NULL, *hci = NULL;
    void *hva = NULL;
    int64_t c_nnz = 0, products = 0;
    int32_t kept = 0;

    if ((s = rvsp_hip_alloc_async(&d[ARP], ((size_t)m + 1) * 4, st)) ||
        (s = rvsp_hip_alloc_async(&d[ACI], (size_t)A->nnz * 4, st)) ||
        (s = rvsp_hip_alloc_async(&d[AVA], (size_t)A->nnz * vin, st)) ||
        (s = rvsp_hip_alloc_async(&d[BRP], ((size_t)k + 1) * 4, st)) ||
        (s = rvsp_hip_alloc_async(&d[BCI], (size_t)B->nnz * 4, st)) ||
        (s = rvsp_hip_alloc_async(&d[BVA], (size_t)B->nnz * vin, st)) ||
        (s = rvsp_hip_copy_async(d[ARP], A->row_ptr, ((size_t)m + 1) * 4, st)) ||
        (s = rvsp_hip_copy_async(d[ACI], A->col_idx, (size_t)A->nnz * 4, st)) ||
        (s = rvsp_hip_copy_async(d[AVA], A->values, (size_t)A->nnz * vin, st)) ||
        (s = rvsp_hip_copy_async(d[BRP], B->row_ptr, ((size_t)k + 1) * 4, st)) ||
        (s = rvsp_hip_copy_async(d[BCI], B->col_idx, (size_t)B->nnz * 4, st)) ||
        (s = rvsp_hip_copy_async(d[BVA], B->values, (size_t)B->nnz * vin, st)) ||
        (s = rvsp_hip_plan_create(&plan, st)) ||
        (s = rvsp_hip_plan_analyse(plan, m, k, n, A->nnz, (const int32_t *)d[ARP],
                                   (const int32_t *)d[ACI], B->nnz, (const int32_t *)d[BRP],
                                   (const int32_t *)d[BCI], RVSP_SELECT_HEURISTIC, &c_nnz,
                                   &products)))
        goto done;
    {
        const size_t cap = c_nnz > 0 ? (size_t)c_nnz : 1;
        if ((s = rvsp_hip_alloc_async(&d[CCI], cap * 4, st)) ||
            (s = rvsp_hip_alloc_async(&d[CVA], cap * vout, st)) ||
            (s = rvsp_hip_alloc_async(&d[RP2], ((size_t)m + 1) * 4, st)) ||
            (s = rvsp_hip_alloc_async(&d[CI2], cap * 4, st)) ||
            (s = rvsp_hip_alloc_async(&d[CV2], cap * vout, st)) ||
            (s = rvsp_hip_plan_fill(plan, (int32_t *)d[CCI])) ||
            (s = rvsp_hip_plan_numeric_typed(plan, in, d[AVA], d[BVA], (const int32_t *)d[CCI],
                                             d[CVA])) ||
            (s = rvsp_hip_plan_check(plan)) ||
            (s = rvsp_hip_plan_drop_zeros(plan, outt, (const int32_t *)d[CCI], d[CVA],
                                          (int32_t *)d[RP2], (int32_t *)d[CI2], d[CV2], &kept)))
            goto done;
        hrp = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
        hci = (int32_t *)malloc((kept > 0 ? (size_t)kept : 1) * sizeof(int32_t));
        hva = malloc((kept > 0 ? (size_t)kept : 1) * vout);
        if (!hrp || !hci || !hva)
        {
            s = RVSP_ERROR_ALLOCATION_FAILED;
            goto done;
        }
        if ((s = rvsp_hip_copy_async(hrp, d[RP2], ((size_t)m + 1) * 4, st)) ||
            (kept > 0 && ((s = rvsp_hip_copy_async(hci, d[CI2], (size_t)kept * 4, st)) ||
                          (s = rvsp_hip_copy_async(hva, d[CV2], (size_t)kept * vout, st)))) ||
            (s = rvsp_hip_stream_sync(st)))
            goto done;
    }
    C->rows = m;
    C->cols = n;
    C->nnz = kept;
    C->row_ptr = hrp;
    C->col_idx = hci;
    C->values = hva;
    C->dtype = outt;
    C->format = RVSP_FORMAT_CSR;
    C->owns_data = 1;
    hrp = NULL;
    hci = NULL;
    hva = NULL;

done:
    if (plan)
        rvsp_hip_plan_destroy(plan);
    for (int i = 0; i < NBUF; i++)
        rvsp_hip_free_async(d[i], st);
    rvsp_hip_stream_sync(st);
    free(hrp);
    free(hci);
    free(hva);
    return s;
}
