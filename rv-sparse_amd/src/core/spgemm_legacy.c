/*
 * spgemm_legacy.c -- the legacy dtypes of rvsp_spgemm_csr on the GPU: FP64
 * in and out, and INT8 in with INT32 out. rv-sparse runs them on its
 * legacy scalar kernels (src/core/spgemm.c:182-194, wrappers in
 * src/kernels/spgemm/legacy/csr_spgemm_wrappers.c:15-120 and :352-400);
 * here they share the FP32 path's analysis and structure fill, run the
 * typed numeric pass, and drop entries that cancel to exactly zero, as the
 * legacy kernels store only nonzero results. Argument checks and their
 * status codes follow the wrappers. Like the legacy kernels
 * (csr_scalar_f64.c:115-160, 206-230), A's rows may hold columns in any
 * order and repeat them: A is walked in stored order, which fixes each
 * output's summation order. B's rows may be unsorted too: they are sorted on
 * the host first, which changes no sum, since a B row adds at most once to
 * each output column. A B row that repeats a column is the one input the
 * legacy kernels accept and this path refuses (RVSP_ERROR_INVALID_CSR): two
 * adds of one k-step into one column have no order on the GPU.
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
    return RVSP_SUCCESS;
}

/* Range checks of the legacy kernels (row pointers, column range): they
 * report RVSP_ERROR_INVALID_ARGUMENT. Column order and repeats are not
 * errors here; *canonical tells whether the rows are already ascending with
 * no repeats. */
static rvsp_status_t legacy_ranges(const rvsp_csr_matrix_t *M, int *canonical)
{
    int32_t bad = -1;
    rvsp_csr_status_t r = rvsp_csr_check(M->rows, M->cols, M->row_ptr, M->col_idx, &bad);
    if (r == RVSP_CSR_OK)
    {
        *canonical = 1;
        return RVSP_SUCCESS;
    }
    *canonical = 0;
    if (r == RVSP_CSR_BAD_ROW_PTR || r == RVSP_CSR_COL_OUT_OF_RANGE)
        return RVSP_ERROR_INVALID_ARGUMENT;
    /* the first violation was an order one: the rows after it still need
       their ranges checked */
    if (M->row_ptr[0] != 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    for (int32_t row = 0; row < M->rows; row++)
    {
        const int32_t s = M->row_ptr[row], e = M->row_ptr[row + 1];
        if (s < 0 || e < s)
            return RVSP_ERROR_INVALID_ARGUMENT;
        for (int32_t p = s; p < e; p++)
            if (M->col_idx[p] < 0 || M->col_idx[p] >= M->cols)
                return RVSP_ERROR_INVALID_ARGUMENT;
    }
    return RVSP_SUCCESS;
}

static int cmp_i64(const void *a, const void *b)
{
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* B with every row sorted by column (stable), into malloc'd arrays;
 * RVSP_ERROR_INVALID_CSR when a row repeats a column. */
static rvsp_status_t sorted_copy(const rvsp_csr_matrix_t *B, size_t vbytes, int32_t **ci_out,
                                 void **va_out)
{
    const int32_t nnz = B->row_ptr[B->rows];
    const size_t cap = nnz > 0 ? (size_t)nnz : 1;
    int32_t *ci = (int32_t *)malloc(cap * sizeof(int32_t));
    char *va = (char *)malloc(cap * vbytes);
    int32_t longest = 0;
    for (int32_t row = 0; row < B->rows; row++)
        if (B->row_ptr[row + 1] - B->row_ptr[row] > longest)
            longest = B->row_ptr[row + 1] - B->row_ptr[row];
    int64_t *key = (int64_t *)malloc((longest > 0 ? (size_t)longest : 1) * sizeof(int64_t));
    if (!ci || !va || !key)
    {
        free(ci);
        free(va);
        free(key);
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    const char *src = (const char *)B->values;
    for (int32_t row = 0; row < B->rows; row++)
    {
        const int32_t s = B->row_ptr[row], len = B->row_ptr[row + 1] - s;
        /* (column, position in row): a stable order by column */
        for (int32_t j = 0; j < len; j++)
            key[j] = ((int64_t)B->col_idx[s + j] << 32) | (int64_t)j;
        qsort(key, (size_t)len, sizeof(int64_t), cmp_i64);
        for (int32_t j = 0; j < len; j++)
        {
            const int32_t from = s + (int32_t)(key[j] & 0xffffffff);
            ci[s + j] = (int32_t)(key[j] >> 32);
            memcpy(va + (size_t)(s + j) * vbytes, src + (size_t)from * vbytes, vbytes);
            if (j > 0 && ci[s + j] == ci[s + j - 1])
            {
                free(ci);
                free(va);
                free(key);
                return RVSP_ERROR_INVALID_CSR;
            }
        }
    }
    free(key);
    *ci_out = ci;
    *va_out = va;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_legacy_hip(const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                                     rvsp_csr_matrix_t *C, rvsp_dtype_t in)
{
    rvsp_status_t s = legacy_checks(A, B, in);
    int a_canon = 0, b_canon = 0;
    if (s == RVSP_SUCCESS)
        s = legacy_ranges(A, &a_canon);
    if (s == RVSP_SUCCESS)
        s = legacy_ranges(B, &b_canon);
    if (s != RVSP_SUCCESS)
        return s;
    (void)a_canon; /* A is used in stored order either way */
    const rvsp_dtype_t outt = in == RVSP_DTYPE_FP64 ? RVSP_DTYPE_FP64 : RVSP_DTYPE_INT32;
    const size_t vin = value_bytes(in), vout = value_bytes(outt);
    /* B's columns ascending per row (the kernels search and window them) */
    int32_t *b_ci = B->col_idx;
    void *b_va = B->values;
    int32_t *b_ci_sorted = NULL;
    void *b_va_sorted = NULL;
    if (!b_canon)
    {
        if ((s = sorted_copy(B, vin, &b_ci_sorted, &b_va_sorted)) != RVSP_SUCCESS)
            return s;
        b_ci = b_ci_sorted;
        b_va = b_va_sorted;
    }
    const int32_t m = A->rows, k = A->cols, n = B->cols;
    if (m == 0 || k == 0 || n == 0 || A->nnz == 0 || B->nnz == 0)
    {
        free(b_ci_sorted);
        free(b_va_sorted);
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
        (s = rvsp_hip_copy_async(d[BCI], b_ci, (size_t)B->nnz * 4, st)) ||
        (s = rvsp_hip_copy_async(d[BVA], b_va, (size_t)B->nnz * vin, st)) ||
        (s = rvsp_hip_plan_create(&plan, st)))
        goto done;
    /* ranges checked above; A's order and repeats are legal here */
    rvsp_hip_plan_operands_checked(plan);
    if ((s = rvsp_hip_plan_analyse(plan, m, k, n, A->nnz, (const int32_t *)d[ARP],
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
    free(b_ci_sorted);
    free(b_va_sorted);
    free(hrp);
    free(hci);
    free(hva);
    return s;
}
