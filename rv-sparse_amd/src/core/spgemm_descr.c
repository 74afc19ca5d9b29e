/*
 * spgemm_descr.c -- the two-phase SpGEMM descriptor on the HIP backend.
 *
 * State machine and argument contract of rv-sparse
 * src/core/spgemm_descr.c:19-380: create -> [set_algo] -> work_estimation
 * (structure of C, op counts, workspace size) -> compute (x N, reusing the
 * structure; C's column indices are written only when C->col_idx differs
 * from the buffer filled last or after invalidate_structure) -> destroy.
 *
 * Where the reference keeps mark/acc arrays in a host workspace, this
 * descriptor owns a device plan (rvsp_hip_abi.h) and, in host-memory mode,
 * device copies of A, B and C. In device-memory mode (rv_sparse_hip.h) the
 * caller's device arrays are used in place and compute() only queues work.
 */
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"
#include "rv_sparse_hip.h"
#include "rvsp_hip_abi.h"
#include "rvsp_spgemm.h"

struct rvsp_spgemm_descr
{
    rvsp_spgemm_algo_t algo;
    int estimated;
    rvsp_memory_t memory;
    int use_graph;
    void *user_stream;

    int32_t c_nnz;
    int64_t products;
    int32_t *h_row_ptr; /* host copy of C's row pointer (host mode) */
    int64_t *h_ops;     /* host copy of per-row product counts (lazy) */
    const int32_t *filled_into;

    int32_t a_rows, a_cols, a_nnz;
    int32_t b_rows, b_cols, b_nnz;
    /* device mode: the structure arrays the analysis read (and the kernels
       keep reading) */
    const int32_t *e_arp, *e_aci, *e_brp, *e_bci;
    size_t workspace_bytes;

    rvsp_hip_plan_t *plan;

    /* host mode: device images of the operands and of C */
    int32_t *d_arp, *d_aci, *d_brp, *d_bci;
    float *d_aval, *d_bval;
    int32_t *d_cci;
    float *d_cval;
    int b_shares_a; /* B is the same host matrix as A (C = A * A) */
    int host_checked; /* the caller ran rvsp_csr_check on the next operands */
};

/* Tells the next host-mode work_estimation that its operands passed
 * rvsp_csr_check already (the one-shot checks them before the descriptor
 * exists, as rv-sparse's does), so it does not scan them again. */
__attribute__((visibility("hidden"))) void rvsp_descr_operands_checked(rvsp_spgemm_descr_t d)
{
    if (d != NULL)
        d->host_checked = 1;
}

static int algo_supported(rvsp_spgemm_algo_t algo)
{
    return algo == RVSP_SPGEMM_ALGO_DEFAULT || algo == RVSP_SPGEMM_ALGO_HIP ||
           algo == RVSP_SPGEMM_ALGO_HIP_LEARNED;
}

static void release_device_images(struct rvsp_spgemm_descr *d)
{
    void *st = d->plan ? rvsp_hip_plan_stream(d->plan) : NULL;
    rvsp_hip_free_async(d->d_arp, st);
    rvsp_hip_free_async(d->d_aci, st);
    if (!d->b_shares_a)
    {
        rvsp_hip_free_async(d->d_brp, st);
        rvsp_hip_free_async(d->d_bci, st);
    }
    rvsp_hip_free_async(d->d_aval, st);
    rvsp_hip_free_async(d->d_bval, st);
    rvsp_hip_free_async(d->d_cci, st);
    rvsp_hip_free_async(d->d_cval, st);
    d->d_arp = d->d_aci = d->d_brp = d->d_bci = NULL;
    d->d_aval = d->d_bval = NULL;
    d->d_cci = NULL;
    d->d_cval = NULL;
    d->b_shares_a = 0;
}

static void reset_structure(struct rvsp_spgemm_descr *d)
{
    free(d->h_row_ptr);
    free(d->h_ops);
    d->h_row_ptr = NULL;
    d->h_ops = NULL;
    d->filled_into = NULL;
    d->estimated = 0;
    d->c_nnz = 0;
    d->products = 0;
    release_device_images(d);
}

rvsp_status_t rvsp_spgemm_descr_create(rvsp_spgemm_descr_t *descr)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    *descr = NULL;
    struct rvsp_spgemm_descr *d = (struct rvsp_spgemm_descr *)calloc(1, sizeof(*d));
    if (d == NULL)
        return RVSP_ERROR_ALLOCATION_FAILED;
    d->algo = RVSP_SPGEMM_ALGO_DEFAULT;
    d->memory = RVSP_MEMORY_HOST;
    *descr = d;
    return RVSP_SUCCESS;
}

void rvsp_spgemm_descr_destroy(rvsp_spgemm_descr_t descr)
{
    if (descr == NULL)
        return;
    reset_structure(descr);
    rvsp_hip_plan_destroy(descr->plan);
    free(descr);
}

rvsp_status_t rvsp_spgemm_set_algo(rvsp_spgemm_descr_t descr, rvsp_spgemm_algo_t algo)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    /* the strategy shapes the analysis, so it is fixed once estimated */
    if (descr->estimated)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (!algo_supported(algo))
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    descr->algo = algo;
    return RVSP_SUCCESS;
}

void rvsp_spgemm_invalidate_structure(rvsp_spgemm_descr_t descr)
{
    if (descr != NULL)
        descr->filled_into = NULL;
}

static rvsp_status_t check_operand(const rvsp_csr_matrix_t *M)
{
    if (M == NULL || M->row_ptr == NULL || M->col_idx == NULL || M->values == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (M->dtype != RVSP_DTYPE_FP32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (M->rows <= 0 || M->cols <= 0 || M->nnz < 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    return RVSP_SUCCESS;
}

static rvsp_status_t ensure_plan(struct rvsp_spgemm_descr *d)
{
    if (d->plan != NULL)
        return RVSP_SUCCESS;
    return rvsp_hip_plan_create(&d->plan, d->user_stream);
}

/* Device images of A's and B's structure (host mode). */
static rvsp_status_t upload_structure(struct rvsp_spgemm_descr *d, const rvsp_csr_matrix_t *A,
                                      const rvsp_csr_matrix_t *B)
{
    void *st = rvsp_hip_plan_stream(d->plan);
    rvsp_status_t s;
    /* C = A * A through one matrix: upload its structure once (values are
     * still uploaded per operand, since compute may pass other values) */
    d->b_shares_a = (A->row_ptr == B->row_ptr && A->col_idx == B->col_idx &&
                     A->rows == B->rows && A->nnz == B->nnz);
    if ((s = rvsp_hip_alloc_async((void **)&d->d_arp, ((size_t)A->rows + 1) * 4, st)) ||
        (s = rvsp_hip_alloc_async((void **)&d->d_aci, (size_t)A->nnz * 4, st)) ||
        (s = rvsp_hip_alloc_async((void **)&d->d_aval, (size_t)A->nnz * 4, st)) ||
        (s = rvsp_hip_alloc_async((void **)&d->d_bval, (size_t)B->nnz * 4, st)))
        return s;
    if (d->b_shares_a)
    {
        d->d_brp = d->d_arp;
        d->d_bci = d->d_aci;
    }
    else if ((s = rvsp_hip_alloc_async((void **)&d->d_brp, ((size_t)B->rows + 1) * 4, st)) ||
             (s = rvsp_hip_alloc_async((void **)&d->d_bci, (size_t)B->nnz * 4, st)))
        return s;
    if ((s = rvsp_hip_copy_async(d->d_arp, A->row_ptr, ((size_t)A->rows + 1) * 4, st)) ||
        (s = rvsp_hip_copy_async(d->d_aci, A->col_idx, (size_t)A->nnz * 4, st)))
        return s;
    if (!d->b_shares_a &&
        ((s = rvsp_hip_copy_async(d->d_brp, B->row_ptr, ((size_t)B->rows + 1) * 4, st)) ||
         (s = rvsp_hip_copy_async(d->d_bci, B->col_idx, (size_t)B->nnz * 4, st))))
        return s;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_work_estimation(rvsp_spgemm_descr_t descr, const rvsp_csr_matrix_t *A,
                                          const rvsp_csr_matrix_t *B, size_t *workspace_bytes,
                                          int32_t *c_nnz_out)
{
    if (descr == NULL || workspace_bytes == NULL || c_nnz_out == NULL)
        return RVSP_ERROR_NULL_POINTER;
    const int checked = descr->host_checked;
    descr->host_checked = 0;
    rvsp_status_t s = check_operand(A);
    if (s != RVSP_SUCCESS)
        return s;
    s = check_operand(B);
    if (s != RVSP_SUCCESS)
        return s;
    if (A->cols != B->rows)
        return RVSP_ERROR_INVALID_ARGUMENT;

    /* the kernels index through the operands' structure: host operands
       are checked here, before any upload, device operands by the analysis
       on the device (rv-sparse leaves non-canonical input to the
       descriptor path undefined; this build rejects it) */
    if (descr->memory == RVSP_MEMORY_HOST && !checked &&
        (rvsp_csr_check(A->rows, A->cols, A->row_ptr, A->col_idx, NULL) != RVSP_CSR_OK ||
         rvsp_csr_check(B->rows, B->cols, B->row_ptr, B->col_idx, NULL) != RVSP_CSR_OK))
        return RVSP_ERROR_INVALID_CSR;

    /* a new estimation replaces whatever the descriptor held */
    reset_structure(descr);
    s = ensure_plan(descr);
    if (s != RVSP_SUCCESS)
        return s;

    const int32_t *arp = A->row_ptr, *aci = A->col_idx, *brp = B->row_ptr, *bci = B->col_idx;
    if (descr->memory == RVSP_MEMORY_HOST)
    {
        rvsp_hip_plan_operands_checked(descr->plan);
        s = upload_structure(descr, A, B);
        if (s != RVSP_SUCCESS)
        {
            reset_structure(descr);
            return s;
        }
        arp = descr->d_arp;
        aci = descr->d_aci;
        brp = descr->d_brp;
        bci = descr->d_bci;
    }

    int64_t c_nnz = 0, products = 0;
    /* RVSP_SPGEMM_ALGO_HIP forces the heuristic bins; the default and the
       learned algo take the learned selectors when their weights are loaded
       (the heuristic otherwise). The result is the same bit for bit. */
    const rvsp_select_t sel = descr->algo == RVSP_SPGEMM_ALGO_HIP ? RVSP_SELECT_HEURISTIC
                                                                  : RVSP_SELECT_LEARNED;
    s = rvsp_hip_plan_analyse(descr->plan, A->rows, A->cols, B->cols, A->nnz, arp, aci, B->nnz,
                              brp, bci, sel, &c_nnz, &products);
    if (s != RVSP_SUCCESS)
    {
        reset_structure(descr);
        return s;
    }

    if (descr->memory == RVSP_MEMORY_HOST)
    {
        void *st = rvsp_hip_plan_stream(descr->plan);
        const size_t cap = c_nnz > 0 ? (size_t)c_nnz : 1;
        descr->h_row_ptr = (int32_t *)malloc(((size_t)A->rows + 1) * sizeof(int32_t));
        if (descr->h_row_ptr == NULL)
        {
            reset_structure(descr);
            return RVSP_ERROR_ALLOCATION_FAILED;
        }
        if ((s = rvsp_hip_copy_async(descr->h_row_ptr, rvsp_hip_plan_row_ptr(descr->plan),
                                     ((size_t)A->rows + 1) * 4, st)) ||
            (s = rvsp_hip_alloc_async((void **)&descr->d_cci, cap * 4, st)) ||
            (s = rvsp_hip_alloc_async((void **)&descr->d_cval, cap * 4, st)) ||
            (s = rvsp_hip_stream_sync(st)))
        {
            reset_structure(descr);
            return s;
        }
    }

    descr->c_nnz = (int32_t)c_nnz;
    descr->products = products;
    descr->a_rows = A->rows;
    descr->a_cols = A->cols;
    descr->a_nnz = A->nnz;
    descr->b_rows = B->rows;
    descr->b_cols = B->cols;
    descr->b_nnz = B->nnz;
    descr->e_arp = A->row_ptr;
    descr->e_aci = A->col_idx;
    descr->e_brp = B->row_ptr;
    descr->e_bci = B->col_idx;
    rvsp_spgemm_buffer_size(B->cols, &descr->workspace_bytes);
    descr->estimated = 1;

    *workspace_bytes = descr->workspace_bytes;
    *c_nnz_out = descr->c_nnz;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_compute(rvsp_spgemm_descr_t descr, const rvsp_csr_matrix_t *A,
                                  const rvsp_csr_matrix_t *B, rvsp_csr_matrix_t *C,
                                  void *workspace)
{
    /* argument precedence as rv-sparse spgemm_descr.c:263-313 */
    if (descr == NULL || C == NULL || workspace == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (!descr->estimated)
        return RVSP_ERROR_INVALID_ARGUMENT;
    rvsp_status_t s = check_operand(A);
    if (s != RVSP_SUCCESS)
        return s;
    s = check_operand(B);
    if (s != RVSP_SUCCESS)
        return s;
    if (A->rows != descr->a_rows || A->cols != descr->a_cols || A->nnz != descr->a_nnz ||
        B->rows != descr->b_rows || B->cols != descr->b_cols || B->nnz != descr->b_nnz)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (C->row_ptr == NULL || C->col_idx == NULL || C->values == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (C->dtype != RVSP_DTYPE_FP32)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (C->rows != descr->a_rows || C->cols != descr->b_cols)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (C->nnz < descr->c_nnz)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (!algo_supported(descr->algo))
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    /* device mode reads A's and B's structure in place through the plan,
       which is sized by the arrays it analysed: other arrays, even equal
       ones, need a new estimation */
    if (descr->memory == RVSP_MEMORY_DEVICE &&
        (A->row_ptr != descr->e_arp || A->col_idx != descr->e_aci ||
         B->row_ptr != descr->e_brp || B->col_idx != descr->e_bci))
        return RVSP_ERROR_INVALID_ARGUMENT;

    rvsp_hip_plan_t *plan = descr->plan;
    void *st = rvsp_hip_plan_stream(plan);
    const size_t rp_bytes = ((size_t)descr->a_rows + 1) * sizeof(int32_t);
    const size_t c_bytes = (size_t)descr->c_nnz * sizeof(int32_t);
    const int refill = descr->filled_into != C->col_idx;

    if (descr->memory == RVSP_MEMORY_DEVICE)
    {
        if ((s = rvsp_hip_copy_async(C->row_ptr, rvsp_hip_plan_row_ptr(plan), rp_bytes, st)))
            return s;
        if (refill)
        {
            /* structure and values in one go: big rows get their columns
               from the numeric pass */
            s = rvsp_hip_plan_fill_numeric(plan, (const float *)A->values,
                                           (const float *)B->values, C->col_idx,
                                           (float *)C->values);
            if (s != RVSP_SUCCESS)
                return s;
            descr->filled_into = C->col_idx;
        }
        else
        {
            if ((s = rvsp_hip_plan_prepare_reuse(plan, C->col_idx)))
                return s;
            if (descr->use_graph)
                s = rvsp_hip_plan_numeric_graph(plan, (const float *)A->values,
                                                (const float *)B->values, C->col_idx,
                                                (float *)C->values);
            else
                s = rvsp_hip_plan_numeric(plan, (const float *)A->values,
                                          (const float *)B->values, C->col_idx,
                                          (float *)C->values);
            if (s != RVSP_SUCCESS)
                return s;
        }
    }
    else
    {
        /* values go up every call: only the pattern may be reused */
        if ((s = rvsp_hip_copy_async(descr->d_aval, A->values, (size_t)A->nnz * 4, st)) ||
            (s = rvsp_hip_copy_async(descr->d_bval, B->values, (size_t)B->nnz * 4, st)))
            return s;
        memcpy(C->row_ptr, descr->h_row_ptr, rp_bytes);
        if (refill)
        {
            /* columns go down after the pass that writes the big rows' */
            if ((s = rvsp_hip_plan_fill_numeric(plan, descr->d_aval, descr->d_bval,
                                                descr->d_cci, descr->d_cval)) ||
                (s = rvsp_hip_copy_async(C->col_idx, descr->d_cci, c_bytes, st)))
                return s;
        }
        else
        {
            if ((s = rvsp_hip_plan_prepare_reuse(plan, descr->d_cci)))
                return s;
            if (descr->use_graph)
                s = rvsp_hip_plan_numeric_graph(plan, descr->d_aval, descr->d_bval,
                                                descr->d_cci, descr->d_cval);
            else
                s = rvsp_hip_plan_numeric(plan, descr->d_aval, descr->d_bval, descr->d_cci,
                                          descr->d_cval);
            if (s != RVSP_SUCCESS)
                return s;
        }
        if ((s = rvsp_hip_copy_async(C->values, descr->d_cval, c_bytes, st)))
            return s;
        if ((s = rvsp_hip_plan_check(plan)))
            return s;
        if (refill)
            descr->filled_into = C->col_idx;
    }

    C->nnz = descr->c_nnz;
    C->format = RVSP_FORMAT_CSR;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_get_op_counts(rvsp_spgemm_descr_t descr, int64_t *op_counts_out,
                                        int32_t n)
{
    if (descr == NULL || op_counts_out == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (!descr->estimated || descr->plan == NULL)
        return RVSP_ERROR_INVALID_ARGUMENT;
    if (n < descr->a_rows)
        return RVSP_ERROR_INVALID_ARGUMENT;
    const size_t bytes = (size_t)descr->a_rows * sizeof(int64_t);
    if (descr->h_ops == NULL)
    {
        int64_t *h = (int64_t *)malloc(bytes);
        if (h == NULL)
            return RVSP_ERROR_ALLOCATION_FAILED;
        void *st = rvsp_hip_plan_stream(descr->plan);
        rvsp_status_t s = rvsp_hip_copy_async(h, rvsp_hip_plan_op_counts(descr->plan), bytes, st);
        if (s == RVSP_SUCCESS)
            s = rvsp_hip_stream_sync(st);
        if (s != RVSP_SUCCESS)
        {
            free(h);
            return s;
        }
        descr->h_ops = h;
    }
    memcpy(op_counts_out, descr->h_ops, bytes);
    return RVSP_SUCCESS;
}

/* ---- rv_sparse_hip.h descriptor extensions ------------------------------ */

rvsp_status_t rvsp_spgemm_set_memory(rvsp_spgemm_descr_t descr, rvsp_memory_t memory)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (descr->estimated || (memory != RVSP_MEMORY_HOST && memory != RVSP_MEMORY_DEVICE))
        return RVSP_ERROR_INVALID_ARGUMENT;
    descr->memory = memory;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_stream(rvsp_spgemm_descr_t descr, void *stream)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (descr->estimated || descr->plan != NULL)
        return RVSP_ERROR_INVALID_ARGUMENT;
    descr->user_stream = stream;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_get_stream(rvsp_spgemm_descr_t descr, void **stream)
{
    if (descr == NULL || stream == NULL)
        return RVSP_ERROR_NULL_POINTER;
    rvsp_status_t s = ensure_plan(descr);
    if (s != RVSP_SUCCESS)
        return s;
    *stream = rvsp_hip_plan_stream(descr->plan);
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_set_graph(rvsp_spgemm_descr_t descr, int enable)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    descr->use_graph = enable != 0;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_get_products(rvsp_spgemm_descr_t descr, int64_t *products)
{
    if (descr == NULL || products == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (!descr->estimated)
        return RVSP_ERROR_INVALID_ARGUMENT;
    *products = descr->products;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_spgemm_wait(rvsp_spgemm_descr_t descr)
{
    if (descr == NULL)
        return RVSP_ERROR_NULL_POINTER;
    if (descr->plan == NULL)
        return RVSP_SUCCESS;
    return rvsp_hip_plan_check(descr->plan);
}
