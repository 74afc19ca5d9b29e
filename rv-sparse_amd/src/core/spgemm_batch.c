/*
 * spgemm_batch.c -- many independent products in few launches.
 *
 * No rv-sparse counterpart (its API computes one product per call). The
 * batch is cut into chunks whose estimated product count stays well inside
 * int32 indexing; each chunk is stacked into one product (rvsp_hip_pack),
 * run through the ordinary descriptor in device memory, and split back
 * (rvsp_hip_unpack_row_ptr). A chunk whose C would still overflow int32 is
 * halved and retried.
 */
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"
#include "rv_sparse_hip.h"
#include "rvsp_hip_abi.h"

/* Estimated products per chunk; RVSP_BATCH_CHUNK_PRODUCTS in the
 * environment overrides it (tests use it to force several chunks). */
static double chunk_products(void)
{
    const char *e = getenv("RVSP_BATCH_CHUNK_PRODUCTS");
    double v = e ? atof(e) : 0.0;
    return v > 0.0 ? v : 1.2e9;
}

/* every allocation, copy and launch of a batch runs on the library's one
   stream, so the pool reuses what the previous call freed */
#define BST rvsp_hip_oneshot_stream()

#define TRY(expr)                                                                                  \
    do                                                                                             \
    {                                                                                              \
        if (s == RVSP_SUCCESS)                                                                     \
            s = (expr);                                                                            \
    } while (0)

struct rvsp_batch_chunk
{
    int32_t *rp;  /* local row pointers of the chunk's problems */
    int32_t *ci;  /* stacked C column indices */
    float *va;    /* stacked C values */
};

struct rvsp_spgemm_batch
{
    int32_t nchunks, cap;
    struct rvsp_batch_chunk *chunks;
};

static rvsp_status_t batch_push(struct rvsp_spgemm_batch *b, struct rvsp_batch_chunk c)
{
    if (b->nchunks == b->cap)
    {
        int32_t nc = b->cap ? 2 * b->cap : 8;
        struct rvsp_batch_chunk *p = realloc(b->chunks, (size_t)nc * sizeof(*p));
        if (!p)
            return RVSP_ERROR_ALLOCATION_FAILED;
        b->chunks = p;
        b->cap = nc;
    }
    b->chunks[b->nchunks++] = c;
    return RVSP_SUCCESS;
}

static void free_chunk(struct rvsp_batch_chunk *c)
{
    rvsp_hip_free_async(c->rp, BST);
    rvsp_hip_free_async(c->ci, BST);
    rvsp_hip_free_async(c->va, BST);
    memset(c, 0, sizeof(*c));
}

void rvsp_spgemm_batch_destroy(rvsp_spgemm_batch_t b)
{
    if (!b)
        return;
    for (int32_t i = 0; i < b->nchunks; i++)
        free_chunk(&b->chunks[i]);
    rvsp_hip_stream_sync(BST);
    free(b->chunks);
    free(b);
}

int32_t rvsp_spgemm_batch_chunks(rvsp_spgemm_batch_t b)
{
    return b ? b->nchunks : 0;
}

/* Stacks problems [p0, p1) and computes them as one product. */
static rvsp_status_t run_chunk(struct rvsp_spgemm_batch *batch, int32_t p0, int32_t p1,
                               const rvsp_csr_matrix_t *A, const rvsp_csr_matrix_t *B,
                               rvsp_csr_matrix_t *C, rvsp_spgemm_algo_t algo)
{
    const int32_t n = p1 - p0;
    int64_t rows = 0, anz = 0, brows = 0, bnz = 0;
    int32_t bcols = 1;
    for (int32_t i = p0; i < p1; i++)
    {
        rows += A[i].rows;
        anz += A[i].nnz;
        brows += B[i].rows;
        bnz += B[i].nnz;
        if (B[i].cols > bcols)
            bcols = B[i].cols;
    }
    if (rows > INT32_MAX - 1 || anz > INT32_MAX || brows > INT32_MAX - 1 || bnz > INT32_MAX)
        return RVSP_ERROR_ALLOCATION_FAILED;

    rvsp_status_t s = RVSP_SUCCESS;
    rvsp_hip_pack_item_t *items = calloc((size_t)n, sizeof(*items));
    int64_t *row_off = calloc((size_t)n, sizeof(int64_t));
    int32_t *nrows = calloc((size_t)n, sizeof(int32_t));
    int64_t *bases = calloc((size_t)n, sizeof(int64_t));
    void *arp = NULL, *aci = NULL, *ava = NULL, *brp = NULL, *bci = NULL, *bva = NULL;
    void *crp = NULL;
    struct rvsp_batch_chunk ch = {0};
    rvsp_spgemm_descr_t d = NULL;
    if (!items || !row_off || !nrows || !bases)
        s = RVSP_ERROR_ALLOCATION_FAILED;

    TRY(rvsp_hip_alloc_async(&arp, ((size_t)rows + 1) * 4, BST));
    TRY(rvsp_hip_alloc_async(&aci, (size_t)anz * 4, BST));
    TRY(rvsp_hip_alloc_async(&ava, (size_t)anz * 4, BST));
    TRY(rvsp_hip_alloc_async(&brp, ((size_t)brows + 1) * 4, BST));
    TRY(rvsp_hip_alloc_async(&bci, (size_t)bnz * 4, BST));
    TRY(rvsp_hip_alloc_async(&bva, (size_t)bnz * 4, BST));
    if (s == RVSP_SUCCESS)
    {
        /* A_i columns shift by the B rows stacked before problem i */
        int64_t ro = 0, no = 0, ko = 0;
        for (int32_t i = 0; i < n; i++)
        {
            const rvsp_csr_matrix_t *M = &A[p0 + i];
            items[i] = (rvsp_hip_pack_item_t){M->row_ptr, M->col_idx, (const float *)M->values,
                                              M->rows, M->nnz, ro, no, (int32_t)ko, M->cols};
            row_off[i] = ro;
            nrows[i] = M->rows;
            ro += M->rows;
            no += M->nnz;
            ko += B[p0 + i].rows;
        }
        s = rvsp_hip_pack(BST, n, items, arp, aci, ava);
    }
    if (s == RVSP_SUCCESS)
    {
        int64_t ro = 0, no = 0;
        for (int32_t i = 0; i < n; i++)
        {
            const rvsp_csr_matrix_t *M = &B[p0 + i];
            items[i] = (rvsp_hip_pack_item_t){M->row_ptr, M->col_idx, (const float *)M->values,
                                              M->rows, M->nnz, ro, no, 0, M->cols};
            ro += M->rows;
            no += M->nnz;
        }
        s = rvsp_hip_pack(BST, n, items, brp, bci, bva);
    }

    rvsp_csr_matrix_t bigA = {(int32_t)rows, (int32_t)brows, (int32_t)anz, arp, aci, ava,
                              RVSP_DTYPE_FP32, RVSP_FORMAT_CSR, 0};
    rvsp_csr_matrix_t bigB = {(int32_t)brows, bcols, (int32_t)bnz, brp, bci, bva,
                              RVSP_DTYPE_FP32, RVSP_FORMAT_CSR, 0};
    size_t ws = 0;
    int32_t cnnz = 0;
    TRY(rvsp_spgemm_descr_create(&d));
    TRY(rvsp_spgemm_set_algo(d, algo));
    TRY(rvsp_spgemm_set_memory(d, RVSP_MEMORY_DEVICE));
    if (rvsp_hip_oneshot_stream())
        TRY(rvsp_spgemm_set_stream(d, rvsp_hip_oneshot_stream()));
    TRY(rvsp_spgemm_work_estimation(d, &bigA, &bigB, &ws, &cnnz));
    if (s == RVSP_ERROR_ALLOCATION_FAILED && n > 1)
    {
        /* C of this chunk overflows int32 (or memory ran out): halve it */
        rvsp_spgemm_descr_destroy(d);
        d = NULL;
        rvsp_hip_free_async(arp, BST);
        rvsp_hip_free_async(aci, BST);
        rvsp_hip_free_async(ava, BST);
        rvsp_hip_free_async(brp, BST);
        rvsp_hip_free_async(bci, BST);
        rvsp_hip_free_async(bva, BST);
        free(items);
        free(row_off);
        free(nrows);
        free(bases);
        const int32_t mid = p0 + n / 2;
        s = run_chunk(batch, p0, mid, A, B, C, algo);
        if (s == RVSP_SUCCESS)
            s = run_chunk(batch, mid, p1, A, B, C, algo);
        return s;
    }
    const size_t cap = cnnz > 0 ? (size_t)cnnz : 1;
    TRY(rvsp_hip_alloc_async(&crp, ((size_t)rows + 1) * 4, BST));
    TRY(rvsp_hip_alloc_async((void **)&ch.ci, cap * 4, BST));
    TRY(rvsp_hip_alloc_async((void **)&ch.va, cap * 4, BST));
    TRY(rvsp_hip_alloc_async((void **)&ch.rp, ((size_t)rows + n) * 4, BST));
    if (s == RVSP_SUCCESS)
    {
        rvsp_csr_matrix_t bigC = {(int32_t)rows, bcols, cnnz, crp, ch.ci, ch.va,
                                  RVSP_DTYPE_FP32, RVSP_FORMAT_CSR, 0};
        char wsb[8];
        rvsp_hip_stream_sync(BST); /* allocations made on the library stream */
        s = rvsp_spgemm_compute(d, &bigA, &bigB, &bigC, wsb);
        if (s == RVSP_SUCCESS)
            s = rvsp_spgemm_wait(d);
    }
    if (s == RVSP_SUCCESS)
        s = rvsp_hip_unpack_row_ptr(BST, n, row_off, nrows, crp, ch.rp, bases);
    if (s == RVSP_SUCCESS)
    {
        int64_t out = 0;
        for (int32_t i = 0; i < n; i++)
        {
            rvsp_csr_matrix_t *Ci = &C[p0 + i];
            const int64_t end = i + 1 < n ? bases[i + 1] : cnnz;
            Ci->rows = A[p0 + i].rows;
            Ci->cols = B[p0 + i].cols;
            Ci->nnz = (int32_t)(end - bases[i]);
            Ci->row_ptr = ch.rp + out;
            Ci->col_idx = ch.ci + bases[i];
            Ci->values = ch.va + bases[i];
            Ci->dtype = RVSP_DTYPE_FP32;
            Ci->format = RVSP_FORMAT_CSR;
            Ci->owns_data = 0;
            out += A[p0 + i].rows + 1;
        }
        s = batch_push(batch, ch);
    }
    if (s != RVSP_SUCCESS)
        free_chunk(&ch);
    rvsp_spgemm_descr_destroy(d);
    rvsp_hip_free_async(arp, BST);
    rvsp_hip_free_async(aci, BST);
    rvsp_hip_free_async(ava, BST);
    rvsp_hip_free_async(brp, BST);
    rvsp_hip_free_async(bci, BST);
    rvsp_hip_free_async(bva, BST);
    rvsp_hip_free_async(crp, BST);
    rvsp_hip_stream_sync(BST);
    free(items);
    free(row_off);
    free(nrows);
    free(bases);
    return s;
}

rvsp_status_t rvsp_spgemm_batched_device(int32_t count, const rvsp_csr_matrix_t *A,
                                         const rvsp_csr_matrix_t *B, rvsp_csr_matrix_t *C,
                                         rvsp_spgemm_algo_t algo, rvsp_spgemm_batch_t *batch)
{
    if (!A || !B || !C || !batch)
        return RVSP_ERROR_NULL_POINTER;
    *batch = NULL;
    if (count <= 0)
        return RVSP_ERROR_INVALID_ARGUMENT;
    for (int32_t i = 0; i < count; i++)
    {
        if (!A[i].row_ptr || !A[i].col_idx || !A[i].values || !B[i].row_ptr ||
            !B[i].col_idx || !B[i].values)
            return RVSP_ERROR_NULL_POINTER;
        if (A[i].dtype != RVSP_DTYPE_FP32 || B[i].dtype != RVSP_DTYPE_FP32)
            return RVSP_ERROR_UNSUPPORTED_DTYPE;
        if (A[i].rows <= 0 || A[i].cols <= 0 || B[i].cols <= 0 || A[i].cols != B[i].rows ||
            A[i].nnz < 0 || B[i].nnz < 0)
            return RVSP_ERROR_INVALID_ARGUMENT;
    }
    struct rvsp_spgemm_batch *b = calloc(1, sizeof(*b));
    if (!b)
        return RVSP_ERROR_ALLOCATION_FAILED;
    rvsp_status_t s = RVSP_SUCCESS;
    const double limit = chunk_products();
    int32_t p0 = 0;
    while (p0 < count && s == RVSP_SUCCESS)
    {
        /* estimated products: nnz(A_i) times the mean row length of B_i */
        double est = 0.0;
        int32_t p1 = p0;
        while (p1 < count)
        {
            const double e = (double)A[p1].nnz * ((double)B[p1].nnz / (double)B[p1].rows);
            if (p1 > p0 && est + e > limit)
                break;
            est += e;
            p1++;
        }
        s = run_chunk(b, p0, p1, A, B, C, algo);
        p0 = p1;
    }
    if (s != RVSP_SUCCESS)
    {
        rvsp_spgemm_batch_destroy(b);
        for (int32_t i = 0; i < count; i++)
            memset(&C[i], 0, sizeof(C[i]));
        return s;
    }
    *batch = b;
    return RVSP_SUCCESS;
}
