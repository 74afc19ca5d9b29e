/*
 * genmat.c -- rv-sparse's synthetic matrix generator (tools/src/genmat.c),
 * the source of its bench's `--gen R C DENSITY SEED` inputs. The matrices
 * are the reference generator's, bit for bit, for every non-symmetric
 * request. A matrix is built line by line (a line is a row, or a column
 * with is_column), each line from its own seeds, so lines are independent:
 *
 *  - degree (genmat.c:56-115): a normal draw when avg - 3 sd > 0, else a
 *    log-normal one, by Box-Muller over two xorshift32 streams seeded from
 *    the line's seed; rounded, then clamped to [min, max];
 *  - pattern: one xorshift32 stream seeded with seed + degree % 10, drawing
 *    `degree - 1` indices (square; the diagonal is always stored) or
 *    `degree` indices (rectangular), each `x % span + lo` over the line's
 *    window, an index drawn twice stored once. A square line's window is
 *    the band [max(0, j - low), min(n, j + up)): genmat.c splits the rows
 *    into three ranges per banding case (:372-580), and each range's window
 *    is this one; the case only sets the seed multiplier (4 when
 *    low > n - up, else 5; rectangular 6);
 *  - values (genmat.c:914-938): glibc's rand() stream after srand(seed),
 *    one draw r per stored entry in line order, (r % 9 + 1.0) / 10. Drawn
 *    here from a private state with initstate_r / random_r (the same TYPE_3
 *    additive generator rand() runs on, seeded the same way), so the
 *    caller's rand() stream is not disturbed.
 *
 * The reference's implicit conversions are spelled out as x86-64 performs
 * them (double -> unsigned through a signed 64-bit truncation, double ->
 * int giving INT_MIN out of range), and its int seed products wrap.
 * Symmetric requests are not built: they return the error matrix. (For an
 * odd order the reference leaves a row pointer unset and draws from
 * rand() ahead of the value stream, genmat.c:214-244, so there is no single
 * result to match.)
 */
#define _DEFAULT_SOURCE /* initstate_r, random_r */
#include "genmat.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"

static uint32_t xs32_next(uint32_t *s)
{
    uint32_t x = *s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return *s = x;
}

/* a zero xorshift state stays zero; the reference substitutes this one */
static uint32_t xs32_state(uint32_t s)
{
    return s ? s : 0x9e3779b9u;
}

static uint32_t dbl_to_u32(double v)
{
    if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)v;
}

static int32_t dbl_to_i32(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0))
        return INT32_MIN;
    return (int32_t)v;
}

/* one Box-Muller draw: genmat.c:56-86 */
static double box_muller(double mean, double sd, int32_t seed)
{
    uint32_t s = xs32_state((uint32_t)seed * 10u);
    double u = 0.0;
    while (u == 0.0)
        u = (double)xs32_next(&s) / 4294967295.0;
    const double r = sqrt(-2.0 * log(u));
    s = xs32_state(dbl_to_u32(floor(mean) * (double)seed));
    double theta = 0.0;
    while (theta == 0.0)
        theta = 6.2831853 * (double)xs32_next(&s) / 4294967295.0;
    const double x = r * cos(theta);
    return x * sd + mean;
}

/* genmat.c:88-115. The log-normal spread is evaluated as the reference
 * writes it, 1 + sd * sd / avg * avg = 1 + ((sd * sd) / avg) * avg. */
static int line_degree(double avg, double sd, int lo, int hi, int32_t seed)
{
    double v;
    if (avg - 3 * sd > 0)
        v = box_muller(avg, sd, seed);
    else
    {
        const double mu = log(avg * avg / sqrt(avg * avg + sd * sd));
        const double sig = sqrt(log(1 + sd * sd / avg * avg));
        v = exp(box_muller(mu, sig, seed));
    }
    int d = dbl_to_i32(round(v));
    if (d < lo)
        d = lo;
    if (d > hi)
        d = hi;
    return d;
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

typedef struct {
    int *idx;
    size_t len, cap;
} ivec_t;

static int ivec_reserve(ivec_t *v, size_t extra)
{
    if (v->len + extra <= v->cap)
        return 0;
    size_t cap = v->cap ? v->cap : 1024;
    while (cap < v->len + extra)
        cap *= 2;
    int *p = (int *)realloc(v->idx, cap * sizeof(int));
    if (!p)
        return -1;
    v->idx = p;
    v->cap = cap;
    return 0;
}

static csr_matrix_t error_matrix(void)
{
    csr_matrix_t m;
    memset(&m, 0, sizeof m);
    m.nrows = -1;
    m.ncols = -1;
    m.nnz = -1;
    return m;
}

genmat_params_t genmat_default_params(int row_cnt, int col_cnt)
{
    genmat_params_t p;
    p.row_cnt = row_cnt;
    p.col_cnt = col_cnt;
    p.density = 0.001;
    p.cv = 0.5;
    p.min = 1;
    p.imbalance = -1.0;
    p.is_symmetric = 0;
    p.is_column = 0;
    p.low_bandwidth = row_cnt - 1;
    p.up_bandwidth = row_cnt - 1;
    p.random_seed = 1;
    return p;
}

csr_matrix_t genmat_generate_csr(genmat_params_t p)
{
    const int square = p.row_cnt == p.col_cnt;
    if (p.is_symmetric || p.row_cnt < 0 || p.col_cnt < 0)
        return error_matrix();
    /* negative bands index before a line's start (an empty matrix has none) */
    if (square && p.row_cnt > 0 && (p.low_bandwidth < 0 || p.up_bandwidth < 0))
        return error_matrix();
    /* lines: rows, or columns with is_column; width: the other extent */
    const int swap = p.is_column && !square;
    const int lines = swap ? p.col_cnt : p.row_cnt;
    const int width = swap ? p.row_cnt : p.col_cnt;
    const double avg = p.density * width;
    const double sd = p.cv * avg;
    int dmax = width - 1;
    if (p.imbalance != -1.0)
        dmax = dbl_to_i32(round(p.imbalance * avg + avg));
    if (dmax > width - 1)
        dmax = width - 1;
    const int low = p.low_bandwidth, up = p.up_bandwidth;
    const uint32_t mult = !square ? 6u : (low > p.row_cnt - up ? 4u : 5u);

    ivec_t ent = {0};
    int *lptr = (int *)malloc(((size_t)lines + 1) * sizeof(int));
    if (!lptr)
        return error_matrix();
    lptr[0] = 0;
    for (int j = 0; j < lines; j++)
    {
        const int32_t seed = (int32_t)((uint32_t)p.random_seed * (uint32_t)(j + (int)mult));
        const int d = line_degree(avg, sd, p.min, dmax, seed);
        uint32_t st = xs32_state((uint32_t)seed + (uint32_t)(d % 10));
        int lo = 0, span = width, draws = d;
        if (square)
        {
            lo = j - low > 0 ? j - low : 0;
            const int hi = (long long)j + up < width ? j + up : width;
            span = hi - lo;
            draws = d - 1;
        }
        if (draws > 0 && span <= 0)
            goto fail; /* the reference divides by zero here */
        const size_t n0 = ent.len;
        if (ivec_reserve(&ent, (size_t)(draws > 0 ? draws : 0) + 1))
            goto fail;
        if (square)
            ent.idx[ent.len++] = j;
        for (int t = 0; t < draws; t++)
            ent.idx[ent.len++] = (int)(xs32_next(&st) % (uint32_t)span) + lo;
        /* ascending, each index once */
        int *L = ent.idx + n0;
        const size_t n = ent.len - n0;
        qsort(L, n, sizeof(int), cmp_int);
        size_t w = 0;
        for (size_t t = 0; t < n; t++)
            if (w == 0 || L[t] != L[w - 1])
                L[w++] = L[t];
        ent.len = n0 + w;
        if (ent.len > (size_t)INT32_MAX)
            goto fail;
        lptr[j + 1] = (int)ent.len;
    }

    const size_t nnz = ent.len;
    csr_matrix_t m;
    m.nrows = p.row_cnt;
    m.ncols = p.col_cnt;
    m.nnz = (long long)nnz;
    m.row_ptr = (int *)calloc((size_t)p.row_cnt + 1, sizeof(int));
    m.col_idx = (int *)malloc((nnz ? nnz : 1) * sizeof(int));
    m.values = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    double *lval = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    if (!m.row_ptr || !m.col_idx || !m.values || !lval)
    {
        free(m.row_ptr);
        free(m.col_idx);
        free(m.values);
        free(lval);
        goto fail;
    }
    /* values in line order */
    {
        struct random_data rd;
        char state[128]; /* 128 bytes: the TYPE_3 generator of rand() */
        memset(&rd, 0, sizeof rd);
        initstate_r((unsigned int)p.random_seed, state, sizeof state, &rd);
        for (size_t t = 0; t < nnz; t++)
        {
            int32_t r;
            random_r(&rd, &r);
            lval[t] = (r % 9 + 1.0) / 10;
        }
    }
    if (!p.is_column)
    {
        memcpy(m.row_ptr, lptr, ((size_t)p.row_cnt + 1) * sizeof(int));
        memcpy(m.col_idx, ent.idx, nnz * sizeof(int));
        memcpy(m.values, lval, nnz * sizeof(double));
    }
    else
    {
        /* lines are columns: entries reach each row in ascending column
         * order, as the reference's stable per-row sort leaves them */
        for (size_t t = 0; t < nnz; t++)
            m.row_ptr[ent.idx[t] + 1]++;
        for (int i = 0; i < p.row_cnt; i++)
            m.row_ptr[i + 1] += m.row_ptr[i];
        int *cur = (int *)malloc(((size_t)p.row_cnt + 1) * sizeof(int));
        if (!cur)
        {
            free(m.row_ptr);
            free(m.col_idx);
            free(m.values);
            free(lval);
            goto fail;
        }
        memcpy(cur, m.row_ptr, ((size_t)p.row_cnt + 1) * sizeof(int));
        for (int c = 0; c < lines; c++)
            for (int t = lptr[c]; t < lptr[c + 1]; t++)
            {
                const int at = cur[ent.idx[t]]++;
                m.col_idx[at] = c;
                m.values[at] = lval[t];
            }
        free(cur);
    }
    free(lval);
    free(lptr);
    free(ent.idx);
    return m;

fail:
    free(lptr);
    free(ent.idx);
    return error_matrix();
}

void genmat_free_csr(csr_matrix_t *m)
{
    if (!m)
        return;
    free(m->row_ptr);
    free(m->col_idx);
    free(m->values);
    memset(m, 0, sizeof *m);
}

int genmat_save_csr_to_mtx(const csr_matrix_t *csr, const char *filename)
{
    if (!csr || !filename || csr->nrows < 0)
        return -1;
    FILE *f = fopen(filename, "w");
    if (!f)
        return -1;
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f, "%d %d %lld\n", csr->nrows, csr->ncols, csr->nnz);
    for (int i = 0; i < csr->nrows; i++)
        for (int t = csr->row_ptr[i]; t < csr->row_ptr[i + 1]; t++)
            fprintf(f, "%d %d %.16e\n", i + 1, csr->col_idx[t] + 1, csr->values[t]);
    return fclose(f) == 0 ? 0 : -1;
}

/* bench/bench.c:391-450 */
static rvsp_status_t to_rvsp(csr_matrix_t *g, rvsp_dtype_t dtype, rvsp_csr_matrix_t *out)
{
    const size_t nnz = (size_t)g->nnz;
    const size_t vb = dtype == RVSP_DTYPE_FP64 ? 8 : dtype == RVSP_DTYPE_INT8 ? 1 : 4;
    int32_t *rp = (int32_t *)malloc(((size_t)g->nrows + 1) * sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc((nnz ? nnz : 1) * sizeof(int32_t));
    void *va = malloc((nnz ? nnz : 1) * vb);
    if (!rp || !ci || !va)
    {
        free(rp);
        free(ci);
        free(va);
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    memcpy(rp, g->row_ptr, ((size_t)g->nrows + 1) * sizeof(int32_t));
    memcpy(ci, g->col_idx, nnz * sizeof(int32_t));
    for (size_t t = 0; t < nnz; t++)
    {
        const double v = g->values[t];
        if (dtype == RVSP_DTYPE_FP32)
            ((float *)va)[t] = (float)v;
        else if (dtype == RVSP_DTYPE_FP64)
            ((double *)va)[t] = v;
        else
        {
            int iv = (int)v % 127;
            ((int8_t *)va)[t] = (int8_t)(iv == 0 ? 1 : iv);
        }
    }
    out->rows = g->nrows;
    out->cols = g->ncols;
    out->nnz = (int32_t)nnz;
    out->row_ptr = rp;
    out->col_idx = ci;
    out->values = va;
    out->dtype = dtype;
    out->format = RVSP_FORMAT_CSR;
    out->owns_data = 1;
    return RVSP_SUCCESS;
}

rvsp_status_t rvsp_genmat_bench_pair(int32_t rows, int32_t cols, double density, int32_t seed,
                                     rvsp_dtype_t dtype, rvsp_csr_matrix_t *A,
                                     rvsp_csr_matrix_t *B)
{
    if (!A || !B)
        return RVSP_ERROR_NULL_POINTER;
    if (dtype != RVSP_DTYPE_FP32 && dtype != RVSP_DTYPE_FP64 && dtype != RVSP_DTYPE_INT8)
        return RVSP_ERROR_UNSUPPORTED_DTYPE;
    if (rows < 0 || cols < 0 || !(density >= 0.0))
        return RVSP_ERROR_INVALID_ARGUMENT;
    memset(A, 0, sizeof *A);
    memset(B, 0, sizeof *B);
    genmat_params_t pa = genmat_default_params(rows, cols);
    pa.density = density;
    pa.random_seed = seed;
    pa.cv = 0.5;
    pa.min = 1;
    genmat_params_t pb = genmat_default_params(cols, rows);
    pb.density = density;
    pb.random_seed = (int32_t)((uint32_t)seed + 1u);
    pb.cv = 0.5;
    pb.min = 1;
    csr_matrix_t ga = genmat_generate_csr(pa);
    csr_matrix_t gb = genmat_generate_csr(pb);
    rvsp_status_t s = RVSP_SUCCESS;
    if (ga.nrows < 0 || gb.nrows < 0 || ga.nnz > INT32_MAX || gb.nnz > INT32_MAX)
        s = RVSP_ERROR_INVALID_ARGUMENT;
    if (s == RVSP_SUCCESS && (s = to_rvsp(&ga, dtype, A)) == RVSP_SUCCESS &&
        (s = to_rvsp(&gb, dtype, B)) != RVSP_SUCCESS)
        rvsp_csr_destroy(A);
    genmat_free_csr(&ga);
    genmat_free_csr(&gb);
    return s;
}
