/*
 * oracle.c -- CPU restatement of rv-sparse's FP32 scalar SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Written from the behaviour of the
 * rv-sparse sources cited per function; the code is this project's own.
 *
 * Must be built with -ffp-contract=off: every `acc + a * b` below is meant
 * as two IEEE-754 binary32 operations, each rounded to nearest-even.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_free(void *p) { free(p); }

/* rvsp_support.c:36-126: the first violation in row order wins; inside a row
 * the range test comes before the duplicate test, which comes before the
 * order test. */
int orc_csr_check(int32_t rows, int32_t cols, const int32_t *row_ptr,
                  const int32_t *col_idx, int32_t *bad_row_out)
{
    if (bad_row_out)
        *bad_row_out = -1;
    if (rows < 0 || cols < 0 || row_ptr == NULL || row_ptr[0] != 0)
        return 1;
    for (int32_t r = 0; r < rows; r++)
    {
        int32_t s = row_ptr[r], e = row_ptr[r + 1];
        if (e < s)
        {
            if (bad_row_out)
                *bad_row_out = r;
            return 1;
        }
        if (e == s)
            continue;
        if (col_idx == NULL)
            return 1;
        int32_t last = -1;
        for (int32_t p = s; p < e; p++)
        {
            int32_t c = col_idx[p];
            int code = 0;
            if (c < 0 || c >= cols)
                code = 2;
            else if (c == last)
                code = 4;
            else if (c < last)
                code = 3;
            if (code)
            {
                if (bad_row_out)
                    *bad_row_out = r;
                return code;
            }
            last = c;
        }
    }
    return 0;
}

/* Per-row distinct-column count with a byte mark array and a touched list
 * (rvsp_symbolic.c:27-73); the list lets each row clear only what it set. */
static int64_t count_row(int32_t r, const int32_t *a_rp, const int32_t *a_ci,
                         const int32_t *b_rp, const int32_t *b_ci,
                         unsigned char *mark, int32_t *touched,
                         int64_t *ops_out)
{
    int32_t nt = 0;
    int64_t ops = 0;
    for (int32_t ap = a_rp[r]; ap < a_rp[r + 1]; ap++)
    {
        int32_t k = a_ci[ap];
        ops += (int64_t)(b_rp[k + 1] - b_rp[k]);
        for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; bp++)
        {
            int32_t c = b_ci[bp];
            if (!mark[c])
            {
                mark[c] = 1;
                touched[nt++] = c;
            }
        }
    }
    for (int32_t t = 0; t < nt; t++)
        mark[touched[t]] = 0;
    *ops_out = ops;
    return nt;
}

int orc_symbolic_count(int32_t m, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const int32_t *b_rp,
                       const int32_t *b_ci, int32_t *c_rp, int64_t *ops,
                       int64_t *total_out)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    unsigned char *mark = (unsigned char *)calloc(nn, 1);
    int32_t *touched = (int32_t *)malloc(nn * sizeof(int32_t));
    if (!mark || !touched)
    {
        free(mark);
        free(touched);
        return -6;
    }
    int64_t total = 0;
    int rc = 0;
    for (int32_t r = 0; r < m; r++)
    {
        int64_t o;
        int64_t cnt = count_row(r, a_rp, a_ci, b_rp, b_ci, mark, touched, &o);
        if (ops)
            ops[r] = o;
        c_rp[r] = (int32_t)cnt;
        total += cnt;
        if (total > INT32_MAX)
        {
            rc = -6;
            break;
        }
    }
    free(mark);
    free(touched);
    if (rc)
        return rc;
    /* exclusive scan in place, total in the last slot (rvsp_symbolic.c:75-84) */
    int32_t run = 0;
    for (int32_t r = 0; r < m; r++)
    {
        int32_t c = c_rp[r];
        c_rp[r] = run;
        run += c;
    }
    c_rp[m] = run;
    if (total_out)
        *total_out = total;
    return 0;
}

/* Row-parallel orc_symbolic_count (OpenMP, per-thread mark arrays): the
 * same per-row counts and op counts, and the same exclusive scan and int32
 * overflow rule, evaluated over `threads` threads. Used to check products
 * whose size makes the sequential count too slow for a test. */
int orc_symbolic_count_par(int32_t m, int32_t n, const int32_t *a_rp,
                           const int32_t *a_ci, const int32_t *b_rp,
                           const int32_t *b_ci, int32_t *c_rp, int64_t *ops,
                           int64_t *total_out, int threads)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    int fail = 0;
    if (threads < 1)
        threads = 1;
#pragma omp parallel num_threads(threads)
    {
        unsigned char *mark = (unsigned char *)calloc(nn, 1);
        int32_t *touched = (int32_t *)malloc(nn * sizeof(int32_t));
        if (!mark || !touched)
        {
#pragma omp atomic write
            fail = 1;
        }
        else
        {
#pragma omp for schedule(dynamic, 64)
            for (int32_t r = 0; r < m; r++)
            {
                int64_t o;
                c_rp[r] = (int32_t)count_row(r, a_rp, a_ci, b_rp, b_ci, mark,
                                             touched, &o);
                if (ops)
                    ops[r] = o;
            }
        }
        free(mark);
        free(touched);
    }
    if (fail)
        return -6;
    int64_t run = 0;
    for (int32_t r = 0; r < m; r++)
    {
        int32_t c = c_rp[r];
        c_rp[r] = (int32_t)run;
        run += c;
        if (run > INT32_MAX)
            return -6;
    }
    c_rp[m] = (int32_t)run;
    if (total_out)
        *total_out = run;
    return 0;
}

/* Ascending sort of distinct column ids: insertion sort for short lists,
 * byte-wise LSD radix sort otherwise (rvsp_sort.h:16-124 uses the same two
 * regimes, with the switch at 40 entries and as many passes as the largest
 * column needs). Any correct sort yields the same order because the keys are
 * distinct; the two regimes are kept so the CPU baseline does comparable
 * work. */
static void sort_cols(int32_t *x, int32_t *tmp, int32_t cnt, int32_t max_col)
{
    if (cnt < 2)
        return;
    if (cnt <= 40)
    {
        for (int32_t i = 1; i < cnt; i++)
        {
            int32_t v = x[i], j = i;
            while (j > 0 && x[j - 1] > v)
            {
                x[j] = x[j - 1];
                j--;
            }
            x[j] = v;
        }
        return;
    }
    int passes = 1;
    while (passes < 4 && (max_col >> (8 * passes)) != 0)
        passes++;
    int32_t *src = x, *dst = tmp;
    for (int p = 0; p < passes; p++)
    {
        int32_t hist[257];
        memset(hist, 0, sizeof(hist));
        int sh = 8 * p;
        for (int32_t i = 0; i < cnt; i++)
            hist[((src[i] >> sh) & 255) + 1]++;
        for (int b = 0; b < 256; b++)
            hist[b + 1] += hist[b];
        for (int32_t i = 0; i < cnt; i++)
            dst[hist[(src[i] >> sh) & 255]++] = src[i];
        int32_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != x)
        memcpy(x, src, (size_t)cnt * sizeof(int32_t));
}

static void fill_row(int32_t r, int32_t n, const int32_t *a_rp,
                     const int32_t *a_ci, const int32_t *b_rp,
                     const int32_t *b_ci, const int32_t *c_rp, int32_t *c_ci,
                     unsigned char *mark, int32_t *tmp)
{
    int32_t *out = c_ci + c_rp[r];
    int32_t cnt = 0;
    for (int32_t ap = a_rp[r]; ap < a_rp[r + 1]; ap++)
    {
        int32_t k = a_ci[ap];
        for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; bp++)
        {
            int32_t c = b_ci[bp];
            if (!mark[c])
            {
                mark[c] = 1;
                out[cnt++] = c;
            }
        }
    }
    sort_cols(out, tmp, cnt, n > 0 ? n - 1 : 0);
    for (int32_t t = 0; t < cnt; t++)
        mark[out[t]] = 0;
}

void orc_symbolic_fill(int32_t m, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const int32_t *b_rp,
                       const int32_t *b_ci, const int32_t *c_rp,
                       int32_t *c_ci)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    unsigned char *mark = (unsigned char *)calloc(nn, 1);
    int32_t *tmp = (int32_t *)malloc(nn * sizeof(int32_t));
    if (mark && tmp)
        for (int32_t r = 0; r < m; r++)
            fill_row(r, n, a_rp, a_ci, b_rp, b_ci, c_rp, c_ci, mark, tmp);
    free(mark);
    free(tmp);
}

/* One output row of the numeric pass. acc[] is zero on entry and left zero
 * on exit, because the drain clears exactly the columns it reads
 * (gustavson_core_f32.inc:31-63). */
static void numeric_row(int32_t r, const int32_t *a_rp, const int32_t *a_ci,
                        const float *a_val, const int32_t *b_rp,
                        const int32_t *b_ci, const float *b_val,
                        const int32_t *c_rp, const int32_t *c_ci,
                        float *c_val, float *acc)
{
    for (int32_t ap = a_rp[r]; ap < a_rp[r + 1]; ap++)
    {
        const int32_t k = a_ci[ap];
        const float a = a_val[ap];
        for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; bp++)
        {
            const float prod = a * b_val[bp];
            acc[b_ci[bp]] = acc[b_ci[bp]] + prod;
        }
    }
    for (int32_t q = c_rp[r]; q < c_rp[r + 1]; q++)
    {
        c_val[q] = acc[c_ci[q]];
        acc[c_ci[q]] = 0.0f;
    }
}

void orc_numeric_f32(int32_t m, int32_t n, const int32_t *a_rp,
                     const int32_t *a_ci, const float *a_val,
                     const int32_t *b_rp, const int32_t *b_ci,
                     const float *b_val, const int32_t *c_rp,
                     const int32_t *c_ci, float *c_val)
{
    float *acc = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
    if (!acc)
        return;
    for (int32_t r = 0; r < m; r++)
        numeric_row(r, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp, c_ci,
                    c_val, acc);
    free(acc);
}

/* One output row of the FP64 yardstick: products and sums in double, the
 * same A-entry order and drain as numeric_row. */
static void numeric_row_f64(int32_t r, const int32_t *a_rp,
                            const int32_t *a_ci, const float *a_val,
                            const int32_t *b_rp, const int32_t *b_ci,
                            const float *b_val, const int32_t *c_rp,
                            const int32_t *c_ci, double *c_val, double *acc)
{
    for (int32_t ap = a_rp[r]; ap < a_rp[r + 1]; ap++)
    {
        int32_t k = a_ci[ap];
        double a = (double)a_val[ap];
        for (int32_t bp = b_rp[k]; bp < b_rp[k + 1]; bp++)
            acc[b_ci[bp]] += a * (double)b_val[bp];
    }
    for (int32_t q = c_rp[r]; q < c_rp[r + 1]; q++)
    {
        c_val[q] = acc[c_ci[q]];
        acc[c_ci[q]] = 0.0;
    }
}

void orc_numeric_f64(int32_t m, int32_t n, const int32_t *a_rp,
                     const int32_t *a_ci, const float *a_val,
                     const int32_t *b_rp, const int32_t *b_ci,
                     const float *b_val, const int32_t *c_rp,
                     const int32_t *c_ci, double *c_val)
{
    double *acc = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!acc)
        return;
    for (int32_t r = 0; r < m; r++)
        numeric_row_f64(r, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp, c_ci,
                        c_val, acc);
    free(acc);
}

/* Row-parallel fill + FP32 numeric (+ FP64 numeric when c_val64 != NULL)
 * on a row_ptr from the symbolic count: every row is computed by the same
 * fill_row / numeric_row code as orc_symbolic_fill / orc_numeric_f32, so the
 * outputs are identical; rows are shared out over `threads` OpenMP threads.
 * Returns 0, or -6 when a thread's scratch cannot be allocated. */
int orc_fill_numeric_par(int32_t m, int32_t n, const int32_t *a_rp,
                         const int32_t *a_ci, const float *a_val,
                         const int32_t *b_rp, const int32_t *b_ci,
                         const float *b_val, const int32_t *c_rp,
                         int32_t *c_ci, float *c_val, double *c_val64,
                         int threads)
{
    size_t nn = (size_t)(n > 0 ? n : 1);
    int fail = 0;
    if (threads < 1)
        threads = 1;
#pragma omp parallel num_threads(threads)
    {
        unsigned char *mark = (unsigned char *)calloc(nn, 1);
        int32_t *tmp = (int32_t *)malloc(nn * sizeof(int32_t));
        float *acc = (float *)calloc(nn, sizeof(float));
        double *acc64 = c_val64 ? (double *)calloc(nn, sizeof(double)) : NULL;
        if (!mark || !tmp || !acc || (c_val64 && !acc64))
        {
#pragma omp atomic write
            fail = 1;
        }
        else
        {
#pragma omp for schedule(dynamic, 16)
            for (int32_t r = 0; r < m; r++)
            {
                fill_row(r, n, a_rp, a_ci, b_rp, b_ci, c_rp, c_ci, mark, tmp);
                numeric_row(r, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp,
                            c_ci, c_val, acc);
                if (c_val64)
                    numeric_row_f64(r, a_rp, a_ci, a_val, b_rp, b_ci, b_val,
                                    c_rp, c_ci, c_val64, acc64);
            }
        }
        free(mark);
        free(tmp);
        free(acc);
        free(acc64);
    }
    return fail ? -6 : 0;
}

int orc_spgemm_f32(int32_t m, int32_t k, int32_t n, const int32_t *a_rp,
                   const int32_t *a_ci, const float *a_val,
                   const int32_t *b_rp, const int32_t *b_ci,
                   const float *b_val, int32_t **c_rp_out,
                   int32_t **c_ci_out, float **c_val_out,
                   int32_t *c_nnz_out)
{
    if (m <= 0 || k <= 0 || n <= 0)
        return -2;
    int32_t *c_rp = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    if (!c_rp)
        return -6;
    int64_t total = 0;
    int rc = orc_symbolic_count(m, n, a_rp, a_ci, b_rp, b_ci, c_rp, NULL,
                                &total);
    if (rc)
    {
        free(c_rp);
        return rc;
    }
    size_t alloc = total > 0 ? (size_t)total : 1;
    int32_t *c_ci = (int32_t *)malloc(alloc * sizeof(int32_t));
    float *c_val = (float *)malloc(alloc * sizeof(float));
    if (!c_ci || !c_val)
    {
        free(c_rp);
        free(c_ci);
        free(c_val);
        return -6;
    }
    orc_symbolic_fill(m, n, a_rp, a_ci, b_rp, b_ci, c_rp, c_ci);
    orc_numeric_f32(m, n, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp, c_ci,
                    c_val);
    *c_rp_out = c_rp;
    *c_ci_out = c_ci;
    *c_val_out = c_val;
    *c_nnz_out = (int32_t)total;
    return 0;
}

int orc_spgemm_f32_omp(int32_t m, int32_t k, int32_t n, const int32_t *a_rp,
                       const int32_t *a_ci, const float *a_val,
                       const int32_t *b_rp, const int32_t *b_ci,
                       const float *b_val, int32_t **c_rp_out,
                       int32_t **c_ci_out, float **c_val_out,
                       int32_t *c_nnz_out, int threads)
{
    if (m <= 0 || k <= 0 || n <= 0)
        return -2;
    if (threads < 1)
        threads = 1;
    size_t nn = (size_t)n;
    int32_t *c_rp = (int32_t *)calloc((size_t)m + 1, sizeof(int32_t));
    if (!c_rp)
        return -6;
    int fail = 0;
#pragma omp parallel num_threads(threads)
    {
        unsigned char *mark = (unsigned char *)calloc(nn, 1);
        int32_t *touched = (int32_t *)malloc(nn * sizeof(int32_t));
        if (!mark || !touched)
        {
#pragma omp atomic write
            fail = 1;
        }
        else
        {
#pragma omp for schedule(dynamic, 64)
            for (int32_t r = 0; r < m; r++)
            {
                int64_t o;
                c_rp[r] = (int32_t)count_row(r, a_rp, a_ci, b_rp, b_ci, mark,
                                             touched, &o);
            }
        }
        free(mark);
        free(touched);
    }
    if (fail)
    {
        free(c_rp);
        return -6;
    }
    int64_t run = 0;
    for (int32_t r = 0; r < m; r++)
    {
        int32_t c = c_rp[r];
        c_rp[r] = (int32_t)run;
        run += c;
        if (run > INT32_MAX)
        {
            free(c_rp);
            return -6;
        }
    }
    c_rp[m] = (int32_t)run;
    size_t alloc = run > 0 ? (size_t)run : 1;
    int32_t *c_ci = (int32_t *)malloc(alloc * sizeof(int32_t));
    float *c_val = (float *)malloc(alloc * sizeof(float));
    if (!c_ci || !c_val)
    {
        free(c_rp);
        free(c_ci);
        free(c_val);
        return -6;
    }
#pragma omp parallel num_threads(threads)
    {
        unsigned char *mark = (unsigned char *)calloc(nn, 1);
        int32_t *tmp = (int32_t *)malloc(nn * sizeof(int32_t));
        float *acc = (float *)calloc(nn, sizeof(float));
        if (!mark || !tmp || !acc)
        {
#pragma omp atomic write
            fail = 1;
        }
        else
        {
#pragma omp for schedule(dynamic, 64)
            for (int32_t r = 0; r < m; r++)
            {
                fill_row(r, n, a_rp, a_ci, b_rp, b_ci, c_rp, c_ci, mark, tmp);
                numeric_row(r, a_rp, a_ci, a_val, b_rp, b_ci, b_val, c_rp,
                            c_ci, c_val, acc);
            }
        }
        free(mark);
        free(tmp);
        free(acc);
    }
    if (fail)
    {
        free(c_rp);
        free(c_ci);
        free(c_val);
        return -6;
    }
    *c_rp_out = c_rp;
    *c_ci_out = c_ci;
    *c_val_out = c_val;
    *c_nnz_out = (int32_t)run;
    return 0;
}
