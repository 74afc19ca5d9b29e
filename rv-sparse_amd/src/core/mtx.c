/*
 * mtx.c -- Matrix Market file to canonical FP32 CSR on the host, for the
 * real-matrix products rv-sparse's bench runs. Follows the reading of
 * rv-sparse tools/src/mtx_to_csr_formatter.c:147-325 (banner, comments,
 * entries, symmetric mirror, per-row sort and duplicate sum,
 * canonicalise_rows at :77-145) and the column inference of
 * bench/bench.c:465-469; see rvsp_mtx.h for the contract.
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rv_sparse.h"
#include "rvsp_mtx.h"

typedef struct
{
    int32_t col;
    int32_t seq; /* position in the row before sorting: keeps the sort stable */
    float val;
} mtx_entry_t;

static int cmp_entry(const void *a, const void *b)
{
    const mtx_entry_t *x = (const mtx_entry_t *)a, *y = (const mtx_entry_t *)b;
    if (x->col != y->col)
        return (x->col > y->col) - (x->col < y->col);
    return (x->seq > y->seq) - (x->seq < y->seq);
}

/* "%%MatrixMarket object format field symmetry" -> pattern / symmetric /
 * complex flags; a line without the banner leaves all three 0. */
static void read_banner(const char *line, int *pattern, int *symmetric, int *complex_field)
{
    char object[64], format[64], field[64], symmetry[64];
    *pattern = *symmetric = *complex_field = 0;
    if (strncmp(line, "%%MatrixMarket", 14) != 0)
        return;
    if (sscanf(line, "%%%%MatrixMarket %63s %63s %63s %63s", object, format, field, symmetry) < 4)
        return;
    *pattern = strcmp(field, "pattern") == 0;
    *complex_field = strcmp(field, "complex") == 0;
    *symmetric = strcmp(symmetry, "symmetric") == 0;
}

rvsp_status_t rvsp_mtx_load(const char *path, rvsp_csr_matrix_t *out)
{
    if (path == NULL || out == NULL)
        return RVSP_ERROR_NULL_POINTER;
    memset(out, 0, sizeof(*out));

    FILE *f = fopen(path, "r");
    if (f == NULL)
        return RVSP_ERROR_INVALID_ARGUMENT;

    /* the first line is the banner; then comment lines up to the sizes */
    char line[2048];
    int pattern = 0, symmetric = 0, complex_field = 0;
    if (fgets(line, sizeof line, f) != NULL)
        read_banner(line, &pattern, &symmetric, &complex_field);
    int have_dims = 0;
    while (fgets(line, sizeof line, f) != NULL)
        if (line[0] != '%')
        {
            have_dims = 1;
            break;
        }
    int M = 0, N = 0, L = 0;
    if (!have_dims || complex_field || sscanf(line, "%d %d %d", &M, &N, &L) != 3 || M <= 0 ||
        L < 0 || (symmetric && L > INT32_MAX / 2))
    {
        fclose(f);
        return RVSP_ERROR_INVALID_ARGUMENT;
    }

    const size_t cap = (size_t)L * (symmetric ? 2 : 1);
    int32_t *er = (int32_t *)malloc((cap ? cap : 1) * sizeof(int32_t));
    int32_t *ec = (int32_t *)malloc((cap ? cap : 1) * sizeof(int32_t));
    float *ev = (float *)malloc((cap ? cap : 1) * sizeof(float));
    int32_t *rp = (int32_t *)calloc((size_t)M + 1, sizeof(int32_t));
    rvsp_status_t st = RVSP_SUCCESS;
    if (!er || !ec || !ev || !rp)
    {
        st = RVSP_ERROR_ALLOCATION_FAILED;
        goto done;
    }

    /* entries: stop at the declared count or the first unreadable one */
    int32_t n = 0, maxcol = -1;
    for (int l = 0; l < L; l++)
    {
        int r, c;
        float v = 1.0f;
        if (pattern ? fscanf(f, "%d %d", &r, &c) != 2 : fscanf(f, "%d %d %f", &r, &c, &v) != 3)
            break;
        if (r < 1 || r > M || c < 1 || (symmetric && c > M))
        {
            st = RVSP_ERROR_INVALID_ARGUMENT;
            goto done;
        }
        er[n] = r - 1;
        ec[n] = c - 1;
        ev[n] = v;
        n++;
    }
    if (symmetric) /* mirror the off-diagonal entries, after all stored ones */
    {
        const int32_t stored = n;
        for (int32_t k = 0; k < stored; k++)
            if (er[k] != ec[k])
            {
                er[n] = ec[k];
                ec[n] = er[k];
                ev[n] = ev[k];
                n++;
            }
    }

    /* rows by counting, entries in the order read (stable) */
    for (int32_t k = 0; k < n; k++)
        rp[er[k] + 1]++;
    for (int32_t i = 0; i < M; i++)
        rp[i + 1] += rp[i];
    int32_t *ci = (int32_t *)malloc((n ? (size_t)n : 1) * sizeof(int32_t));
    float *va = (float *)malloc((n ? (size_t)n : 1) * sizeof(float));
    int32_t *cur = (int32_t *)malloc((size_t)M * sizeof(int32_t));
    int32_t maxlen = 0;
    for (int32_t i = 0; i < M; i++)
        if (rp[i + 1] - rp[i] > maxlen)
            maxlen = rp[i + 1] - rp[i];
    mtx_entry_t *buf = (mtx_entry_t *)malloc((maxlen ? (size_t)maxlen : 1) * sizeof(mtx_entry_t));
    if (!ci || !va || !cur || !buf)
    {
        free(ci);
        free(va);
        free(cur);
        free(buf);
        st = RVSP_ERROR_ALLOCATION_FAILED;
        goto done;
    }
    memcpy(cur, rp, (size_t)M * sizeof(int32_t));
    for (int32_t k = 0; k < n; k++)
    {
        const int32_t d = cur[er[k]]++;
        ci[d] = ec[k];
        va[d] = ev[k];
    }
    free(cur);

    /* per row: sort by column (stable), sum repeated columns in order */
    int32_t w = 0;
    for (int32_t i = 0; i < M; i++)
    {
        const int32_t s = rp[i], len = rp[i + 1] - s;
        for (int32_t t = 0; t < len; t++)
        {
            buf[t].col = ci[s + t];
            buf[t].seq = t;
            buf[t].val = va[s + t];
        }
        qsort(buf, (size_t)len, sizeof(mtx_entry_t), cmp_entry);
        rp[i] = w;
        for (int32_t t = 0; t < len;)
        {
            const int32_t col = buf[t].col;
            float acc = buf[t].val;
            for (t++; t < len && buf[t].col == col; t++)
                acc += buf[t].val;
            ci[w] = col;
            va[w] = acc;
            if (col > maxcol)
                maxcol = col;
            w++;
        }
    }
    rp[M] = w;
    free(buf);

    out->rows = M;
    out->cols = w > 0 ? maxcol + 1 : N;
    out->nnz = w;
    out->row_ptr = rp;
    out->col_idx = ci;
    out->values = va;
    out->dtype = RVSP_DTYPE_FP32;
    out->format = RVSP_FORMAT_CSR;
    out->owns_data = 1;
    rp = NULL; /* handed over */

done:
    fclose(f);
    free(er);
    free(ec);
    free(ev);
    free(rp);
    return st;
}
