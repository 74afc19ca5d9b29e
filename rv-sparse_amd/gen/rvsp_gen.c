/*
 * rvsp_gen.c -- seeded synthetic CSR generators for the benchmark workloads.
 *
 * One project PRNG (splitmix64 seeding xoshiro256**) so the CPU checker and
 * the GPU see identical inputs for a given seed. Every generator emits
 * canonical CSR (ascending, unique columns per row) in malloc'd arrays that
 * the caller releases with rvsp_gen_free. Values are either U[-1,1) on a
 * 2^-23 grid or N(0,1) by Box-Muller, drawn in stored order.
 *
 * Workloads:
 *   random rows   each row draws L in [lo, hi] distinct uniform columns
 *                 (Floyd's sampling, then sorted)
 *   R-MAT         2^scale square, `edges` recursive quadrant draws with
 *                 probabilities (a, b, c, 1-a-b-c), duplicates merged
 *   stencil27     27-point stencil on a g^3 grid, index x + g*y + g*g*z
 *   zipf rows     m x n, rank-based lengths clamp(round(c*r^-s), 1, maxlen)
 *                 with c solved so the mean length is `mean`, rows shuffled
 *   banded        row k takes `per_row` distinct columns in
 *                 [stride*k - half, stride*k + half) clipped to [0, n)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct
{
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++)
        r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t out = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return out;
}

/* uniform integer in [0, bound) by 128-bit multiply (Lemire, no rejection;
 * the bias is below 2^-32 for the bounds used here) */
static inline uint64_t rng_below(rng_t *r, uint64_t bound)
{
    return (uint64_t)(((unsigned __int128)rng_next(r) * bound) >> 64);
}

static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* U[-1, 1) on a 2^-23 grid: exact in binary32 */
static inline float rng_sym(rng_t *r)
{
    int32_t q = (int32_t)(rng_next(r) >> 40); /* 24 bits */
    return (float)(q - (1 << 23)) * 0x1.0p-23f;
}

static inline float rng_normal(rng_t *r)
{
    double u1 = rng_unit(r), u2 = rng_unit(r);
    if (u1 < 1e-300)
        u1 = 1e-300;
    return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

void rvsp_gen_free(void *p) { free(p); }

/* dist: 0 = U[-1,1), 1 = N(0,1) */
void rvsp_gen_fill_values(float *val, int64_t count, uint64_t seed, int dist)
{
    rng_t r;
    rng_seed(&r, seed);
    for (int64_t i = 0; i < count; i++)
        val[i] = dist ? rng_normal(&r) : rng_sym(&r);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Floyd's sampling of `want` distinct values in [lo, lo + span), sorted,
 * written to out. mark is a zeroed bitmap of >= span bits, left zeroed. */
static void sample_distinct(rng_t *r, int32_t lo, int32_t span, int32_t want,
                            int32_t *out, uint64_t *mark)
{
    int32_t got = 0;
    for (int32_t j = span - want; j < span; j++)
    {
        int32_t t = (int32_t)rng_below(r, (uint64_t)j + 1);
        int32_t pick = (mark[t >> 6] >> (t & 63)) & 1 ? j : t;
        mark[pick >> 6] |= 1ull << (pick & 63);
        out[got++] = pick;
    }
    for (int32_t i = 0; i < got; i++)
        mark[out[i] >> 6] &= ~(1ull << (out[i] & 63));
    qsort(out, (size_t)got, sizeof(int32_t), cmp_i32);
    for (int32_t i = 0; i < got; i++)
        out[i] += lo;
}

/* Rows with L ~ U{lo..hi} distinct uniform columns in [0, n). Returns nnz
 * or -1 on allocation failure. */
int64_t rvsp_gen_random_rows(int32_t m, int32_t n, int32_t lo, int32_t hi,
                             uint64_t seed, int dist, int32_t **rp_out,
                             int32_t **ci_out, float **val_out)
{
    rng_t r;
    rng_seed(&r, seed);
    if (hi > n)
        hi = n;
    if (lo > hi)
        lo = hi;
    int32_t *rp = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    if (!rp)
        return -1;
    rp[0] = 0;
    for (int32_t i = 0; i < m; i++)
        rp[i + 1] = rp[i] + lo + (int32_t)rng_below(&r, (uint64_t)(hi - lo + 1));
    int64_t nnz = rp[m];
    int32_t *ci = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    float *val = (float *)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    uint64_t *mark = (uint64_t *)calloc(((size_t)n + 63) / 64, 8);
    if (!ci || !val || !mark)
    {
        free(rp);
        free(ci);
        free(val);
        free(mark);
        return -1;
    }
    for (int32_t i = 0; i < m; i++)
        sample_distinct(&r, 0, n, rp[i + 1] - rp[i], ci + rp[i], mark);
    for (int64_t p = 0; p < nnz; p++)
        val[p] = dist ? rng_normal(&r) : rng_sym(&r);
    free(mark);
    *rp_out = rp;
    *ci_out = ci;
    *val_out = val;
    return nnz;
}

static void radix_sort_u64(uint64_t *key, uint64_t *tmp, int64_t n, int bits)
{
    uint64_t *src = key, *dst = tmp;
    for (int sh = 0; sh < bits; sh += 11)
    {
        int64_t *h = (int64_t *)calloc(2049, sizeof(int64_t));
        for (int64_t i = 0; i < n; i++)
            h[((src[i] >> sh) & 2047) + 1]++;
        for (int b = 0; b < 2048; b++)
            h[b + 1] += h[b];
        for (int64_t i = 0; i < n; i++)
            dst[h[(src[i] >> sh) & 2047]++] = src[i];
        free(h);
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != key)
        memcpy(key, src, (size_t)n * sizeof(uint64_t));
}

/* R-MAT, 2^scale x 2^scale. Quadrant per level from one uniform draw:
 * [0,a) top-left, [a,a+b) top-right, [a+b,a+b+c) bottom-left, else
 * bottom-right. Duplicates merged; values drawn per unique entry in sorted
 * order. */
int64_t rvsp_gen_rmat(int32_t scale, int64_t edges, double a, double b,
                      double c, uint64_t seed, int dist, int32_t **rp_out,
                      int32_t **ci_out, float **val_out)
{
    rng_t r;
    rng_seed(&r, seed);
    int64_t nrow = (int64_t)1 << scale;
    uint64_t *key = (uint64_t *)malloc((size_t)edges * sizeof(uint64_t));
    uint64_t *tmp = (uint64_t *)malloc((size_t)edges * sizeof(uint64_t));
    if (!key || !tmp)
    {
        free(key);
        free(tmp);
        return -1;
    }
    const double ab = a + b, abc = a + b + c;
    for (int64_t e = 0; e < edges; e++)
    {
        uint64_t row = 0, col = 0;
        for (int l = 0; l < scale; l++)
        {
            double u = rng_unit(&r);
            uint64_t rb = u >= ab, cb = (u >= a && u < ab) || u >= abc;
            row = (row << 1) | rb;
            col = (col << 1) | cb;
        }
        key[e] = (row << scale) | col;
    }
    radix_sort_u64(key, tmp, edges, 2 * scale);
    free(tmp);
    int64_t nnz = 0;
    for (int64_t e = 0; e < edges; e++)
        if (e == 0 || key[e] != key[e - 1])
            key[nnz++] = key[e];
    int32_t *rp = (int32_t *)calloc((size_t)nrow + 1, sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    float *val = (float *)malloc((size_t)(nnz ? nnz : 1) * sizeof(float));
    if (!rp || !ci || !val)
    {
        free(key);
        free(rp);
        free(ci);
        free(val);
        return -1;
    }
    const uint64_t cmask = ((uint64_t)1 << scale) - 1;
    for (int64_t p = 0; p < nnz; p++)
    {
        rp[(key[p] >> scale) + 1]++;
        ci[p] = (int32_t)(key[p] & cmask);
        val[p] = dist ? rng_normal(&r) : rng_sym(&r);
    }
    for (int64_t i = 0; i < nrow; i++)
        rp[i + 1] += rp[i];
    free(key);
    *rp_out = rp;
    *ci_out = ci;
    *val_out = val;
    return nnz;
}

/* 27-point stencil on a g^3 grid. Columns come out ascending because the
 * neighbour offsets are visited z-major, then y, then x. */
int64_t rvsp_gen_stencil27(int32_t g, uint64_t seed, int dist,
                           int32_t **rp_out, int32_t **ci_out,
                           float **val_out)
{
    int64_t m = (int64_t)g * g * g;
    int32_t *rp = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc((size_t)m * 27 * sizeof(int32_t));
    if (!rp || !ci)
    {
        free(rp);
        free(ci);
        return -1;
    }
    int64_t p = 0;
    rp[0] = 0;
    for (int32_t z = 0; z < g; z++)
        for (int32_t y = 0; y < g; y++)
            for (int32_t x = 0; x < g; x++)
            {
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++)
                        {
                            int32_t xx = x + dx, yy = y + dy, zz = z + dz;
                            if (xx < 0 || yy < 0 || zz < 0 || xx >= g || yy >= g || zz >= g)
                                continue;
                            ci[p++] = xx + g * (yy + g * zz);
                        }
                rp[x + g * (y + g * z) + 1] = (int32_t)p;
            }
    float *val = (float *)malloc((size_t)(p ? p : 1) * sizeof(float));
    if (!val)
    {
        free(rp);
        free(ci);
        return -1;
    }
    rvsp_gen_fill_values(val, p, seed, dist);
    *rp_out = rp;
    *ci_out = ci;
    *val_out = val;
    return p;
}

static int64_t zipf_total(double cc, double s, int32_t m, int32_t maxlen)
{
    int64_t t = 0;
    for (int32_t r = 1; r <= m; r++)
    {
        double v = floor(cc * pow((double)r, -s) + 0.5);
        if (v < 1)
            v = 1;
        if (v > maxlen)
            v = maxlen;
        t += (int64_t)v;
    }
    return t;
}

/* m x n rows with rank-based power-law lengths whose total is exactly
 * round(mean*m), rows shuffled, distinct uniform columns. */
int64_t rvsp_gen_zipf_rows(int32_t m, int32_t n, double s, double mean,
                           int32_t maxlen, uint64_t seed, int dist,
                           int32_t **rp_out, int32_t **ci_out,
                           float **val_out)
{
    rng_t r;
    rng_seed(&r, seed);
    if (maxlen > n)
        maxlen = n;
    int64_t target = (int64_t)llround(mean * (double)m);
    double lo = 0.0, hi = (double)maxlen * pow((double)m, s) + 1.0;
    for (int it = 0; it < 200; it++)
    {
        double mid = 0.5 * (lo + hi);
        if (zipf_total(mid, s, m, maxlen) < target)
            lo = mid;
        else
            hi = mid;
    }
    int32_t *len = (int32_t *)malloc((size_t)m * sizeof(int32_t));
    if (!len)
        return -1;
    int64_t tot = 0;
    for (int32_t q = 0; q < m; q++)
    {
        double v = floor(hi * pow((double)(q + 1), -s) + 0.5);
        if (v < 1)
            v = 1;
        if (v > maxlen)
            v = maxlen;
        len[q] = (int32_t)v;
        tot += len[q];
    }
    /* exact total: take the surplus off the longest rows above 1, or add the
     * deficit to the longest rows below maxlen, one at a time in rank order */
    for (int32_t q = 0; tot > target && q < m; q++)
        if (len[q] > 1)
        {
            len[q]--;
            tot--;
        }
    for (int32_t q = 0; tot < target && q < m; q++)
        if (len[q] < maxlen)
        {
            len[q]++;
            tot++;
        }
    for (int32_t q = m - 1; q > 0; q--)
    {
        int32_t j = (int32_t)rng_below(&r, (uint64_t)q + 1);
        int32_t t = len[q];
        len[q] = len[j];
        len[j] = t;
    }
    int32_t *rp = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc((size_t)(tot ? tot : 1) * sizeof(int32_t));
    float *val = (float *)malloc((size_t)(tot ? tot : 1) * sizeof(float));
    uint64_t *mark = (uint64_t *)calloc(((size_t)n + 63) / 64, 8);
    if (!rp || !ci || !val || !mark)
    {
        free(len);
        free(rp);
        free(ci);
        free(val);
        free(mark);
        return -1;
    }
    rp[0] = 0;
    for (int32_t q = 0; q < m; q++)
        rp[q + 1] = rp[q] + len[q];
    free(len);
    for (int32_t q = 0; q < m; q++)
        sample_distinct(&r, 0, n, rp[q + 1] - rp[q], ci + rp[q], mark);
    for (int64_t p = 0; p < tot; p++)
        val[p] = dist ? rng_normal(&r) : rng_sym(&r);
    free(mark);
    *rp_out = rp;
    *ci_out = ci;
    *val_out = val;
    return tot;
}

int64_t rvsp_gen_banded(int32_t m, int32_t n, int32_t per_row, int32_t stride,
                        int32_t half, uint64_t seed, int dist,
                        int32_t **rp_out, int32_t **ci_out, float **val_out)
{
    rng_t r;
    rng_seed(&r, seed);
    int32_t *rp = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    int32_t *ci = (int32_t *)malloc((size_t)m * per_row * sizeof(int32_t));
    uint64_t *mark = (uint64_t *)calloc(((size_t)2 * half + 63) / 64 + 1, 8);
    if (!rp || !ci || !mark)
    {
        free(rp);
        free(ci);
        free(mark);
        return -1;
    }
    int64_t p = 0;
    rp[0] = 0;
    for (int32_t k = 0; k < m; k++)
    {
        int64_t lo = (int64_t)stride * k - half, hi = (int64_t)stride * k + half;
        if (lo < 0)
            lo = 0;
        if (hi > n)
            hi = n;
        int32_t span = (int32_t)(hi - lo);
        int32_t want = per_row < span ? per_row : span;
        if (want > 0)
            sample_distinct(&r, (int32_t)lo, span, want, ci + p, mark);
        p += want;
        rp[k + 1] = (int32_t)p;
    }
    free(mark);
    float *val = (float *)malloc((size_t)(p ? p : 1) * sizeof(float));
    if (!val)
    {
        free(rp);
        free(ci);
        return -1;
    }
    for (int64_t q = 0; q < p; q++)
        val[q] = dist ? rng_normal(&r) : rng_sym(&r);
    *rp_out = rp;
    *ci_out = ci;
    *val_out = val;
    return p;
}

/* Sizes floor(exp(U[ln lo, ln hi])) for a batch of `count` problems. */
void rvsp_gen_loguniform_sizes(int32_t count, int32_t lo, int32_t hi,
                               uint64_t seed, int32_t *sizes)
{
    rng_t r;
    rng_seed(&r, seed);
    double a = log((double)lo), b = log((double)hi);
    for (int32_t i = 0; i < count; i++)
    {
        int32_t v = (int32_t)floor(exp(a + (b - a) * rng_unit(&r)));
        sizes[i] = v < lo ? lo : (v > hi ? hi : v);
    }
}
