/*
 * genmat.h -- synthetic sparse matrices straight into CSR, the generator
 * behind rv-sparse's `bench --gen R C DENSITY SEED` inputs. Drop-in for
 * rv-sparse's tools/include/genmat.h:18-53: the same two structs, field for
 * field, and the same four functions, so bench.c and any other caller
 * compile unchanged against this header. The generated matrices are the
 * reference generator's, bit for bit, for every non-symmetric request
 * (rows or columns as lines, banded or not, square or rectangular).
 *
 *   genmat_params_t p = genmat_default_params(rows, cols);
 *   p.density = 0.01;
 *   csr_matrix_t m = genmat_generate_csr(p);
 *   if (m.nrows < 0) { error }
 *   ... m.row_ptr / m.col_idx / m.values ...
 *   genmat_free_csr(&m);
 */
#ifndef RVSP_GENMAT_H
#define RVSP_GENMAT_H

#include <stdint.h>

#include "rv_sparse_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* tools/include/genmat.h:18-25 */
typedef struct {
    int nrows;
    int ncols;
    long long nnz;
    int *row_ptr;   /* nrows + 1 */
    int *col_idx;   /* nnz, ascending within each row */
    double *values; /* nnz */
} csr_matrix_t;

/* tools/include/genmat.h:27-39 */
typedef struct {
    int row_cnt;
    int col_cnt;
    double density;    /* stored fraction of each line, default 0.001 */
    double cv;         /* coefficient of variation of line degrees, default 0.5 */
    int min;           /* fewest entries per line, default 1 */
    double imbalance;  /* (max - avg) / avg; -1.0 (default) caps at width - 1 */
    int is_symmetric;  /* 1: symmetric pattern and values (not built: error) */
    int is_column;     /* 1: degrees are drawn per column instead of per row */
    int low_bandwidth; /* default row_cnt - 1 (no band) */
    int up_bandwidth;  /* default row_cnt - 1 (no band) */
    int random_seed;   /* default 1 */
} genmat_params_t;

/* Defaults for the given shape. tools/src/genmat.c:626-640 */
genmat_params_t genmat_default_params(int row_cnt, int col_cnt);

/* The matrix, or one with nrows == -1 and NULL arrays on error: a symmetric
 * request (not built), a symmetric-but-not-square request (as the reference),
 * negative sizes or bandwidths, a band so narrow that a line must draw from
 * an empty window (where the reference divides by zero), or no memory.
 * Reentrant: the value stream runs on a private state and leaves the
 * caller's rand() alone. tools/src/genmat.c:854-981 */
csr_matrix_t genmat_generate_csr(genmat_params_t params);

/* Matrix Market "coordinate real general", values as %.16e; 0 or -1.
 * tools/src/genmat.c:996-1013 */
int genmat_save_csr_to_mtx(const csr_matrix_t *csr, const char *filename);

/* Frees the arrays and zeroes the struct. tools/src/genmat.c:983-994 */
void genmat_free_csr(csr_matrix_t *m);

/* rv-sparse's bench inputs for `--gen R C DENSITY SEED` (bench/bench.c:
 * 677-707, conversions :391-450): A = R x C with `seed`, B = C x R with
 * `seed + 1`, both cv 0.5 and min 1, values converted to `dtype` (FP32:
 * rounded from the double; FP64: kept; INT8: (int)v % 127, 0 -> 1).
 * A and B own host arrays (owns_data = 1; release with rvsp_csr_destroy).
 * No counterpart in the reference library: it restates the bench's own
 * setup so every caller builds the same pair. */
rvsp_status_t rvsp_genmat_bench_pair(int32_t rows, int32_t cols, double density, int32_t seed,
                                     rvsp_dtype_t dtype, rvsp_csr_matrix_t *A,
                                     rvsp_csr_matrix_t *B);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_GENMAT_H */
