/*
 * rvsp_mtx.h -- Matrix Market files to canonical FP32 CSR (host memory).
 *
 * Replaces rv-sparse's tools/src/mtx_to_csr_formatter.c:147
 * (assemble_csr_matrix) together with bench/bench.c:453 (build_from_mtx),
 * the path rv-sparse's bench and bench/csr_check.c use to read real
 * matrices. Same reading of the file:
 *   - the banner's field "pattern" gives every entry the value 1.0f, and its
 *     symmetry "symmetric" mirrors every off-diagonal entry (any other
 *     symmetry is read as general);
 *   - comment lines after the banner are skipped, then "rows cols entries";
 *   - entries are 1-based "row col [value]", values parsed as float; reading
 *     stops at the declared count or at the first entry that does not parse;
 *   - each row is sorted by column (stable: equal columns keep file order,
 *     stored entries before their mirrors) and repeated columns are summed
 *     in that order in float arithmetic;
 *   - rows = the declared row count, cols = the largest column index + 1
 *     (bench.c:465-469 infers it the same way).
 * Unlike the reference tool, an entry outside the declared rows or a column
 * below 1 is an error (RVSP_ERROR_INVALID_ARGUMENT) instead of a write out
 * of bounds. No counterpart in rv_sparse.h.
 */
#ifndef RVSP_MTX_H
#define RVSP_MTX_H

#include "rv_sparse_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reads `path` into *out: host arrays allocated by the library
 * (owns_data = 1, release with rvsp_csr_destroy), FP32, canonical CSR.
 * RVSP_ERROR_NULL_POINTER for a null argument; RVSP_ERROR_INVALID_ARGUMENT
 * when the file cannot be opened, has no dimensions line or holds an entry
 * out of range; RVSP_ERROR_ALLOCATION_FAILED when memory runs out. */
rvsp_status_t rvsp_mtx_load(const char *path, rvsp_csr_matrix_t *out);

#ifdef __cplusplus
}
#endif

#endif /* RVSP_MTX_H */
