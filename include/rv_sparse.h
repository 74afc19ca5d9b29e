/*
 * rv_sparse.h -- public C API of the MI355X build of rv_sparse.
 *
 * Every entry point here has the signature and the status behaviour of the
 * rv-sparse function it replaces (include/rv_sparse.h:27-142 in
 * merledu/rv-sparse); the citation on each declaration names that function.
 * Arrays passed through this header are HOST memory, exactly as in rv-sparse.
 * Device-resident variants, streams, graphs and batching live in
 * rv_sparse_hip.h.
 *
 * Threading: distinct descriptors are independent; one descriptor must not be
 * used from two threads at once (compute() updates its structure cache).
 */
#ifndef RV_SPARSE_H
#define RV_SPARSE_H

#include <stddef.h>

#include "rv_sparse_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RVSP_VERSION_MAJOR 0
#define RVSP_VERSION_MINOR 1
#define RVSP_VERSION_PATCH 0

/* Replaces rvsp_get_version (rv_sparse.h:27, src/core/error.c:17-20). */
const char *rvsp_get_version(void);

/* Replaces rvsp_status_to_string (rv_sparse.h:29, src/core/error.c:22-43). */
const char *rvsp_status_to_string(rvsp_status_t status);

/*
 * Replaces rvsp_csr_create (rv_sparse.h:38-46, src/core/matrix.c:18-49).
 * Borrows the three arrays (owns_data = 0) and runs rvsp_csr_validate.
 */
rvsp_status_t rvsp_csr_create(rvsp_csr_matrix_t *A,
                              int32_t rows,
                              int32_t cols,
                              int32_t nnz,
                              int32_t *row_ptr,
                              int32_t *col_idx,
                              void *values,
                              rvsp_dtype_t dtype);

/*
 * Replaces rvsp_csr_validate (rv_sparse.h:48, src/core/matrix.c:51-90).
 * Checks non-NULL arrays, positive dims, row_ptr[0] == 0,
 * row_ptr[rows] == nnz, monotone row_ptr and in-range columns. It does not
 * check sortedness or duplicates (rvsp_csr_check in rvsp_spgemm.h does).
 */
rvsp_status_t rvsp_csr_validate(const rvsp_csr_matrix_t *A);

/*
 * Replaces rvsp_csr_destroy (rv_sparse.h:50, src/core/matrix.c:92-115).
 * Frees the arrays when the library owns them (host or device), then resets
 * every field. NULL is accepted.
 */
void rvsp_csr_destroy(rvsp_csr_matrix_t *A);

/*
 * Replaces rvsp_spgemm_csr (rv_sparse.h:63-67, src/core/spgemm.c:135-211).
 * One-shot C = A * B with library-allocated host arrays for C
 * (owns_data = 1). options == NULL means SCALAR / FP32 / FP32. FP32 with
 * backend SCALAR, SCALAR_UNROLL4 or HIP runs on the GPU; inputs are checked
 * for canonical CSR first (RVSP_ERROR_INVALID_CSR otherwise). On failure C's
 * arrays are released and C is reset to empty.
 */
rvsp_status_t rvsp_spgemm_csr(const rvsp_csr_matrix_t *A,
                              const rvsp_csr_matrix_t *B,
                              rvsp_csr_matrix_t *C,
                              const rvsp_spgemm_options_t *options);

/* Replaces rvsp_spgemm_descr_create (rv_sparse.h:81,
 * src/core/spgemm_descr.c:66-88). */
rvsp_status_t rvsp_spgemm_descr_create(rvsp_spgemm_descr_t *descr);

/* Replaces rvsp_spgemm_descr_destroy (rv_sparse.h:88,
 * src/core/spgemm_descr.c:90-100). Releases host and device state; NULL ok. */
void rvsp_spgemm_descr_destroy(rvsp_spgemm_descr_t descr);

/* Replaces rvsp_spgemm_set_algo (rv_sparse.h:95-96,
 * src/core/spgemm_descr.c:102-125). INVALID_ARGUMENT after estimation;
 * UNSUPPORTED_BACKEND for strategies this build lacks (RVV, CONTIG,
 * ADAPTIVE). */
rvsp_status_t rvsp_spgemm_set_algo(rvsp_spgemm_descr_t descr,
                                   rvsp_spgemm_algo_t algo);

/*
 * Replaces rvsp_spgemm_work_estimation (rv_sparse.h:107-111,
 * src/core/spgemm_descr.c:162-255). Uploads A's and B's structure, runs the
 * symbolic pass on the GPU and keeps C's row pointer and per-row product
 * counts. Reports the exact nnz(C) and the workspace compute() needs.
 * nnz(C) above INT32_MAX returns RVSP_ERROR_ALLOCATION_FAILED, as the
 * reference's count does (src/kernels/spgemm/rvsp_symbolic.c:69-72).
 */
rvsp_status_t rvsp_spgemm_work_estimation(rvsp_spgemm_descr_t descr,
                                          const rvsp_csr_matrix_t *A,
                                          const rvsp_csr_matrix_t *B,
                                          size_t *workspace_bytes,
                                          int32_t *c_nnz_out);

/*
 * Replaces rvsp_spgemm_compute (rv_sparse.h:121-125,
 * src/core/spgemm_descr.c:257-355). Writes C's row pointer, column indices
 * (only when C->col_idx differs from the buffer filled last, or after
 * invalidate_structure) and values. Argument checks and their precedence
 * follow the reference (spgemm_descr.c:263-313).
 */
rvsp_status_t rvsp_spgemm_compute(rvsp_spgemm_descr_t descr,
                                  const rvsp_csr_matrix_t *A,
                                  const rvsp_csr_matrix_t *B,
                                  rvsp_csr_matrix_t *C,
                                  void *workspace);

/* Replaces rvsp_spgemm_get_op_counts (rv_sparse.h:133-135,
 * src/core/spgemm_descr.c:357-380). */
rvsp_status_t rvsp_spgemm_get_op_counts(rvsp_spgemm_descr_t descr,
                                        int64_t *op_counts_out,
                                        int32_t n);

/* Replaces rvsp_spgemm_invalidate_structure (rv_sparse.h:142,
 * src/core/spgemm_descr.c:129-135). NULL is accepted. */
void rvsp_spgemm_invalidate_structure(rvsp_spgemm_descr_t descr);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_H */
