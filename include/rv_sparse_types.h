/*
 * rv_sparse_types.h -- value types of the rv_sparse C API, MI355X build.
 *
 * Binary layout and enumerator values are identical to the rv-sparse public
 * types (include/rv_sparse_types.h:21-96 in merledu/rv-sparse), so a caller
 * compiled against that header links against this library unchanged. New
 * enumerators are only ever appended, never renumbered.
 */
#ifndef RV_SPARSE_TYPES_H
#define RV_SPARSE_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Zero is success; every error is negative.
 * Same values as rv-sparse include/rv_sparse_types.h:21-30; the last one is
 * appended by this build (no counterpart in rv-sparse). */
typedef enum
{
    RVSP_SUCCESS = 0,
    RVSP_ERROR_NULL_POINTER = -1,
    RVSP_ERROR_INVALID_ARGUMENT = -2,
    RVSP_ERROR_INVALID_CSR = -3,
    RVSP_ERROR_UNSUPPORTED_DTYPE = -4,
    RVSP_ERROR_UNSUPPORTED_BACKEND = -5,
    RVSP_ERROR_ALLOCATION_FAILED = -6,
    /* the product needs more device scratch than the allocation-free entry
       point's fixed buffer holds (rv_sparse_hip.h) */
    RVSP_ERROR_SCRATCH_EXCEEDED = -7
} rvsp_status_t;

/* Element types (rv_sparse_types.h:32-39). Only FP32 runs on the GPU path. */
typedef enum
{
    RVSP_DTYPE_INT8,
    RVSP_DTYPE_INT32,
    RVSP_DTYPE_BF16,
    RVSP_DTYPE_FP32,
    RVSP_DTYPE_FP64
} rvsp_dtype_t;

/* Storage formats (rv_sparse_types.h:41-44). */
typedef enum
{
    RVSP_FORMAT_CSR
} rvsp_format_t;

/* Execution backends (rv_sparse_types.h:46-51) plus the appended HIP backend.
 *
 * In this build SCALAR and SCALAR_UNROLL4 are served by the HIP kernels, which
 * reproduce the scalar kernel's arithmetic (ascending-k, separately rounded
 * multiply and add) bit for bit. RVV_INTRINSICS is reported unsupported, as in
 * any non-RISC-V build of rv-sparse. */
typedef enum
{
    RVSP_BACKEND_SCALAR,
    RVSP_BACKEND_SCALAR_UNROLL4,
    RVSP_BACKEND_RVV_INTRINSICS,
    RVSP_BACKEND_HIP /* appended: MI355X (gfx950) kernels */
} rvsp_backend_t;

/* CSR matrix, field for field the rv-sparse struct (rv_sparse_types.h:53-67).
 *
 * owns_data: 0 = arrays borrowed from the caller, 1 = host arrays allocated by
 * the library (free() on destroy), RVSP_OWNS_DEVICE = device arrays allocated
 * by the library (released on the GPU on destroy; device one-shot only). */
typedef struct
{
    int32_t rows;
    int32_t cols;
    int32_t nnz;

    int32_t *row_ptr;
    int32_t *col_idx;
    void *values;

    rvsp_dtype_t dtype;
    rvsp_format_t format;

    int owns_data;
} rvsp_csr_matrix_t;

#define RVSP_OWNS_DEVICE 2

/* One-shot options (rv_sparse_types.h:69-74). */
typedef struct
{
    rvsp_backend_t backend;
    rvsp_dtype_t input_dtype;
    rvsp_dtype_t output_dtype;
} rvsp_spgemm_options_t;

/* Opaque two-phase descriptor (rv_sparse_types.h:82). */
typedef struct rvsp_spgemm_descr *rvsp_spgemm_descr_t;

/* Accumulation strategies (rv_sparse_types.h:90-96) plus appended HIP ones.
 *
 * DEFAULT           the build's default: here the HIP heuristic-binned path.
 * RVV, CONTIG,      RISC-V Vector strategies; unsupported in this build
 * ADAPTIVE          (set_algo returns RVSP_ERROR_UNSUPPORTED_BACKEND).
 * HIP               HIP kernels, rows binned by exact symbolic counts with
 *                   fixed heuristic thresholds.
 * HIP_LEARNED       HIP kernels, each row's accumulator chosen by the small
 *                   MLP selector; the exact symbolic counts still size every
 *                   table, so a misprediction costs time, never correctness. */
typedef enum
{
    RVSP_SPGEMM_ALGO_DEFAULT = 0,
    RVSP_SPGEMM_ALGO_RVV,
    RVSP_SPGEMM_ALGO_CONTIG,
    RVSP_SPGEMM_ALGO_ADAPTIVE,
    RVSP_SPGEMM_ALGO_HIP,        /* appended */
    RVSP_SPGEMM_ALGO_HIP_LEARNED /* appended */
} rvsp_spgemm_algo_t;

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_TYPES_H */
