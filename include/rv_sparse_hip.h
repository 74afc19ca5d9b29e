/*
 * rv_sparse_hip.h -- MI355X extensions to the rv_sparse C API.
 *
 * Nothing in this header has a counterpart in rv-sparse: it adds what a GPU
 * caller needs around the unchanged rv_sparse.h surface.
 *   - device-resident operands for the descriptor and one-shot paths, so
 *     callers that keep matrices in HBM skip the PCIe copies the host API
 *     implies;
 *   - the HIP stream a descriptor launches on, and hipGraph replay of the
 *     numeric pass for descriptor reuse;
 *   - a batched entry point that runs many independent products in one
 *     persistent launch;
 *   - small device-memory and timing helpers, so callers without a HIP
 *     toolchain (C through dlopen, Python through ctypes) can stage data.
 * All functions are extern "C", take plain pointers and sizes, and return
 * rvsp_status_t. A missing or unusable GPU is reported as
 * RVSP_ERROR_UNSUPPORTED_BACKEND; there is no CPU fallback.
 */
#ifndef RV_SPARSE_HIP_H
#define RV_SPARSE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device and memory helpers ---------------------------------------- */

/* RVSP_SUCCESS when a HIP device is usable, RVSP_ERROR_UNSUPPORTED_BACKEND
 * otherwise. *count (optional) receives the number of visible devices. */
rvsp_status_t rvsp_hip_available(int *count);

rvsp_status_t rvsp_hip_malloc(void **ptr, size_t bytes);
rvsp_status_t rvsp_hip_free(void *ptr);
rvsp_status_t rvsp_hip_memcpy_h2d(void *dst, const void *src, size_t bytes);
rvsp_status_t rvsp_hip_memcpy_d2h(void *dst, const void *src, size_t bytes);
rvsp_status_t rvsp_hip_memcpy_d2d(void *dst, const void *src, size_t bytes);
rvsp_status_t rvsp_hip_memset(void *dst, int value, size_t bytes);
/* Waits for all work on the device. */
rvsp_status_t rvsp_hip_synchronize(void);

/* Events on a given stream (NULL = the descriptor-less default stream). */
typedef struct rvsp_hip_event *rvsp_hip_event_t;
rvsp_status_t rvsp_hip_event_create(rvsp_hip_event_t *ev);
rvsp_status_t rvsp_hip_event_record(rvsp_hip_event_t ev, void *stream);
rvsp_status_t rvsp_hip_event_synchronize(rvsp_hip_event_t ev);
rvsp_status_t rvsp_hip_event_elapsed_ms(rvsp_hip_event_t start,
                                        rvsp_hip_event_t stop,
                                        float *ms);
void rvsp_hip_event_destroy(rvsp_hip_event_t ev);

/* Phase timing (process-wide). When on, each analysis, structure fill and
 * numeric pass records a pair of HIP events on the stream it runs on.
 * rvsp_hip_get_timing waits for them and returns, per phase
 * (0 analysis, 1 fill, 2 numeric, 3 unused), the summed milliseconds and
 * the number of calls. */
void rvsp_hip_set_timing(int on);
rvsp_status_t rvsp_hip_get_timing(double *ms, long long *calls, int n);
void rvsp_hip_reset_timing(void);

/* ---- descriptor extensions --------------------------------------------- */

typedef enum
{
    RVSP_MEMORY_HOST = 0,  /* operands and C hold host pointers (default) */
    RVSP_MEMORY_DEVICE = 1 /* operands and C hold device pointers         */
} rvsp_memory_t;

/* Where A, B and C live for this descriptor. Call before
 * work_estimation(); INVALID_ARGUMENT afterwards. With RVSP_MEMORY_DEVICE the
 * pointers in A, B (and C at compute) must be device allocations that stay
 * valid until the descriptor is destroyed or re-estimated, and compute()
 * returns once the work is queued on the descriptor's stream (no host
 * synchronisation); C->row_ptr then receives a device-to-device copy.
 * work_estimation checks device operands for canonical CSR on the device
 * (RVSP_ERROR_INVALID_CSR) before anything reads through their indices;
 * compute must be given the same row_ptr and col_idx arrays the estimation
 * saw (RVSP_ERROR_INVALID_ARGUMENT otherwise); values may change freely. */
rvsp_status_t rvsp_spgemm_set_memory(rvsp_spgemm_descr_t descr,
                                     rvsp_memory_t memory);

/* Use an existing HIP stream (hipStream_t passed as void *). The library
 * never destroys a stream it did not create. Call before work_estimation. */
rvsp_status_t rvsp_spgemm_set_stream(rvsp_spgemm_descr_t descr, void *stream);

/* The stream the descriptor launches on (created at descr_create unless set). */
rvsp_status_t rvsp_spgemm_get_stream(rvsp_spgemm_descr_t descr, void **stream);

/* Capture the numeric pass in a hipGraph on the first compute() that does
 * not refill C's structure, and replay it on later calls whose pointers
 * match. Default off. Only meaningful with RVSP_MEMORY_DEVICE. */
rvsp_status_t rvsp_spgemm_set_graph(rvsp_spgemm_descr_t descr, int enable);

/* Waits for the descriptor's queued work and reports a kernel-side
 * consistency error (RVSP_ERROR_INVALID_CSR: the operands were not the
 * canonical CSR the analysis saw). Needed after device-memory compute(). */
rvsp_status_t rvsp_spgemm_wait(rvsp_spgemm_descr_t descr);

/* Total intermediate products (sum of op counts) of the last estimation. */
rvsp_status_t rvsp_spgemm_get_products(rvsp_spgemm_descr_t descr,
                                       int64_t *products);

/* ---- device one-shot ---------------------------------------------------- */

/*
 * C = A * B with A, B device-resident. C receives device arrays allocated by
 * the library (owns_data = RVSP_OWNS_DEVICE); release them with
 * rvsp_csr_destroy. Operands that are not canonical CSR are found on the
 * device and reported as RVSP_ERROR_INVALID_CSR before any product kernel
 * runs. options as for rvsp_spgemm_csr, plus
 * algo selection through rvsp_spgemm_csr_device_algo. Synchronous.
 */
rvsp_status_t rvsp_spgemm_csr_device(const rvsp_csr_matrix_t *A,
                                     const rvsp_csr_matrix_t *B,
                                     rvsp_csr_matrix_t *C,
                                     const rvsp_spgemm_options_t *options);

rvsp_status_t rvsp_spgemm_csr_device_algo(const rvsp_csr_matrix_t *A,
                                          const rvsp_csr_matrix_t *B,
                                          rvsp_csr_matrix_t *C,
                                          rvsp_spgemm_algo_t algo);

/* ---- autotuned launch parameters ------------------------------------------ */

/*
 * No counterpart in rv-sparse (whose tunables are compile-time macros).
 * Parameters that decide how a product is binned and scheduled, never what
 * it computes: every setting gives bit-identical C.
 *   big_window_outputs/chunks  big-row numeric window: one of 2048x16,
 *                              4096x32, 2048x32, 1024x16, 1024x8
 *   sym_big_threads            big-row count block: 256, 512 or 1024
 *   seg_products               products per big-row segment (one wave),
 *                              1024 .. 2^22
 *   hash_max_u                 rows whose min(products, column span) is at
 *                              most this use the LDS hash count, 64 .. 2048
 *   fan_min_products/_fill     products (structure fill: outputs) above
 *                              which a pass runs its row bins on side streams
 * One entry per workload class (rvsp_tuning_class: estimated products
 * nnz(A) * nnz(B) / k in four size buckets, times whether B has more than
 * 16384 columns), written by the autotuner (python -m rvsp.autotune) to
 * tuning/table.txt beside the library, or RVSP_TUNING_TABLE, and loaded on
 * first use. Environment overrides (RVSP_BIG_WINDOW, RVSP_SYM_BIG_THREADS)
 * still take precedence.
 */
#define RVSP_TUNING_CLASSES 8

typedef struct
{
    int32_t big_window_outputs;
    int32_t big_window_chunks;
    int32_t sym_big_threads;
    int32_t seg_products;
    int32_t hash_max_u;
    int64_t fan_min_products;
    int64_t fan_min_fill;
} rvsp_tuning_t;

/* Class of a product C = A * B (A m x k, B k x n) from its shapes alone. */
int32_t rvsp_tuning_class(int64_t nnz_a, int64_t nnz_b, int32_t k, int32_t n);
/* The built-in defaults (the values before any tuning). */
rvsp_status_t rvsp_tuning_defaults(rvsp_tuning_t *out);
rvsp_status_t rvsp_tuning_get(int32_t cls, rvsp_tuning_t *out);
/* Replaces one class's entry (INVALID_ARGUMENT for values outside the
 * ranges above). */
rvsp_status_t rvsp_tuning_set(int32_t cls, const rvsp_tuning_t *t);
/* Uses t for every class until called with NULL (the autotuner's trials). */
rvsp_status_t rvsp_tuning_force(const rvsp_tuning_t *t);
/* Loads a table file (format: '#' comments, then per class the line
 * "class outputs chunks sym_big_threads seg_products hash_max_u
 * fan_min_products fan_min_fill"); classes it omits get the defaults. */
rvsp_status_t rvsp_tuning_load_file(const char *path);

/* ---- allocation-free device one-shot --------------------------------------- */

/*
 * No counterpart in rv-sparse. For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): C = A * B with A,
 * B device-resident FP32 CSR, as rvsp_spgemm_csr_device with
 * RVSP_SPGEMM_ALGO_HIP (heuristic row bins, bit-identical results), but
 * without any device or host allocation. Every device array the product
 * needs besides its inputs and outputs comes from one fixed device buffer
 * of RVSP_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends.
 *   C: the caller's device arrays, row_ptr with A->rows + 1 entries and
 *      col_idx / values with room for C->nnz entries (C->nnz = capacity on
 *      input). On success C->nnz = nnz(C), rows, cols, dtype and format are
 *      set and owns_data = 0. When nnz(C) exceeds the capacity:
 *      RVSP_ERROR_INVALID_ARGUMENT with C->nnz = the entries needed.
 *   stream: a hipStream_t as void * (NULL: the default stream); the call
 *      returns when the product is complete.
 * RVSP_ERROR_SCRATCH_EXCEEDED when the product needs more than the fixed
 * buffer (plan arrays grow with A's rows, B's entries and the column spans
 * of long rows); RVSP_ERROR_INVALID_CSR for non-canonical operands.
 * Calls on one GPU share its buffer and each overwrites what the previous
 * one left: they are serialised within a process; work on the same device
 * that the caller overlaps with a call must not touch the outputs.
 */
#define RVSP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

rvsp_status_t rvsp_spgemm_csr_device_noalloc(const rvsp_csr_matrix_t *A,
                                             const rvsp_csr_matrix_t *B,
                                             rvsp_csr_matrix_t *C, void *stream);

/* ---- batched products ---------------------------------------------------- */

typedef struct rvsp_spgemm_batch *rvsp_spgemm_batch_t;

/*
 * C[i] = A[i] * B[i] for i < count, all device-resident FP32 CSR, computed
 * as one stacked product per chunk (chunks keep each stacked C within
 * int32 indexing), so every kernel launch covers many problems. Results are
 * bit-identical to separate rvsp_spgemm_csr_device calls. Every problem's
 * operands are checked for canonical CSR on the device first
 * (RVSP_ERROR_INVALID_CSR). C[i] receives
 * device pointers into storage owned by *batch (owns_data = 0); release it
 * with rvsp_spgemm_batch_destroy. Synchronous.
 */
rvsp_status_t rvsp_spgemm_batched_device(int32_t count, const rvsp_csr_matrix_t *A,
                                         const rvsp_csr_matrix_t *B, rvsp_csr_matrix_t *C,
                                         rvsp_spgemm_algo_t algo, rvsp_spgemm_batch_t *batch);

/* Number of stacked products (chunks) the batch was computed in. */
int32_t rvsp_spgemm_batch_chunks(rvsp_spgemm_batch_t batch);

void rvsp_spgemm_batch_destroy(rvsp_spgemm_batch_t batch);

/* ---- learned row selectors ------------------------------------------------ */

/*
 * RVSP_SPGEMM_ALGO_HIP_LEARNED, and RVSP_SPGEMM_ALGO_DEFAULT, use two small
 * MLPs (features -> 16 -> 16 -> outputs, ReLU) in place of fixed binning
 * rules where their parameters are loaded; RVSP_SPGEMM_ALGO_HIP always uses
 * the rules:
 *   SIZER   4 features -> 1: predicted log2(nnz(C_i) / min(ops_i, span_i)),
 *           sizes each row's symbolic LDS hash table (times the margin);
 *           rows whose exact count overflows are recounted in their
 *           heuristic bin, so results never depend on the prediction;
 *   CHOOSER 5 features -> 8: predicted log2(ns per row) of each numeric
 *           accumulator (LDS hash table, or the rank or dense window the
 *           rules would use, 8..64 lanes); the smallest feasible prediction
 *           wins unless it beats the rules' own pick by less than the
 *           margin.
 * Feature vectors are 5 floats per row: log2(1+alen)/16, log2(1+ops)/32,
 * log2(1+span)/32, log2(1+ops/alen)/16, log2(1+nnz)/32 (the sizer reads
 * the first four). Parameters are stored layer by layer: W1[16][nf], b1,
 * W2[16][16], b2, W3[no][16], b3.
 * Without loaded parameters both behave as RVSP_SPGEMM_ALGO_HIP. Results do
 * not depend on the choice: they are the same bit for bit.
 * Parameters load on first use from the file named by RVSP_SELECTOR_SIZER /
 * RVSP_SELECTOR_CHOOSER, else from selector/{sizer,chooser}.txt beside the
 * library's directory (text: the margin, then the parameters).
 */
#define RVSP_SELECTOR_SIZER 0
#define RVSP_SELECTOR_CHOOSER 1
#define RVSP_SIZER_PARAMS 369
#define RVSP_CHOOSER_PARAMS 504

/* Number of parameters of a selector (-1 for an unknown one). */
int32_t rvsp_selector_param_count(int32_t which);
/* Installs parameters (host array of param_count floats) and margin > 0. */
rvsp_status_t rvsp_selector_set(int32_t which, const float *params, int32_t n, float margin);
rvsp_status_t rvsp_selector_get(int32_t which, float *params, int32_t n, float *margin);
/* Reads a parameter file (format above) and installs it. */
rvsp_status_t rvsp_selector_load_file(int32_t which, const char *path);
/* Turns a selector off (the learned algo then uses the heuristic for it). */
rvsp_status_t rvsp_selector_clear(int32_t which);

/* Network outputs for n feature rows (device, n x 5) at the given host
 * parameters: pred (device) receives n x outputs floats. Synchronous. */
rvsp_status_t rvsp_selector_predict(int32_t which, const float *feats, int32_t n,
                                    const float *params_host, float *pred);
/* Gradient at params_host of L = (1/n) * sum over samples and outputs of
 * (pred - target)^2, NaN targets skipped; feats/target are device arrays
 * (n x 5, n x outputs). grad_host receives param_count floats, *loss L. */
rvsp_status_t rvsp_selector_grad(int32_t which, const float *feats, const float *target,
                                 int32_t n, const float *params_host, float *grad_host,
                                 float *loss);
/* Adam (beta 0.9/0.999) on L over minibatches of `batch` samples, reshuffled
 * each epoch from `seed`; params_host is the start point and receives the
 * result; *loss_out the last epoch's loss per sample. Runs on the GPU. */
rvsp_status_t rvsp_selector_train(int32_t which, const float *feats, const float *target,
                                  int32_t n, int32_t epochs, int32_t batch, float lr,
                                  uint64_t seed, float *params_host, float *loss_out);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_HIP_H */
