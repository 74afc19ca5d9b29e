// rvsp_validate.hip -- canonical-CSR check of device-resident operands.
//
// The analysis kernels index B's row pointer by A's column indices and walk
// row ranges taken from row pointers, so an operand that is not canonical
// CSR could make them read outside its arrays. Host operands are checked on
// the host (rvsp_csr_check, the rule of rv-sparse's one-shot,
// src/core/spgemm.c:41-51); device operands are checked here, before any
// kernel that trusts them runs. Two streaming passes, each reading only
// indices that are in bounds whatever the input holds:
//   rows     one thread per row pointer entry: row_ptr[0] == 0, every row
//            pointer in [0, nnz], non-decreasing, row_ptr[rows] == nnz; marks
//            each row's first entry in a bitmap over the entries;
//   entries  one thread per entry: 0 <= col < cols, and a column is
//            strictly below the next one unless the next entry starts a row
//            (catches unsorted rows and repeated columns alike).
// A violation sets kErrOperand in the error word; the host reads it once
// both operands are checked and reports RVSP_ERROR_INVALID_CSR.
#include <hip/hip_runtime.h>

#include "rvsp_device.hpp"
#include "rvsp_kernels.hpp"

namespace rvsp {

__global__ __launch_bounds__(256) void k_csr_rows(int32_t rows, int32_t nnz,
                                                  const int32_t *__restrict__ rp,
                                                  uint32_t *__restrict__ starts,
                                                  int *__restrict__ err)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r > rows) return;
    const int32_t v = rp[r];
    bool bad = v < 0 || v > nnz;
    if (r == 0) bad |= v != 0;
    if (r == rows) bad |= v != nnz;
    if (r < rows) bad |= rp[r + 1] < v;
    if (bad) {
        atomicOr(err, kErrOperand);
        return;
    }
    // entry v opens row r (empty rows mark the same entry as the next row)
    if (r > 0 && r < rows && v < nnz) atomicOr(&starts[v >> 5], 1u << (v & 31));
}

__global__ __launch_bounds__(256) void k_csr_entries(int32_t cols, int32_t nnz,
                                                     const int32_t *__restrict__ ci,
                                                     const uint32_t *__restrict__ starts,
                                                     int *__restrict__ err)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= nnz) return;
    const int32_t c = ci[p];
    bool bad = c < 0 || c >= cols;
    const int64_t q = p + 1;
    if (q < nnz && !((starts[q >> 5] >> (q & 31)) & 1u)) bad |= ci[q] <= c;
    if (bad) atomicOr(err, kErrOperand);
}

hipError_t launch_csr_validate(hipStream_t st, int32_t rows, int32_t cols, int32_t nnz,
                               const int32_t *rp, const int32_t *ci, uint32_t *starts, int *err)
{
    const size_t words = ((size_t)nnz + 32) / 32;
    hipError_t e = hipMemsetAsync(starts, 0, words * 4, st);
    if (e != hipSuccess) return e;
    k_csr_rows<<<(unsigned)(((int64_t)rows + 1 + 255) / 256), 256, 0, st>>>(rows, nnz, rp, starts, err);
    if (nnz > 0)
        k_csr_entries<<<(unsigned)(((int64_t)nnz + 255) / 256), 256, 0, st>>>(cols, nnz, ci, starts, err);
    return hipGetLastError();
}

} // namespace rvsp
