// rvsp_batch.hip -- packing many independent CSR products into one.
//
// A batch C_i = A_i * B_i (i < count) is run as ONE product of stacked
// matrices: A_big stacks the A_i row blocks with their column indices
// shifted by the running B row offset, B_big stacks the B_i row blocks with
// their own (local) column indices. Row r of A_big then only reaches B rows
// of its own problem, so C_big's rows are exactly the C_i rows, with local
// column indices, computed by the same kernels in the same order as one
// product at a time: results are bit-identical to separate calls. Every
// analysis and numeric launch then covers all problems at once.
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "rvsp_hip_abi.h"

namespace {

constexpr int kPackTile = 4096;

struct PackJob {
    const int32_t *rp;
    const int32_t *ci;
    const float *va;
    int32_t rows;
    int32_t nnz;
    int64_t row_off;
    int64_t nnz_off;
    int32_t col_shift;
    int32_t cols;
    int32_t tile0;
};

__device__ __forceinline__ int find_job(const PackJob *jobs, int njobs, int tile)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One block per tile of kPackTile elements of one job. Writes row pointers
// (shifted by the job's nnz offset, including the closing entry, which the
// next job writes with the same value) and the entries (columns shifted).
// Every read stays inside the job's arrays (rows + 1 pointers, nnz entries)
// whatever they hold; a row pointer outside [0, nnz], a decreasing one, a
// first one other than 0 or a last one other than nnz, and a column outside
// [0, cols) set *err, so the stacked product is never analysed then.
__global__ __launch_bounds__(256) void k_pack(const PackJob *__restrict__ jobs, int njobs,
                                              int32_t *__restrict__ orp, int32_t *__restrict__ oci,
                                              float *__restrict__ ova, int *__restrict__ err)
{
    const int j = find_job(jobs, njobs, blockIdx.x);
    const PackJob jb = jobs[j];
    const int64_t t0 = (int64_t)(blockIdx.x - jb.tile0) * kPackTile;
    bool bad = false;
    for (int64_t e = t0 + threadIdx.x; e < t0 + kPackTile; e += 256) {
        if (e <= jb.rows) {
            const int32_t v = jb.rp[e];
            bad |= v < 0 || v > jb.nnz || (e == 0 && v != 0) || (e == jb.rows && v != jb.nnz) ||
                   (e < jb.rows && jb.rp[e + 1] < v);
            orp[jb.row_off + e] = (int32_t)(v + jb.nnz_off);
        }
        if (e < jb.nnz) {
            const int32_t c = jb.ci[e];
            bad |= c < 0 || c >= jb.cols;
            oci[jb.nnz_off + e] = c + jb.col_shift;
            ova[jb.nnz_off + e] = jb.va[e];
        }
    }
    if (bad) atomicOr(err, 1);
}

struct UnpackJob {
    int64_t row_off;   // first row of the problem in the packed product
    int64_t out_off;   // first entry of its local row pointer
    int32_t rows;
    int32_t tile0;
};

__global__ __launch_bounds__(256) void k_unpack_rp(const UnpackJob *__restrict__ jobs, int njobs,
                                                   const int32_t *__restrict__ big_rp,
                                                   int32_t *__restrict__ out_rp,
                                                   int64_t *__restrict__ base_out)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= (int)blockIdx.x)
            lo = mid;
        else
            hi = mid - 1;
    }
    const UnpackJob jb = jobs[lo];
    const int64_t t0 = (int64_t)(blockIdx.x - jb.tile0) * kPackTile;
    const int32_t base = big_rp[jb.row_off];
    if (t0 == 0 && threadIdx.x == 0) base_out[lo] = base;
    for (int64_t e = t0 + threadIdx.x; e < t0 + kPackTile && e <= jb.rows; e += 256)
        out_rp[jb.out_off + e] = big_rp[jb.row_off + e] - base;
}

} // namespace

extern "C" {

rvsp_status_t rvsp_hip_pack(void *stream, int32_t count, const rvsp_hip_pack_item_t *items,
                            int32_t *out_rp, int32_t *out_ci, float *out_va)
{
    if (count <= 0) return RVSP_SUCCESS;
    if (!items || !out_rp || !out_ci || !out_va) return RVSP_ERROR_NULL_POINTER;
    hipStream_t st = (hipStream_t)stream;
    std::vector<PackJob> jobs((size_t)count);
    int tiles = 0;
    for (int i = 0; i < count; ++i) {
        const rvsp_hip_pack_item_t &it = items[i];
        const int64_t work = (int64_t)(it.rows + 1 > it.nnz ? it.rows + 1 : it.nnz);
        jobs[i] = PackJob{it.row_ptr, it.col_idx, it.values, it.rows, it.nnz,
                          it.row_off, it.nnz_off, it.col_shift, it.cols, tiles};
        tiles += (int)((work + kPackTile - 1) / kPackTile);
    }
    // the job table, then the error word
    const size_t table = (jobs.size() * sizeof(PackJob) + 255) & ~(size_t)255;
    char *d = nullptr;
    if (hipMallocAsync((void **)&d, table + 256, st) != hipSuccess)
        return RVSP_ERROR_ALLOCATION_FAILED;
    int *derr = (int *)(d + table);
    if (hipMemcpyAsync(d, jobs.data(), jobs.size() * sizeof(PackJob), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemsetAsync(derr, 0, 4, st) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        (void)hipFreeAsync(d, st);
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    k_pack<<<tiles, 256, 0, st>>>((const PackJob *)d, count, out_rp, out_ci, out_va, derr);
    hipError_t e = hipGetLastError();
    int herr = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, derr, 4, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);  // the host job table must outlive the copy
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return RVSP_ERROR_UNSUPPORTED_BACKEND;
    return herr ? RVSP_ERROR_INVALID_CSR : RVSP_SUCCESS;
}

rvsp_status_t rvsp_hip_unpack_row_ptr(void *stream, int32_t count, const int64_t *row_off,
                                      const int32_t *rows, const int32_t *big_rp,
                                      int32_t *out_rp, int64_t *bases_host)
{
    if (count <= 0) return RVSP_SUCCESS;
    hipStream_t st = (hipStream_t)stream;
    std::vector<UnpackJob> jobs((size_t)count);
    int tiles = 0;
    int64_t out = 0;
    for (int i = 0; i < count; ++i) {
        jobs[i] = UnpackJob{row_off[i], out, rows[i], tiles};
        tiles += (rows[i] + 1 + kPackTile - 1) / kPackTile;
        out += rows[i] + 1;
    }
    UnpackJob *d = nullptr;
    int64_t *dbase = nullptr;
    if (hipMallocAsync((void **)&d, jobs.size() * sizeof(UnpackJob), st) != hipSuccess)
        return RVSP_ERROR_ALLOCATION_FAILED;
    if (hipMallocAsync((void **)&dbase, (size_t)count * 8, st) != hipSuccess) {
        (void)hipFreeAsync(d, st);
        return RVSP_ERROR_ALLOCATION_FAILED;
    }
    rvsp_status_t s = RVSP_SUCCESS;
    if (hipMemcpyAsync(d, jobs.data(), jobs.size() * sizeof(UnpackJob), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    if (s == RVSP_SUCCESS) {
        k_unpack_rp<<<tiles, 256, 0, st>>>(d, count, big_rp, out_rp, dbase);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(bases_host, dbase, (size_t)count * 8, hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    (void)hipFreeAsync(d, st);
    (void)hipFreeAsync(dbase, st);
    return s;
}

} // extern "C"
