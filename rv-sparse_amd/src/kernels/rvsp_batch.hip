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
__global__ __launch_bounds__(256) void k_pack(const PackJob *__restrict__ jobs, int njobs,
                                              int32_t *__restrict__ orp, int32_t *__restrict__ oci,
                                              float *__restrict__ ova)
{
    const int j = find_job(jobs, njobs, blockIdx.x);
    const PackJob jb = jobs[j];
    const int64_t t0 = (int64_t)(blockIdx.x - jb.tile0) * kPackTile;
    for (int64_t e = t0 + threadIdx.x; e < t0 + kPackTile; e += 256) {
        if (e <= jb.rows) orp[jb.row_off + e] = (int32_t)(jb.rp[e] + jb.nnz_off);
        if (e < jb.nnz) {
            oci[jb.nnz_off + e] = jb.ci[e] + jb.col_shift;
            ova[jb.nnz_off + e] = jb.va[e];
        }
    }
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
                          it.row_off, it.nnz_off, it.col_shift, tiles};
        tiles += (int)((work + kPackTile - 1) / kPackTile);
    }
    PackJob *d = nullptr;
    if (hipMallocAsync((void **)&d, jobs.size() * sizeof(PackJob), st) != hipSuccess)
        return RVSP_ERROR_ALLOCATION_FAILED;
    if (hipMemcpyAsync(d, jobs.data(), jobs.size() * sizeof(PackJob), hipMemcpyHostToDevice, st) !=
        hipSuccess) {
        (void)hipFreeAsync(d, st);
        return RVSP_ERROR_UNSUPPORTED_BACKEND;
    }
    k_pack<<<tiles, 256, 0, st>>>(d, count, out_rp, out_ci, out_va);
    const hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(st);  // the host job table must outlive the copy
    (void)hipFreeAsync(d, st);
    return e == hipSuccess ? RVSP_SUCCESS : RVSP_ERROR_UNSUPPORTED_BACKEND;
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
