// rvsp_kernels.hpp -- host-side declarations shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rvsp {

// Counters the analysis kernels accumulate; copied to the host once per
// analysis phase.
struct AnalysisCounters {
    int sym_hist[16];
    int num_hist[32];
    unsigned long long products;
    unsigned long long big_cursor_cap;  // sum over big rows of alen*(nwin+1)
    unsigned long long big_windows;     // sum over big rows of nwin
    unsigned long long cursor_fill;     // running allocation of cursors
    int nitems;                         // window items emitted
    int pad;
    long long nnz_total;                // scan total
};

struct SymArgs {
    const int32_t *arp, *aci, *brp, *bci;
    const int32_t *lo, *hi;
    int32_t *cnt;
    const int32_t *crp;
    int32_t *cci;
    int64_t *cur_off;
    int32_t *cursors;
    unsigned long long *cursor_fill;
    int *err;
};

struct NumArgs {
    const int32_t *arp, *aci;
    const float *aval;
    const int32_t *brp, *bci;
    const float *bval;
    const int32_t *crp, *cci;
    float *cval;
    const int32_t *lo;
    const int64_t *cur_off;
    const int32_t *cursors;
    int *err;
};

void launch_row_stats(hipStream_t st, int32_t m, const int32_t *arp, const int32_t *aci,
                      const int32_t *brp, const int32_t *bci, int64_t *ops, int32_t *lo,
                      int32_t *hi, uint8_t *sbin, AnalysisCounters *ctr);
void launch_bin_scatter(hipStream_t st, int32_t m, const uint8_t *bin, const int32_t *off,
                        int32_t *cursor, int32_t *perm, int nb);
void launch_sym_count(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                      const int32_t *off);
void launch_sym_fill(hipStream_t st, const SymArgs &a, const int32_t *perm, const int32_t *count,
                     const int32_t *off);
void launch_scan(hipStream_t st, int32_t m, const int32_t *cnt, int64_t *tsum, int64_t *total,
                 int32_t *crp);
int scan_tiles(int32_t m);
void launch_num_bins(hipStream_t st, int32_t m, const int32_t *crp, const int64_t *ops,
                     const int32_t *arp, const uint8_t *sbin, const int32_t *lo,
                     const int32_t *hi, uint8_t *nbin, int2 *items, AnalysisCounters *ctr);
void launch_numeric(hipStream_t st, const NumArgs &a, const int32_t *perm, const int32_t *count,
                    const int32_t *off, const int2 *items, int nitems);

} // namespace rvsp
