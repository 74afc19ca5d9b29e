// rvsp_mlp.hpp -- device side of the learned row selectors: the MLP
// forward pass, the per-row features and the chooser's candidate
// accumulators. Shared by the selector kernels and the numeric binning.
#pragma once

#include "rvsp_device.hpp"
#include "rv_sparse_hip.h"

namespace rvsp {

template <int NF, int NO>
struct Mlp {
    static constexpr int H1 = 16, H2 = 16;
    static constexpr int W1 = 0;
    static constexpr int B1 = W1 + H1 * NF;
    static constexpr int W2 = B1 + H1;
    static constexpr int B2 = W2 + H2 * H1;
    static constexpr int W3 = B2 + H2;
    static constexpr int B3 = W3 + NO * H2;
    static constexpr int NP = B3 + NO;

    __device__ static void forward(const float *P, const float *f, float *z1, float *h1, float *z2,
                                   float *h2, float *y)
    {
#pragma unroll
        for (int i = 0; i < H1; ++i) {
            float z = P[B1 + i];
#pragma unroll
            for (int j = 0; j < NF; ++j) z += P[W1 + i * NF + j] * f[j];
            z1[i] = z;
            h1[i] = z > 0.0f ? z : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < H2; ++i) {
            float z = P[B2 + i];
#pragma unroll
            for (int j = 0; j < H1; ++j) z += P[W2 + i * H1 + j] * h1[j];
            z2[i] = z;
            h2[i] = z > 0.0f ? z : 0.0f;
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            float v = P[B3 + o];
#pragma unroll
            for (int i = 0; i < H2; ++i) v += P[W3 + o * H2 + i] * h2[i];
            y[o] = v;
        }
    }
};

using Sizer = Mlp<4, 1>;
using Chooser = Mlp<5, 8>;
static_assert(Sizer::NP == RVSP_SIZER_PARAMS, "sizer parameter count");
static_assert(Chooser::NP == RVSP_CHOOSER_PARAMS, "chooser parameter count");

constexpr int kFeat = 5;  // feature rows are stored 5 wide for both nets

__device__ __forceinline__ void row_features(int64_t ops, int alen, int64_t span, int nz,
                                             float f[kFeat])
{
    const float a = (float)max(alen, 1);
    f[0] = log2f(1.0f + (float)alen) * (1.0f / 16.0f);
    f[1] = log2f(1.0f + (float)ops) * (1.0f / 32.0f);
    f[2] = log2f(1.0f + (float)span) * (1.0f / 32.0f);
    f[3] = log2f(1.0f + (float)ops / a) * (1.0f / 16.0f);
    f[4] = log2f(1.0f + (float)nz) * (1.0f / 32.0f);
}

// Numeric bin of chooser output o for a row, or -1 when not feasible
// (only the instantiated (H, G), (W, G) and (CAP, G) triples are). Outputs
// 0-3: hash table; 4-7: window (rank window when the row fits one, G is 32
// or 64 and its LDS footprint is no larger than a dense window's, else
// dense window).
__device__ __forceinline__ int chooser_bin(int o, int nz, int64_t span)
{
    const int lg = 3 + (o & 3);
    if (o < 4) {
        if (nz > kHashMaxNnz) return -1;
        const int H = max(64, next_pow2((3 * nz + 1) / 2));
        if ((1 << lg) < (H >= 4096 ? 32 : H >= 2048 ? 16 : 8)) return -1;
        return NB_HASH0 + 4 * (ilog2(H) - 6) + (lg - 3);
    }
    // window kind by LDS footprint, as the heuristic decides it: a rank
    // window (32 or 64 lanes) unless a dense window over the span is smaller
    const bool rank_ok = span <= kRankMaxCols && nz <= kRankMaxNnz && lg >= 5;
    const int C = max(512, next_pow2(nz));
    const int W = span <= kDenseMaxCols ? max(512, next_pow2((int)span)) : 0;
    const bool dense_ok =
        W > 0 && (1 << lg) >= (W >= 16384 ? 64 : W >= 8192 ? 32 : W >= 4096 ? 16 : 8);
    if (rank_ok && (!dense_ok || kRankWords * 6 + 4 * C <= 4 * W))
        return NB_RANK0 + 2 * (ilog2(C) - 9) + (lg == 6 ? 1 : 0);
    if (!dense_ok) return -1;
    return NB_DENSE0 + 4 * (ilog2(W) - 9) + (lg - 3);
}

// The feasible accumulator with the smallest predicted time (-1: none).
// When the heuristic's pick `hb` is one of the outputs, the chooser keeps it
// unless the best prediction is faster by more than lg_margin (log2 of the
// chooser's margin: 0 keeps the plain argmin).
__device__ inline int chooser_pick(const float *P, int64_t ops, int alen, int64_t span, int nz,
                                   int hb, float lg_margin)
{
    float f[kFeat], z1[16], h1[16], z2[16], h2[16], y[8];
    row_features(ops, alen, span, nz, f);
    Chooser::forward(P, f, z1, h1, z2, h2, y);
    int best = -1, oh = -1;
    float by = 0.0f;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int b = chooser_bin(o, nz, span);
        if (b >= 0 && (best < 0 || y[o] < by)) {
            best = b;
            by = y[o];
        }
        if (b >= 0 && b == hb) oh = o;
    }
    if (oh >= 0 && !(by + lg_margin < y[oh])) return hb;
    return best;
}

} // namespace rvsp
