"""The five benchmark workloads, with fixed seeds, and scaled-down versions
of each for tests. Every matrix comes from librvsp_gen.so (one seeded PRNG),
so a given (workload, scale) is identical wherever it is generated.

  uniform   A 4096 x 4096, 16 distinct uniform columns per row, U[-1,1);
            C = A*A                                        (CPU reference case)
  rmat      A 2^18 x 2^18 R-MAT (.57,.19,.19,.05), 16*2^18 edge draws,
            duplicates merged, U[-1,1); C = A*A            (skewed rows)
  stencil   27-point stencil on 64^3, N(0,1); analysis once, numeric passes
            with values redrawn per pass                   (descriptor reuse)
  batched   512 square problems, n_i = floor(exp(U[ln 256, ln 16384])),
            nnz/row U{4..64} uniform columns, U[-1,1)      (many uneven products)
  mixed     A 2^20 x 2^16 power-law rows (rank r length ~ r^-1.2, mean 8,
            max 4096) x B 2^16 x 2^20 banded (12 of [16k-16, 16k+16)), N(0,1)
"""
from __future__ import annotations

from . import (Csr, gen_banded, gen_random_rows, gen_rmat, gen_stencil27, gen_zipf_rows,
               loguniform_sizes)

SEED_UNIFORM = 1
SEED_RMAT = 2
SEED_STENCIL = 3
SEED_BATCH = 4
SEED_MIXED_A = 5
SEED_MIXED_B = 6


def uniform(n=4096, per_row=16, seed=SEED_UNIFORM):
    A = gen_random_rows(n, n, per_row, per_row, seed)
    return A, A


def rmat(scale=18, edge_factor=16, seed=SEED_RMAT):
    A = gen_rmat(scale, edge_factor << scale, 0.57, 0.19, 0.19, seed)
    return A, A


def stencil(g=64, seed=SEED_STENCIL):
    A = gen_stencil27(g, seed, normal=True)
    return A, A


def batched(count=512, lo=256, hi=16384, nnz_lo=4, nnz_hi=64, seed=SEED_BATCH):
    sizes = loguniform_sizes(count, lo, hi, seed)
    out = []
    for i, n in enumerate(sizes):
        n = int(n)
        A = gen_random_rows(n, n, nnz_lo, nnz_hi, seed * 1_000_003 + 2 * i + 1)
        B = gen_random_rows(n, n, nnz_lo, nnz_hi, seed * 1_000_003 + 2 * i + 2)
        out.append((A, B))
    return out


def mixed(m=1 << 20, k=1 << 16, n=1 << 20, mean=8.0, maxlen=4096, per_row=12, seed_a=SEED_MIXED_A,
          seed_b=SEED_MIXED_B):
    A = gen_zipf_rows(m, k, 1.2, mean, maxlen, seed_a, normal=True)
    stride = max(1, n // k)
    B = gen_banded(k, n, per_row, stride, 16, seed_b, normal=True)
    return A, B


# scaled-down versions used by the parity tests (seconds on the CPU oracle)
SMALL = {
    "uniform": lambda: uniform(n=1024, per_row=16, seed=11),
    "rmat": lambda: rmat(scale=12, seed=12),
    "stencil": lambda: stencil(g=12, seed=13),
    "batched": lambda: batched(count=6, lo=64, hi=2048, seed=14),
    "mixed": lambda: mixed(m=1 << 13, k=1 << 9, n=1 << 13, maxlen=256, seed_a=15, seed_b=16),
}


def products_bytes(A: Csr, B: Csr, nnz_c: int, same: bool):
    """Compulsory HBM bytes of one product: A (and B if distinct) read once,
    C's row pointer, column indices and values written once."""
    b = 4 * (A.rows + 1) + 8 * A.nnz
    if not same:
        b += 4 * (B.rows + 1) + 8 * B.nnz
    b += 4 * (A.rows + 1) + 8 * nnz_c
    return b
