"""Writes tests/golden/ref_vectors.npz: small inputs and the outputs of
rv-sparse's own scalar kernel (rvsp_spgemm_scalar_f32, compiled from the
reference sources by oracle/build_ref.sh). Run in a container that has the
reference tree; the vectors are data and travel without it.

    python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "rv-sparse_amd")]

import rvsp  # noqa: E402
from oracle import pyoracle  # noqa: E402
from rvsp import workloads as W  # noqa: E402


def rand_pair(seed, m, k, n, dens):
    rng = np.random.default_rng(seed)

    def mk(r, c):
        mask = rng.random((r, c)) < dens
        rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
        ci = np.nonzero(mask)[1].astype(np.int32)
        va = rng.uniform(-1, 1, ci.size).astype(np.float32)
        return rvsp.Csr.make(r, c, rp, ci, va)

    return mk(m, k), mk(k, n)


def main():
    if not pyoracle.ref_available():
        sys.exit("oracle/_ref/librvsp_ref.so missing: run oracle/build_ref.sh first")
    cases = {
        "rand_a": rand_pair(1, 40, 30, 50, 0.1),
        "rand_b": rand_pair(2, 120, 90, 100, 0.04),
        "rand_tall": rand_pair(3, 200, 20, 30, 0.2),
        "stencil_g5": W.stencil(g=5, seed=21),
        "uniform_256": W.uniform(n=256, per_row=8, seed=22),
        "rmat_s8": W.rmat(scale=8, seed=23),
        "mixed_tiny": W.mixed(m=1 << 9, k=1 << 6, n=1 << 9, maxlen=32, seed_a=24, seed_b=25),
    }
    out = {}
    for name, (A, B) in cases.items():
        rp, ci, va, ops = pyoracle.ref_spgemm(A, B, want_ops=True)
        for tag, M in (("a", A), ("b", B)):
            out[f"{name}__{tag}_shape"] = np.array([M.rows, M.cols], np.int64)
            out[f"{name}__{tag}_rp"] = M.row_ptr
            out[f"{name}__{tag}_ci"] = M.col_idx
            out[f"{name}__{tag}_va"] = M.values
        out[f"{name}__c_rp"] = rp
        out[f"{name}__c_ci"] = ci
        out[f"{name}__c_va"] = va
        out[f"{name}__ops"] = ops
        print(f"{name}: m={A.rows} nnzA={A.nnz} nnzC={len(ci)} products={int(ops.sum())}")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
