"""Shared test helpers: fixture loading and CSR comparison."""
import json
import os

import numpy as np

import rvsp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixtures():
    with open(os.path.join(GOLDEN, "csr_fixtures.json")) as f:
        return json.load(f)


def csr_of(d) -> rvsp.Csr:
    return rvsp.Csr.make(d["rows"], d["cols"], d["row_ptr"], d["col_idx"], d["values"])


def dense_of(M: rvsp.Csr) -> np.ndarray:
    D = np.zeros((M.rows, M.cols), np.float64)
    for r in range(M.rows):
        for p in range(M.row_ptr[r], M.row_ptr[r + 1]):
            D[r, M.col_idx[p]] += M.values[p]
    return D


def structure_bruteforce(A: rvsp.Csr, B: rvsp.Csr):
    """Set-union definition of C's structure: cols(C_i) = sorted union of
    cols(B_k) over k in A_i; op_i = sum of |B_k|."""
    rp = [0]
    ci = []
    ops = []
    for r in range(A.rows):
        cols = set()
        o = 0
        for p in range(A.row_ptr[r], A.row_ptr[r + 1]):
            k = A.col_idx[p]
            cols.update(B.col_idx[B.row_ptr[k]:B.row_ptr[k + 1]].tolist())
            o += int(B.row_ptr[k + 1] - B.row_ptr[k])
        ci.extend(sorted(cols))
        rp.append(len(ci))
        ops.append(o)
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(ops, np.int64)


def assert_bitexact(got: rvsp.Csr, rp, ci, va, what=""):
    np.testing.assert_array_equal(got.row_ptr, rp, err_msg=f"{what}: row_ptr")
    np.testing.assert_array_equal(got.col_idx, ci, err_msg=f"{what}: col_idx")
    # bit-exact values (compare the binary32 patterns: +0 vs -0 counts)
    gb = got.values.view(np.uint32)
    wb = np.asarray(va, np.float32).view(np.uint32)
    bad = np.nonzero(gb != wb)[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} values differ, first at {bad[:5].tolist()}: "
        f"got {got.values[bad[:5]].tolist()} want {np.asarray(va)[bad[:5]].tolist()}")


def rel_err_stats(got32, want64):
    got = np.asarray(got32, np.float64)
    den = np.maximum(np.abs(want64), 1e-300)
    rel = np.abs(got - want64) / den
    nz = np.abs(want64) > 0
    if not nz.any():
        return 0.0, 0.0
    return float(rel[nz].max()), float(rel[nz].mean())
