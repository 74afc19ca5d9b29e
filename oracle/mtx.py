"""Matrix Market reading, restated for checking rvsp_mtx_load, and the
binding of rv-sparse's own reader (oracle/_ref/librvsp_mtx_ref.so, built from
tools/src/mtx_to_csr_formatter.c and tools/src/vec.c by build_ref.sh).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The restatement follows rv-sparse tools/src/mtx_to_csr_formatter.c:
  banner (:36-56): field "pattern" -> values 1.0, symmetry "symmetric" ->
      mirror the off-diagonal entries (:224-247), after all stored ones;
  comments skipped up to the "rows cols entries" line (:27-34, :170-178);
  entries 1-based, read until the declared count or the first that does not
      parse (:196-215), values parsed as float;
  rows built in the order read (:249-298), then each row sorted by column
      and repeated columns summed in float in sorted order (canonicalise_rows,
      :77-145; glibc's qsort is a merge sort there, so equal columns keep
      their order);
and bench/bench.c:465-469 for the column count (largest column + 1).
Values are parsed through Python's float then rounded to float32, which can
differ from strtof by one unit in the last place for decimals within a
hair of a float32 rounding boundary; the fixtures avoid such decimals and
the reference binding checks the rest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_MTX_PATH = os.path.join(HERE, "_ref", "librvsp_mtx_ref.so")


def load(path):
    """(rows, cols, row_ptr, col_idx, values float32) as rv-sparse reads `path`."""
    with open(path) as f:
        lines = f.read().split("\n")
    first = lines[0] if lines else ""
    pattern = symmetric = False
    parts = first.split()
    if first.startswith("%%MatrixMarket") and len(parts) >= 5:
        pattern = parts[3] == "pattern"
        symmetric = parts[4] == "symmetric"
    i = 1
    while i < len(lines) and lines[i].startswith("%"):
        i += 1
    M, N, L = (int(x) for x in lines[i].split()[:3])
    toks = " ".join(lines[i + 1:]).split()
    per = 2 if pattern else 3
    rows, cols, vals = [], [], []
    for e in range(L):
        t = toks[e * per:(e + 1) * per]
        if len(t) < per:
            break
        rows.append(int(t[0]) - 1)
        cols.append(int(t[1]) - 1)
        vals.append(np.float32(1.0) if pattern else np.float32(float(t[2])))
    if symmetric:
        n = len(rows)
        for k in range(n):
            if rows[k] != cols[k]:
                rows.append(cols[k])
                cols.append(rows[k])
                vals.append(vals[k])
    per_row = [[] for _ in range(M)]
    for r, c, v in zip(rows, cols, vals):
        per_row[r].append((c, v))
    rp, ci, va = [0], [], []
    for lst in per_row:
        lst.sort(key=lambda cv: cv[0])  # stable
        t = 0
        while t < len(lst):
            c, acc = lst[t][0], np.float32(lst[t][1])
            t += 1
            while t < len(lst) and lst[t][0] == c:
                acc = np.float32(acc + np.float32(lst[t][1]))
                t += 1
            ci.append(c)
            va.append(acc)
        rp.append(len(ci))
    ncols = max(ci) + 1 if ci else N
    return (M, ncols, np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float32))


class _Vec(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_size_t), ("capacity", C.c_size_t),
                ("elem_size", C.c_size_t)]


class _CSR(C.Structure):
    _fields_ = [("row_ptr", C.POINTER(_Vec)), ("col_ind", C.POINTER(_Vec)), ("val", C.POINTER(_Vec))]


_lib = None


def ref_available() -> bool:
    return os.path.exists(REF_MTX_PATH)


def ref_load(path):
    """rv-sparse's assemble_csr_matrix on `path`, with bench.c's column
    inference: (rows, cols, row_ptr, col_idx, values) or None if it failed."""
    global _lib
    if _lib is None:
        _lib = C.CDLL(REF_MTX_PATH)
        _lib.assemble_csr_matrix.restype = _CSR
        _lib.assemble_csr_matrix.argtypes = [C.c_char_p]
        _lib.vector_free.restype = None
        _lib.vector_free.argtypes = [C.POINTER(_Vec)]
    m = _lib.assemble_csr_matrix(path.encode())
    if not m.row_ptr or not m.col_ind:
        return None

    def arr(v, dt):
        n = v.contents.size
        out = np.ctypeslib.as_array(C.cast(v.contents.data, C.POINTER(
            C.c_int32 if dt == np.int32 else C.c_float)), shape=(n,)).copy() if n else np.zeros(0, dt)
        return out.astype(dt)

    rp, ci, va = arr(m.row_ptr, np.int32), arr(m.col_ind, np.int32), arr(m.val, np.float32)
    for v in (m.row_ptr, m.col_ind, m.val):
        _lib.vector_free(v)
    rows = rp.size - 1
    cols = int(ci.max()) + 1 if ci.size else 0
    return rows, cols, rp, ci, va
