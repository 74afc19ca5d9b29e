"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, when it was
built, of rv-sparse's own scalar library (oracle/_ref/librvsp_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker or the timed CPU
baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "librvsp_ref.so")

_o = None
_r = None


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run make -C oracle")
        o = C.CDLL(ORACLE_PATH)
        I32, I64, V = C.c_int32, C.c_int64, C.c_void_p
        o.orc_csr_check.restype = C.c_int
        o.orc_csr_check.argtypes = [I32, I32, V, V, C.POINTER(I32)]
        o.orc_symbolic_count.restype = C.c_int
        o.orc_symbolic_count.argtypes = [I32, I32, V, V, V, V, V, V, C.POINTER(I64)]
        o.orc_symbolic_fill.restype = None
        o.orc_symbolic_fill.argtypes = [I32, I32, V, V, V, V, V, V]
        o.orc_numeric_f32.restype = None
        o.orc_numeric_f32.argtypes = [I32, I32, V, V, V, V, V, V, V, V, V]
        o.orc_numeric_f64.restype = None
        o.orc_numeric_f64.argtypes = [I32, I32, V, V, V, V, V, V, V, V, V]
        o.orc_symbolic_count_par.restype = C.c_int
        o.orc_symbolic_count_par.argtypes = [I32, I32, V, V, V, V, V, V, C.POINTER(I64), C.c_int]
        o.orc_fill_numeric_par.restype = C.c_int
        o.orc_fill_numeric_par.argtypes = [I32, I32, V, V, V, V, V, V, V, V, V, V, C.c_int]
        for name in ("orc_spgemm_f32", "orc_spgemm_f32_omp"):
            f = getattr(o, name)
            f.restype = C.c_int
            extra = [C.c_int] if name.endswith("omp") else []
            f.argtypes = [I32, I32, I32, V, V, V, V, V, V, C.POINTER(V), C.POINTER(V), C.POINTER(V), C.POINTER(I32)] + extra
        o.orc_free.restype = None
        o.orc_free.argtypes = [V]
        o.orc_now.restype = C.c_double
        o.orc_now.argtypes = []
        _o = o
    return _o


def _p(a):
    return a.ctypes.data


def threads_default() -> int:
    """Host threads for the row-parallel oracle: the CPUs this process may
    use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def symbolic(A, B, threads=0):
    """(row_ptr int32[m+1], ops int64[m], total) of C = A*B; raises on
    overflow. threads > 0: the row-parallel count (same outputs)."""
    o = oracle()
    rp = np.zeros(A.rows + 1, np.int32)
    ops = np.zeros(max(A.rows, 1), np.int64)
    tot = C.c_int64(0)
    if threads > 0:
        rc = o.orc_symbolic_count_par(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(B.row_ptr),
                                      _p(B.col_idx), _p(rp), _p(ops), C.byref(tot), threads)
    else:
        rc = o.orc_symbolic_count(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(B.row_ptr),
                                  _p(B.col_idx), _p(rp), _p(ops), C.byref(tot))
    if rc != 0:
        raise OverflowError(f"oracle symbolic count status {rc}")
    return rp, ops[: A.rows], tot.value


def spgemm(A, B, f64=False):
    """(row_ptr, col_idx, values_f32[, values_f64]) of C = A*B."""
    o = oracle()
    rp, ops, tot = symbolic(A, B)
    ci = np.zeros(max(tot, 1), np.int32)
    o.orc_symbolic_fill(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(B.row_ptr),
                        _p(B.col_idx), _p(rp), _p(ci))
    va = np.zeros(max(tot, 1), np.float32)
    o.orc_numeric_f32(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(A.values), _p(B.row_ptr),
                      _p(B.col_idx), _p(B.values), _p(rp), _p(ci), _p(va))
    out = [rp, ci[:tot], va[:tot]]
    if f64:
        vd = np.zeros(max(tot, 1), np.float64)
        o.orc_numeric_f64(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(A.values),
                          _p(B.row_ptr), _p(B.col_idx), _p(B.values), _p(rp), _p(ci), _p(vd))
        out.append(vd[:tot])
    return tuple(out)


def spgemm_par(A, B, threads=None, f64=True, sym=None):
    """spgemm() over row-parallel oracle passes: (row_ptr, col_idx, values_f32
    [, values_f64]). `sym` = a (row_ptr, ops, total) already counted."""
    o = oracle()
    threads = threads or threads_default()
    rp, _, tot = sym if sym is not None else symbolic(A, B, threads=threads)
    ci = np.zeros(max(tot, 1), np.int32)
    va = np.zeros(max(tot, 1), np.float32)
    vd = np.zeros(max(tot, 1), np.float64) if f64 else None
    rc = o.orc_fill_numeric_par(A.rows, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(A.values),
                                _p(B.row_ptr), _p(B.col_idx), _p(B.values), _p(rp), _p(ci),
                                _p(va), _p(vd) if f64 else None, threads)
    if rc != 0:
        raise MemoryError(f"oracle fill/numeric status {rc}")
    out = [rp, ci[:tot], va[:tot]]
    if f64:
        out.append(vd[:tot])
    return tuple(out)


def time_oneshot(A, B, threads=0):
    """Seconds for one full oracle one-shot (count, malloc C, fill+sort,
    numeric), the work rv-sparse's bench times. threads>0: OpenMP rows."""
    o = oracle()
    rp, ci, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int32(0)
    t0 = o.orc_now()
    if threads > 0:
        rc = o.orc_spgemm_f32_omp(A.rows, A.cols, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(A.values),
                                  _p(B.row_ptr), _p(B.col_idx), _p(B.values), C.byref(rp), C.byref(ci),
                                  C.byref(va), C.byref(nnz), threads)
    else:
        rc = o.orc_spgemm_f32(A.rows, A.cols, B.cols, _p(A.row_ptr), _p(A.col_idx), _p(A.values),
                              _p(B.row_ptr), _p(B.col_idx), _p(B.values), C.byref(rp), C.byref(ci),
                              C.byref(va), C.byref(nnz))
    t1 = o.orc_now()
    for p in (rp, ci, va):
        o.orc_free(p)
    if rc != 0:
        raise RuntimeError(f"oracle one-shot status {rc}")
    return t1 - t0


# ---------------------------------------------------------------------------
# rv-sparse's own scalar library, compiled from its sources by build_ref.sh
# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _r
    if _r is None:
        r = C.CDLL(REF_PATH)
        I32, I64, V, SZ = C.c_int32, C.c_int64, C.c_void_p, C.c_size_t
        r.rvsp_spgemm_scalar_f32.restype = C.c_int
        r.rvsp_spgemm_scalar_f32.argtypes = [I32, I32, I32, V, V, V, V, V, V, V, SZ,
                                             C.POINTER(V), C.POINTER(V), C.POINTER(V),
                                             C.POINTER(I32), V]
        r.rvsp_spgemm_buffer_size.restype = C.c_int
        r.rvsp_spgemm_buffer_size.argtypes = [I32, C.POINTER(SZ)]
        r.rvsp_csr_check.restype = C.c_int
        r.rvsp_csr_check.argtypes = [I32, I32, V, V, C.POINTER(I32)]
        _r = r
    return _r


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def ref_spgemm(A, B, want_ops=False, timed=False):
    """rv-sparse's rvsp_spgemm_scalar_f32 raw one-shot on host arrays.
    Returns (rp, ci, va[, ops][, seconds])."""
    r = ref()
    ws_b = C.c_size_t(0)
    assert r.rvsp_spgemm_buffer_size(B.cols, C.byref(ws_b)) == 0
    ws = np.zeros(ws_b.value, np.uint8)
    rp, ci, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int32(0)
    ops = np.zeros(max(A.rows, 1), np.int64) if want_ops else None
    t0 = oracle().orc_now()
    rc = r.rvsp_spgemm_scalar_f32(A.rows, A.cols, B.cols, _p(A.row_ptr), _p(A.col_idx),
                                  _p(A.values), _p(B.row_ptr), _p(B.col_idx), _p(B.values),
                                  _p(ws), ws_b.value, C.byref(rp), C.byref(ci), C.byref(va),
                                  C.byref(nnz), _p(ops) if want_ops else None)
    t1 = oracle().orc_now()
    if rc != 0:
        raise RuntimeError(f"reference scalar kernel status {rc}")
    n = nnz.value
    out_rp = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (A.rows + 1,)).copy()
    out_ci = np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), (max(n, 1),))[:n].copy()
    out_va = np.ctypeslib.as_array(C.cast(va, C.POINTER(C.c_float)), (max(n, 1),))[:n].copy()
    for p in (rp, ci, va):
        _libc.free(p)
    res = [out_rp, out_ci, out_va]
    if want_ops:
        res.append(ops[: A.rows])
    if timed:
        res.append(t1 - t0)
    return tuple(res)
