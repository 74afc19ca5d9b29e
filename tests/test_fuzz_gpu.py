"""Seeded randomized parity: forty FP32 products whose structure comes from the
library's genmat with parameters drawn per case (square and rectangular,
banded, column-wise degrees, log-normal and normal degree spreads, imbalance
caps), signed U[-1,1) FP32 values so that sums cancel, run through the host
one-shot (heuristic bins) and the device one-shot (learned bins) and
compared bit for bit with the oracle; and twenty-four FP64 and INT8 -> INT32
products of the same kind through the legacy dtypes, compared bit for bit
with rv-sparse's own legacy kernels (oracle/_ref). The cases print their
parameters, so a failure names its case."""
import ctypes as C
import numpy as np
import pytest

import rvsp
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def gen_csr(rows, cols, rng, **kw):
    got = rvsp.genmat(rows, cols, **kw)
    assert got is not None, (rows, cols, kw)
    r, c, rp, ci, _ = got
    va = rng.uniform(-1, 1, ci.size).astype(np.float32)
    return rvsp.Csr(r, c, rp, ci, va)


def draw_params(rng, rows, cols):
    kw = dict(density=float(rng.choice([0.002, 0.01, 0.05, 0.2])),
              cv=float(rng.choice([0.1, 0.5, 1.0])), min=int(rng.integers(1, 3)),
              random_seed=int(rng.integers(1, 1 << 30)))
    if rng.random() < 0.3:
        kw["imbalance"] = float(rng.choice([1.0, 4.0, 50.0]))
    if rng.random() < 0.25:
        kw["is_column"] = 1
    if rows == cols and rng.random() < 0.4:
        kw["low_bandwidth"] = int(rng.integers(1, rows + 1))
        kw["up_bandwidth"] = int(rng.integers(1, rows + 1))
    return kw


def assert_same(C, want, what):
    rp, ci, va = want[:3]
    np.testing.assert_array_equal(C.row_ptr, rp, err_msg=what)
    np.testing.assert_array_equal(C.col_idx, ci, err_msg=what)
    np.testing.assert_array_equal(C.values.view(np.uint32), np.asarray(va, np.float32).view(np.uint32),
                                  err_msg=what)


@pytest.mark.parametrize("case", range(40))
def test_random_generated_products(case):
    rng = np.random.default_rng(1000 + case)
    m = int(rng.integers(1, 3000))
    k = m if rng.random() < 0.5 else int(rng.integers(1, 3000))
    n = int(rng.integers(1, 6000))
    # at most 4e7 products, so the oracle answers in about a second: denser
    # draws are thinned until the pair fits
    for _ in range(8):
        ka, kb = draw_params(rng, m, k), draw_params(rng, k, n)
        A, B = gen_csr(m, k, rng, **ka), gen_csr(k, n, rng, **kb)
        if int(np.asarray(pyoracle.symbolic(A, B)[1]).sum()) <= 4e7:  # products (>= nnz)
            break
        m, n = max(1, m // 2), max(1, n // 2)
    what = f"case {case}: A {m}x{k} {ka}, B {k}x{n} {kb}"
    print(what, flush=True)
    want = pyoracle.spgemm(A, B)
    st, C1 = rvsp.spgemm_csr(A, B, raise_on_error=False)
    assert st == rvsp.SUCCESS, what
    assert_same(C1, want, what + " (host one-shot)")
    C2 = rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B), algo=rvsp.ALGO_HIP_LEARNED)
    assert_same(C2, want, what + " (device one-shot, learned)")


LEGACY = {rvsp.DTYPE_FP64: (rvsp.DTYPE_FP64, np.float64, C.c_double),
          rvsp.DTYPE_INT8: (rvsp.DTYPE_INT32, np.int32, C.c_int32)}


def legacy_struct(M, dtype, rng):
    if dtype == rvsp.DTYPE_FP64:
        va = rng.uniform(-1, 1, M.nnz)
    else:
        va = rng.integers(-128, 128, M.nnz).astype(np.int8)
    keep = (M.row_ptr, M.col_idx, va)
    s = rvsp.CsrStruct(M.rows, M.cols, M.nnz, M.row_ptr.ctypes.data, M.col_idx.ctypes.data,
                       va.ctypes.data, dtype, rvsp.FORMAT_CSR, 0)
    return s, keep


def legacy_product(lib, a, b, din):
    dout, npt, ct = LEGACY[din]
    c = rvsp.CsrStruct()
    opt = rvsp.Options(rvsp.BACKEND_SCALAR, din, dout)
    st = lib.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt))
    assert st == rvsp.SUCCESS, st
    n = c.nnz
    rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (c.rows + 1,)).copy()
    ci = (np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (n,)).copy()
          if n else np.zeros(0, np.int32))
    va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), (n,)).copy() if n else np.zeros(0, npt)
    lib.rvsp_csr_destroy(C.byref(c))
    return rp, ci, va.view(np.uint64) if npt == np.float64 else va


@pytest.mark.parametrize("din", [rvsp.DTYPE_FP64, rvsp.DTYPE_INT8])
@pytest.mark.parametrize("case", range(12))
def test_random_generated_legacy_products(case, din):
    if not pyoracle.ref_available():
        pytest.skip("rv-sparse reference library not built (oracle/build_ref.sh)")
    R = pyoracle.ref()
    R.rvsp_spgemm_csr.restype = C.c_int
    R.rvsp_spgemm_csr.argtypes = [C.c_void_p] * 4
    R.rvsp_csr_destroy.restype = None
    R.rvsp_csr_destroy.argtypes = [C.c_void_p]
    rng = np.random.default_rng(5000 + case)
    m, k, n = (int(x) for x in rng.integers(1, 1500, 3))
    ka, kb = draw_params(rng, m, k), draw_params(rng, k, n)
    what = f"legacy dtype {din} case {case}: A {m}x{k} {ka}, B {k}x{n} {kb}"
    print(what, flush=True)
    A, B = gen_csr(m, k, rng, **ka), gen_csr(k, n, rng, **kb)
    (a, ka_), (b, kb_) = legacy_struct(A, din, rng), legacy_struct(B, din, rng)
    got = legacy_product(rvsp.lib(), a, b, din)
    want = legacy_product(R, a, b, din)
    for g, w, name in zip(got, want, ("row_ptr", "col_idx", "values")):
        np.testing.assert_array_equal(g, w, err_msg=f"{what}: {name}")
