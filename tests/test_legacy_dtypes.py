"""Legacy dtypes of rvsp_spgemm_csr (FP64 in and out; INT8 in, INT32 out) on
the GPU, checked bit for bit against rv-sparse's own legacy scalar kernels
(src/kernels/spgemm/legacy/csr_scalar_f64.c, csr_scalar_i8.c, built into
oracle/_ref): same structure after dropping entries that cancel to zero,
same values (FP64 accumulated in ascending k, rounded per operation).
Argument checks run on the CPU; the products need a GPU."""
import ctypes as C

import numpy as np
import pytest

import rvsp
from rvsp import workloads as W
from oracle import pyoracle

FP64, INT8, INT32 = rvsp.DTYPE_FP64, rvsp.DTYPE_INT8, rvsp.DTYPE_INT32
NP = {FP64: np.float64, INT8: np.int8, INT32: np.int32, rvsp.DTYPE_FP32: np.float32}


class Mat:
    """A host CSR struct over numpy arrays of any legacy dtype."""

    def __init__(self, rows, cols, rp, ci, va, dtype):
        self.rp = np.ascontiguousarray(rp, np.int32)
        self.ci = np.ascontiguousarray(ci, np.int32)
        self.va = np.ascontiguousarray(va, NP[dtype])
        self.s = rvsp.CsrStruct(rows, cols, int(self.rp[-1]), self.rp.ctypes.data,
                                self.ci.ctypes.data, self.va.ctypes.data, dtype, rvsp.FORMAT_CSR, 0)


def with_values(A, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == FP64:
        va = rng.uniform(-1, 1, A.nnz)
    else:
        va = rng.integers(-128, 128, A.nnz).astype(np.int8)
    return Mat(A.rows, A.cols, A.row_ptr, A.col_idx, va, dtype)


def run(lib, A, B, din, dout, backend=rvsp.BACKEND_SCALAR):
    c = rvsp.CsrStruct()
    opt = rvsp.Options(backend, din, dout)
    st = lib.rvsp_spgemm_csr(C.byref(A.s), C.byref(B.s), C.byref(c), C.byref(opt))
    if st != 0:
        return st, None
    n = c.nnz
    rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (c.rows + 1,)).copy()
    ci = (np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (n,)).copy()
          if n else np.zeros(0, np.int32))
    ct = C.c_double if dout == FP64 else C.c_int32
    va = (np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), (n,)).copy()
          if n else np.zeros(0, NP[dout]))
    assert c.dtype == dout and c.owns_data == 1
    lib.rvsp_csr_destroy(C.byref(c))
    return 0, (rp, ci, va)


def ref_lib():
    if not pyoracle.ref_available():
        pytest.skip("rv-sparse reference library not built (oracle/build_ref.sh)")
    r = pyoracle.ref()
    r.rvsp_spgemm_csr.restype = C.c_int
    r.rvsp_spgemm_csr.argtypes = [C.c_void_p] * 4
    r.rvsp_csr_destroy.restype = None
    r.rvsp_csr_destroy.argtypes = [C.c_void_p]
    return r


def assert_same(got, want, what):
    for name, g, w in zip(("row_ptr", "col_idx", "values"), got, want):
        if name == "values" and g.dtype == np.float64:
            g, w = g.view(np.uint64), w.view(np.uint64)
        np.testing.assert_array_equal(g, w, err_msg=f"{what}: {name}")


# ---------------------------------------------------------------------------
# argument checks (CPU: they fail before any device work)
# ---------------------------------------------------------------------------
def test_checks_follow_the_legacy_wrappers():
    A = rvsp.gen_random_rows(20, 30, 1, 4, 1)
    B = rvsp.gen_random_rows(30, 25, 1, 4, 2)
    L = rvsp.lib()
    f32A, f32B = Mat(A.rows, A.cols, A.row_ptr, A.col_idx, A.values, rvsp.DTYPE_FP32), \
        Mat(B.rows, B.cols, B.row_ptr, B.col_idx, B.values, rvsp.DTYPE_FP32)
    assert run(L, f32A, f32B, FP64, FP64)[0] == rvsp.ERROR_INVALID_ARGUMENT  # dtype (wrapper :362)
    assert run(L, f32A, f32B, INT8, INT32)[0] == rvsp.ERROR_UNSUPPORTED_DTYPE  # wrapper :43
    a64 = with_values(A, FP64, 3)
    assert run(L, a64, a64, FP64, FP64)[0] == rvsp.ERROR_INVALID_ARGUMENT  # 30 cols x 20 rows
    # a B row that repeats a column: refused on the host, before any device work
    rep = B.col_idx.copy()
    r = int(np.nonzero(np.diff(B.row_ptr) >= 2)[0][0])
    rep[B.row_ptr[r] + 1] = rep[B.row_ptr[r]]
    b_rep = Mat(B.rows, B.cols, B.row_ptr, rep, np.ones(B.nnz), FP64)
    assert run(L, with_values(A, FP64, 3), b_rep, FP64, FP64)[0] == rvsp.ERROR_INVALID_CSR
    # a column out of range after an order violation still counts as the
    # range error the legacy kernels report
    late = B.col_idx.copy()
    late[B.row_ptr[r]], late[B.row_ptr[r] + 1] = late[B.row_ptr[r] + 1], late[B.row_ptr[r]]
    late[-1] = B.cols
    b_late = Mat(B.rows, B.cols, B.row_ptr, late, np.ones(B.nnz), FP64)
    assert run(L, with_values(A, FP64, 3), b_late, FP64, FP64)[0] == rvsp.ERROR_INVALID_ARGUMENT
    oob = A.col_idx.copy()
    oob[0] = 999
    assert run(L, Mat(A.rows, A.cols, A.row_ptr, oob, np.zeros(A.nnz), FP64), with_values(B, FP64, 5),
               FP64, FP64)[0] == rvsp.ERROR_INVALID_ARGUMENT
    # RVV with legacy dtypes stays unsupported, as in a non-vector rv-sparse build
    assert run(L, a64, with_values(B, FP64, 6), FP64, FP64,
               backend=rvsp.BACKEND_RVV_INTRINSICS)[0] == rvsp.ERROR_UNSUPPORTED_BACKEND


# ---------------------------------------------------------------------------
# products on the GPU against the reference legacy kernels
# ---------------------------------------------------------------------------
def pairs():
    yield "uniform", rvsp.gen_random_rows(400, 500, 2, 30, 11), rvsp.gen_random_rows(500, 600, 2, 30, 12)
    A, _ = W.rmat(scale=14, seed=21)  # wide rows: the typed big-row kernel
    yield "rmat14", A, A
    yield "wide", rvsp.gen_random_rows(40, 3000, 700, 700, 78), \
        rvsp.gen_random_rows(3000, 100000, 250, 350, 77)


@pytest.mark.gpu
@pytest.mark.parametrize("din,dout", [(FP64, FP64), (INT8, INT32)])
def test_products_match_reference_legacy_kernels(din, dout):
    R = ref_lib()
    L = rvsp.lib()
    for i, (name, A, B) in enumerate(pairs()):
        a, b = with_values(A, din, 100 + i), with_values(B, din, 200 + i)
        print(f"legacy {name} dtype={din}", flush=True)
        st, got = run(L, a, b, din, dout)
        assert st == 0, rvsp.status_string(st)
        st, want = run(R, a, b, din, dout)
        assert st == 0
        assert_same(got, want, f"{name} dtype {din}")
        # the hip backend name gives the same result
        assert_same(run(L, a, b, din, dout, backend=rvsp.BACKEND_HIP)[1], want, f"{name} HIP")


def shuffled_rows(M, seed, repeat_a=False):
    """M with every row's entries in a seeded random order (values moved
    with their columns); repeat_a also repeats each row's first entry at
    its end (A may hold a column twice, as the legacy kernels allow)."""
    rng = np.random.default_rng(seed)
    rp = [0]
    ci, va = [], []
    for r in range(M.rows):
        s, e = M.row_ptr[r], M.row_ptr[r + 1]
        order = rng.permutation(e - s) + s
        c, v = list(M.col_idx[order]), list(M.values[order])
        if repeat_a and e > s:
            c.append(c[0])
            v.append(v[0])
        ci += c
        va += v
        rp.append(len(ci))
    return rvsp.Csr(M.rows, M.cols, np.array(rp, np.int32), np.array(ci, np.int32),
                    np.array(va, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("din,dout", [(FP64, FP64), (INT8, INT32)])
def test_unsorted_rows_match_reference_legacy_kernels(din, dout):
    """Unsorted A rows (with repeated columns) and unsorted B rows, which the
    legacy kernels accept: same C bit for bit, for every row bin and the
    typed big-row kernel."""
    R = ref_lib()
    L = rvsp.lib()
    for i, (name, A, B) in enumerate(pairs()):
        As = shuffled_rows(A, 300 + i, repeat_a=True)
        Bs = shuffled_rows(B, 400 + i)
        a, b = with_values(As, din, 500 + i), with_values(Bs, din, 600 + i)
        print(f"legacy unsorted {name} dtype={din}", flush=True)
        st, got = run(L, a, b, din, dout)
        assert st == 0, rvsp.status_string(st)
        st, want = run(R, a, b, din, dout)
        assert st == 0
        assert_same(got, want, f"unsorted {name} dtype {din}")


@pytest.mark.gpu
def test_cancelled_entries_are_dropped():
    """C[0,0] = 2*3 + (-2)*3 = 0 and C[0,1] = -0.5*0 ... cancel exactly and
    are not stored (legacy kernels store nonzero results only); C[0,2] stays."""
    R = ref_lib()
    L = rvsp.lib()
    # A: 1x2 [2, -2]; B: 2x3 rows [3, 1, 0] and [3, 0, 4] (structure (0,0),(0,1) / (1,0),(1,2))
    a = Mat(1, 2, [0, 2], [0, 1], [2.0, -2.0], FP64)
    b = Mat(2, 3, [0, 2, 4], [0, 1, 0, 2], [3.0, -0.0, 3.0, 4.0], FP64)
    st, got = run(L, a, b, FP64, FP64)
    assert st == 0
    np.testing.assert_array_equal(got[0], [0, 1])
    np.testing.assert_array_equal(got[1], [2])
    assert got[2][0] == -8.0
    assert_same(got, run(R, a, b, FP64, FP64)[1], "cancellation")
    ai = Mat(1, 2, [0, 2], [0, 1], [5, -5], INT8)
    bi = Mat(2, 3, [0, 2, 4], [0, 1, 0, 2], [7, 0, 7, 1], INT8)
    st, got = run(L, ai, bi, INT8, INT32)
    assert st == 0
    np.testing.assert_array_equal(got[1], [2])
    assert got[2][0] == -5
    assert_same(got, run(R, ai, bi, INT8, INT32)[1], "int cancellation")
