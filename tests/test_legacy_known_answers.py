"""The known-answer cases of rv-sparse's own tests under tests/legacy/, run
through rvsp_spgemm_csr on the GPU with the options those tests pass:
test_spgemm_csr_f32.c (SCALAR, FP32), test_spgemm_csr_unroll4_f32.c
(SCALAR and SCALAR_UNROLL4 give the same C; its product also checked here
against the hand-computed closed form) and test_spgemm_csr_i8.c (SCALAR,
INT8 in, INT32 out)."""
import ctypes as C

import numpy as np
import pytest

import rvsp

pytestmark = pytest.mark.gpu

NP = {rvsp.DTYPE_FP32: (np.float32, C.c_float), rvsp.DTYPE_INT8: (np.int8, C.c_int8),
      rvsp.DTYPE_INT32: (np.int32, C.c_int32)}


class Mat:
    def __init__(self, rows, cols, rp, ci, va, dtype):
        self.rp = np.asarray(rp, np.int32)
        self.ci = np.asarray(ci, np.int32)
        self.va = np.asarray(va, NP[dtype][0])
        self.s = rvsp.CsrStruct()
        st = rvsp.lib().rvsp_csr_create(C.byref(self.s), rows, cols, len(self.ci), self.rp.ctypes.data,
                                        self.ci.ctypes.data, self.va.ctypes.data, dtype)
        assert st == rvsp.SUCCESS


def product(A, B, backend, din, dout):
    L = rvsp.lib()
    c = rvsp.CsrStruct()
    opt = rvsp.Options(backend, din, dout)
    assert L.rvsp_spgemm_csr(C.byref(A.s), C.byref(B.s), C.byref(c), C.byref(opt)) == rvsp.SUCCESS
    assert c.dtype == dout and c.owns_data == 1
    n = c.nnz
    rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (c.rows + 1,)).copy()
    ci = np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (n,)).copy()
    va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(NP[dout][1])), (n,)).copy()
    shape = (c.rows, c.cols)
    L.rvsp_csr_destroy(C.byref(c))
    return shape, rp, ci, va


def test_f32_two_by_two():
    """test_spgemm_csr_f32.c: [[1,2],[0,3]] x [[4,0],[0,5]] = [[4,10],[0,15]]."""
    A = Mat(2, 2, [0, 2, 3], [0, 1, 1], [1, 2, 3], rvsp.DTYPE_FP32)
    B = Mat(2, 2, [0, 1, 2], [0, 1], [4, 5], rvsp.DTYPE_FP32)
    for backend in (rvsp.BACKEND_SCALAR, rvsp.BACKEND_HIP):
        shape, rp, ci, va = product(A, B, backend, rvsp.DTYPE_FP32, rvsp.DTYPE_FP32)
        assert shape == (2, 2)
        assert rp.tolist() == [0, 2, 3] and ci.tolist() == [0, 1, 1]
        assert va.tolist() == [4.0, 10.0, 15.0]


def test_unroll4_matches_scalar_and_closed_form():
    """test_spgemm_csr_unroll4_f32.c's pair: A 2x4, B 4x3. Row 0 =
    1*B_0 + 2*B_1 + 4*B_3 = [1, 30, 30]; row 1 = 3*B_1 + 5*B_2 = [20, 9, 25]."""
    A = Mat(2, 4, [0, 3, 5], [0, 1, 3, 1, 2], [1, 2, 4, 3, 5], rvsp.DTYPE_FP32)
    B = Mat(4, 3, [0, 2, 3, 5, 7], [0, 2, 1, 0, 2, 1, 2], [1, 2, 3, 4, 5, 6, 7], rvsp.DTYPE_FP32)
    s = product(A, B, rvsp.BACKEND_SCALAR, rvsp.DTYPE_FP32, rvsp.DTYPE_FP32)
    u = product(A, B, rvsp.BACKEND_SCALAR_UNROLL4, rvsp.DTYPE_FP32, rvsp.DTYPE_FP32)
    assert s[0] == u[0] == (2, 3)
    for x, y in zip(s[1:], u[1:]):
        np.testing.assert_array_equal(x, y)
    assert s[1].tolist() == [0, 3, 6] and s[2].tolist() == [0, 1, 2, 0, 1, 2]
    assert s[3].tolist() == [1.0, 30.0, 30.0, 20.0, 9.0, 25.0]


def test_int8_to_int32():
    """test_spgemm_csr_i8.c: [[1,2],[0,-3]] x [[4,0],[0,5]] = [[4,10],[0,-15]]."""
    A = Mat(2, 2, [0, 2, 3], [0, 1, 1], [1, 2, -3], rvsp.DTYPE_INT8)
    B = Mat(2, 2, [0, 1, 2], [0, 1], [4, 5], rvsp.DTYPE_INT8)
    shape, rp, ci, va = product(A, B, rvsp.BACKEND_SCALAR, rvsp.DTYPE_INT8, rvsp.DTYPE_INT32)
    assert shape == (2, 2)
    assert rp.tolist() == [0, 2, 3] and ci.tolist() == [0, 1, 1]
    assert va.tolist() == [4, 10, -15]
