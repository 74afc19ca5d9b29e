"""Host-side contract of the C library, runnable without a GPU.

Mirrors the parts of rv-sparse tests/test_spgemm_descr_f32.c that do not
need the kernels (create/destroy, set_algo ordering, compute/op counts
before estimation, NULL handling), the matrix lifecycle and canonical
check of src/core/matrix.c and rvsp_support.c, the dispatch rules of
src/core/spgemm.c, and that the library exports every entry point its
headers declare. Without a GPU every call that needs one must fail loudly
(RVSP_ERROR_UNSUPPORTED_BACKEND), never fall back to a CPU path.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import rvsp
from helpers import csr_of, load_fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FX = load_fixtures()


def declared_functions():
    names = set()
    for h in ("rv_sparse.h", "rv_sparse_hip.h", "rvsp_spgemm.h", "rvsp_hip_abi.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b(rvsp_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = rvsp.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) > 40


def test_version_and_status_strings():
    L = rvsp.lib()
    assert L.rvsp_get_version() == b"0.1.0"
    assert L.rvsp_status_to_string(0) == b"success"
    assert L.rvsp_status_to_string(-6) == b"allocation failed"
    assert L.rvsp_status_to_string(-99) == b"unknown error"


def test_struct_layout_matches_reference_abi(tmp_path):
    """The binding's struct equals what a C compiler makes of the header."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "rv_sparse.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(rvsp_csr_matrix_t),'
        'offsetof(rvsp_csr_matrix_t,row_ptr), offsetof(rvsp_csr_matrix_t,values),'
        'offsetof(rvsp_csr_matrix_t,owns_data), sizeof(rvsp_spgemm_options_t));return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(rvsp.CsrStruct), rvsp.CsrStruct.row_ptr.offset, rvsp.CsrStruct.values.offset,
            rvsp.CsrStruct.owns_data.offset, C.sizeof(rvsp.Options)]
    assert got == want


def test_csr_create_validate_destroy():
    L = rvsp.lib()
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 1, 1], np.int32)
    va = np.array([1, 2, 3], np.float32)
    A = rvsp.CsrStruct()
    assert L.rvsp_csr_create(C.byref(A), 2, 2, 3, rp.ctypes.data, ci.ctypes.data, va.ctypes.data,
                             rvsp.DTYPE_FP32) == 0
    assert A.owns_data == 0 and A.format == rvsp.FORMAT_CSR
    assert L.rvsp_csr_create(None, 2, 2, 3, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3) == -1
    assert L.rvsp_csr_create(C.byref(A), 0, 2, 3, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3) == -2
    bad = np.array([0, 2, 4], np.int32)  # row_ptr[rows] != nnz
    assert L.rvsp_csr_create(C.byref(A), 2, 2, 3, bad.ctypes.data, ci.ctypes.data, va.ctypes.data, 3) == -3
    badc = np.array([0, 1, 2], np.int32)  # column 2 out of range for cols=2
    assert L.rvsp_csr_create(C.byref(A), 2, 2, 3, rp.ctypes.data, badc.ctypes.data, va.ctypes.data, 3) == -3
    # validate does not check order (rv-sparse docs/api_design.md): unsorted passes
    uns = np.array([1, 0, 1], np.int32)
    assert L.rvsp_csr_create(C.byref(A), 2, 2, 3, rp.ctypes.data, uns.ctypes.data, va.ctypes.data, 3) == 0
    L.rvsp_csr_destroy(C.byref(A))  # borrowed arrays are not freed
    assert A.rows == 0 and A.row_ptr is None and A.owns_data == 0
    assert rp.tolist() == [0, 2, 3]
    L.rvsp_csr_destroy(None)


def test_csr_check_reasons_and_rows():
    L = rvsp.lib()

    def chk(rows, cols, rp, ci):
        rp = np.asarray(rp, np.int32)
        ci = np.asarray(ci if len(ci) else [0], np.int32)
        bad = C.c_int32(99)
        rc = L.rvsp_csr_check(rows, cols, rp.ctypes.data, ci.ctypes.data, C.byref(bad))
        return rc, bad.value

    assert chk(2, 3, [0, 1, 2], [0, 2]) == (rvsp.CSR_OK, -1)
    assert chk(2, 3, [1, 1, 2], [0, 2]) == (rvsp.CSR_BAD_ROW_PTR, -1)
    assert chk(2, 3, [0, 2, 1], [0, 2]) == (rvsp.CSR_BAD_ROW_PTR, 1)
    assert chk(2, 3, [0, 1, 2], [0, 3]) == (rvsp.CSR_COL_OUT_OF_RANGE, 1)
    assert chk(1, 3, [0, 2], [2, 1]) == (rvsp.CSR_UNSORTED, 0)
    assert chk(1, 3, [0, 2], [1, 1]) == (rvsp.CSR_DUPLICATE, 0)
    assert L.rvsp_csr_status_string(rvsp.CSR_DUPLICATE) == b"duplicate column in row"


def test_buffer_size():
    L = rvsp.lib()
    n = C.c_size_t(0)
    assert L.rvsp_spgemm_buffer_size(10, C.byref(n)) == 0 and n.value > 0
    assert L.rvsp_spgemm_buffer_size(0, C.byref(n)) == -2
    assert L.rvsp_spgemm_buffer_size(10, None) == -1


def test_descriptor_contract_without_gpu_work():
    """test_spgemm_descr_f32.c:74-103: lifecycle and ordering rules."""
    L = rvsp.lib()
    pair = FX["pairs"][0]
    A, B = csr_of(pair["a"]), csr_of(pair["b"])
    a, b = A.struct(), B.struct()
    h = C.c_void_p()
    assert L.rvsp_spgemm_descr_create(C.byref(h)) == 0 and h.value
    assert L.rvsp_spgemm_descr_create(None) == rvsp.ERROR_NULL_POINTER
    assert L.rvsp_spgemm_set_algo(h, rvsp.ALGO_DEFAULT) == 0
    assert L.rvsp_spgemm_set_algo(None, rvsp.ALGO_DEFAULT) == rvsp.ERROR_NULL_POINTER
    for algo in (rvsp.ALGO_RVV, rvsp.ALGO_CONTIG, rvsp.ALGO_ADAPTIVE):
        assert L.rvsp_spgemm_set_algo(h, algo) == rvsp.ERROR_UNSUPPORTED_BACKEND
    assert L.rvsp_spgemm_set_algo(h, rvsp.ALGO_HIP) == 0
    assert L.rvsp_spgemm_set_algo(h, rvsp.ALGO_HIP_LEARNED) == 0
    rp = np.zeros(3, np.int32)
    ci = np.zeros(3, np.int32)
    va = np.zeros(3, np.float32)
    c = rvsp.CsrStruct(2, 2, 3, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    ws = np.zeros(4096, np.uint8)
    assert L.rvsp_spgemm_compute(h, C.byref(a), C.byref(b), C.byref(c), ws.ctypes.data) == -2
    ops = np.zeros(2, np.int64)
    assert L.rvsp_spgemm_get_op_counts(h, ops.ctypes.data, 2) == -2
    assert L.rvsp_spgemm_compute(h, C.byref(a), C.byref(b), C.byref(c), None) == -1
    assert L.rvsp_spgemm_compute(h, C.byref(a), C.byref(b), None, ws.ctypes.data) == -1
    wsb = C.c_size_t(0)
    nnz = C.c_int32(0)
    assert L.rvsp_spgemm_work_estimation(h, None, C.byref(b), C.byref(wsb), C.byref(nnz)) == -1
    assert L.rvsp_spgemm_work_estimation(h, C.byref(a), C.byref(b), None, C.byref(nnz)) == -1
    # dtype and shape checks come before any device work
    a64 = A.struct()
    a64.dtype = rvsp.DTYPE_FP64
    assert L.rvsp_spgemm_work_estimation(h, C.byref(a64), C.byref(b), C.byref(wsb), C.byref(nnz)) == -4
    bt = B.struct()
    bt.rows = 5
    assert L.rvsp_spgemm_work_estimation(h, C.byref(a), C.byref(bt), C.byref(wsb), C.byref(nnz)) == -2
    L.rvsp_spgemm_invalidate_structure(None)
    L.rvsp_spgemm_descr_destroy(None)
    L.rvsp_spgemm_descr_destroy(h)


def test_memory_mode_setters_before_estimation():
    L = rvsp.lib()
    h = C.c_void_p()
    assert L.rvsp_spgemm_descr_create(C.byref(h)) == 0
    assert L.rvsp_spgemm_set_memory(h, rvsp.MEMORY_DEVICE) == 0
    assert L.rvsp_spgemm_set_memory(h, 7) == -2
    assert L.rvsp_spgemm_set_graph(h, 1) == 0
    assert L.rvsp_spgemm_set_memory(None, 0) == -1
    v = C.c_int64(0)
    assert L.rvsp_spgemm_get_products(h, C.byref(v)) == -2
    L.rvsp_spgemm_descr_destroy(h)


def test_oneshot_dispatch_rules():
    """src/core/spgemm.c:135-211: NULLs, legacy dtypes and RVV."""
    L = rvsp.lib()
    pair = FX["pairs"][0]
    A, B = csr_of(pair["a"]), csr_of(pair["b"])
    a, b = A.struct(), B.struct()
    c = rvsp.CsrStruct()
    assert L.rvsp_spgemm_csr(None, C.byref(b), C.byref(c), None) == -1
    # legacy dtypes over FP32 matrices fail the legacy wrappers' dtype checks
    # (csr_spgemm_wrappers.c:362 INVALID_ARGUMENT, :43 UNSUPPORTED_DTYPE)
    for be, ind, outd, want in ((0, rvsp.DTYPE_FP64, rvsp.DTYPE_FP64, rvsp.ERROR_INVALID_ARGUMENT),
                                (0, rvsp.DTYPE_INT8, rvsp.DTYPE_INT32, rvsp.ERROR_UNSUPPORTED_DTYPE)):
        opt = rvsp.Options(be, ind, outd)
        assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt)) == want
    # a legacy dtype pair the legacy dispatch does not know
    opt = rvsp.Options(0, rvsp.DTYPE_FP64, rvsp.DTYPE_INT32)
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt)) == -5
    # RVV reaches set_algo after the input checks, then is unsupported
    opt = rvsp.Options(rvsp.BACKEND_RVV_INTRINSICS, 3, 3)
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt)) == -5
    # canonical check precedes any device work
    bad = rvsp.Csr.make(2, 2, [0, 2, 3], [1, 0, 1], [1, 2, 3])
    assert L.rvsp_spgemm_csr(C.byref(bad.struct()), C.byref(b), C.byref(c), None) == -3
    assert c.row_ptr is None and c.owns_data == 0


@pytest.mark.skipif(rvsp.gpu_available(), reason="checks the no-GPU behaviour")
def test_no_gpu_fails_loudly():
    pair = FX["pairs"][0]
    A, B = csr_of(pair["a"]), csr_of(pair["b"])
    st, C_ = rvsp.spgemm_csr(A, B, raise_on_error=False)
    assert st == rvsp.ERROR_UNSUPPORTED_BACKEND and C_ is None
    d = rvsp.Descr()
    st, _, _ = d.estimate(A.struct(), B.struct())
    assert st == rvsp.ERROR_UNSUPPORTED_BACKEND
    d.close()
