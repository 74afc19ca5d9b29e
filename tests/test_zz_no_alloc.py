"""The allocation-free device one-shot, rvsp_spgemm_csr_device_noalloc,
through the C ABI against the allocating one (rvsp_spgemm_csr_device_algo
with ALGO_HIP, which it shares its kernels with): same C bit for bit, on two
calls in a row, the second with a smaller input, so a stale part of the
shared scratch buffer would show. Also: an output too small reports the size
it needs, a product too large for the fixed buffer reports
RVSP_ERROR_SCRATCH_EXCEEDED, and non-canonical operands are still refused.
"""
import ctypes as C

import numpy as np
import pytest

import rvsp
from rvsp import workloads as W

pytestmark = pytest.mark.gpu


def reference(dA, dB):
    """C from the allocating device one-shot (host copy)."""
    return rvsp.spgemm_device(dA, dB, algo=rvsp.ALGO_HIP)


def noalloc(dA, dB, cap, stream=None):
    """(status, C struct, host copy or None) of the allocation-free one-shot
    into fresh device outputs with room for `cap` entries."""
    L = rvsp.lib()
    rp = rvsp.DeviceBuffer(4 * (dA.rows + 1))
    ci = rvsp.DeviceBuffer(4 * max(cap, 1))
    va = rvsp.DeviceBuffer(4 * max(cap, 1))
    c = rvsp.CsrStruct(0, 0, cap, rp.ptr, ci.ptr, va.ptr, rvsp.DTYPE_FP32, 0, 0)
    a, b = dA.struct(), dB.struct()
    st = L.rvsp_spgemm_csr_device_noalloc(C.byref(a), C.byref(b), C.byref(c), stream)
    host = rvsp.device_csr_to_host(c) if st == 0 else None
    return st, c, host, (rp, ci, va)


def same(got, want, what):
    np.testing.assert_array_equal(got.row_ptr, want.row_ptr, err_msg=f"{what}: row_ptr")
    np.testing.assert_array_equal(got.col_idx, want.col_idx, err_msg=f"{what}: col_idx")
    np.testing.assert_array_equal(got.values.view(np.uint32), want.values.view(np.uint32),
                                  err_msg=f"{what}: values")


def test_status_string():
    assert rvsp.status_string(rvsp.ERROR_SCRATCH_EXCEEDED) == "scratch buffer too small"


@pytest.mark.parametrize("big,small", [("rmat14", "uniform"), ("mixed", "stencil")])
def test_noalloc_matches_allocating_oneshot_twice(big, small):
    """Two calls in a row, the second smaller; both bit-identical to the
    allocating entry point. The first input has big rows (bitmaps, windows,
    cursors in the scratch buffer), the second reuses the buffer from 0."""
    inputs = {
        "rmat14": lambda: W.rmat(scale=14, seed=31),
        "uniform": lambda: W.uniform(n=1024, per_row=16, seed=32),
        "mixed": lambda: W.mixed(m=1 << 14, k=1 << 10, n=1 << 14, maxlen=1024, seed_a=33, seed_b=34),
        "stencil": lambda: W.stencil(g=12, seed=35),
    }
    for name in (big, small):
        print(f"noalloc {name}", flush=True)
        A, B = inputs[name]()
        dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
        want = reference(dA, dB)
        st, c, got, _ = noalloc(dA, dB, want.nnz)
        assert st == 0, rvsp.status_string(st)
        assert (c.rows, c.cols, c.nnz, c.owns_data) == (A.rows, B.cols, want.nnz, 0)
        same(got, want, name)
        # and a second call on the same input into fresh outputs
        st, c, again, _ = noalloc(dA, dB, want.nnz + 17)
        assert st == 0 and c.nnz == want.nnz
        same(again, want, f"{name} again")


def test_noalloc_on_a_caller_stream():
    A, B = W.uniform(n=512, per_row=8, seed=36)
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    want = reference(dA, dB)
    d = rvsp.Descr()
    stream = d.stream()  # a stream the library created for a descriptor
    st, _, got, _ = noalloc(dA, dB, want.nnz, stream=stream)
    assert st == 0
    same(got, want, "caller stream")
    d.close()


def test_noalloc_reports_capacity_and_scratch_limits():
    A, B = W.uniform(n=1024, per_row=16, seed=37)
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    want = reference(dA, dB)
    st, c, _, _ = noalloc(dA, dB, want.nnz - 1)
    assert st == rvsp.ERROR_INVALID_ARGUMENT
    assert c.nnz == want.nnz  # the size the caller needs
    # R-MAT 2^18: its big-row bitmaps alone exceed the fixed buffer
    R, _ = W.rmat()
    dR = rvsp.DeviceCsr(R)
    st, _, _, _ = noalloc(dR, dR, 1)
    assert st == rvsp.ERROR_SCRATCH_EXCEEDED
    # still usable afterwards
    st, _, got, _ = noalloc(dA, dB, want.nnz)
    assert st == 0
    same(got, want, "after a refused call")


def test_noalloc_refuses_non_canonical_operands():
    A, B = W.uniform(n=256, per_row=8, seed=38)
    ci = A.col_idx.copy()
    ci[1] = ci[0]  # a repeated column in row 0
    X = rvsp.Csr(A.rows, A.cols, A.row_ptr, ci, A.values)
    st, _, _, _ = noalloc(rvsp.DeviceCsr(X), rvsp.DeviceCsr(B), 1 << 20)
    assert st == rvsp.ERROR_INVALID_CSR
