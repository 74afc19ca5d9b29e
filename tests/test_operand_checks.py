"""Operands that are not canonical CSR never reach the product kernels.

rv-sparse checks canonical CSR in its one-shot because its kernels assume it
(src/core/spgemm.c:41-51). Here every entry point checks before its kernels
read through the operands' indices:
  * host operands: rvsp_csr_check on the host, in the one-shot and in
    work_estimation (so also in the raw rvsp_spgemm_hip_f32), before any
    upload or launch (CPU tests below: no device is needed to be refused);
  * device operands (device descriptor, device one-shot): two bounded
    streaming kernels check row pointers, column range and strict order on
    the device before the analysis (GPU tests);
  * batched operands: each problem's row pointer and column range are
    checked while it is stacked, the stacked product's order by the
    analysis (GPU tests).
Each malformed operand below has arrays of exactly its declared sizes
(rows + 1 row pointers, nnz entries), the only reads the checks make.
"""
import ctypes as C

import numpy as np
import pytest

import rvsp


def good(n=64, per_row=6, seed=3):
    return rvsp.gen_random_rows(n, n, per_row, per_row, seed)


def broken(M, kind):
    """A copy of M (square, every row with >= 2 entries) with one defect."""
    rp, ci = M.row_ptr.copy(), M.col_idx.copy()
    s = int(rp[5])
    if kind == "col_too_big":
        ci[s] = M.cols
    elif kind == "col_negative":
        ci[s] = -1
    elif kind == "unsorted":
        ci[s], ci[s + 1] = ci[s + 1], ci[s]
    elif kind == "duplicate":
        ci[s + 1] = ci[s]
    elif kind == "row_ptr_decreasing":
        rp[6] = rp[5] - 1
    elif kind == "row_ptr_first":
        rp[0] = 1
    elif kind == "row_ptr_last":
        rp[-1] = rp[-1] - 1
    elif kind == "row_ptr_beyond":
        rp[3] = M.nnz + 100
    else:
        raise ValueError(kind)
    return rvsp.Csr(M.rows, M.cols, rp, ci, M.values.copy())


KINDS = ["col_too_big", "col_negative", "unsorted", "duplicate", "row_ptr_decreasing",
         "row_ptr_first", "row_ptr_last", "row_ptr_beyond"]


def struct_nnz(M, nnz):
    """CsrStruct over M's host arrays with the declared nnz = len(col_idx)."""
    return rvsp.CsrStruct(M.rows, M.cols, nnz, M.row_ptr.ctypes.data, M.col_idx.ctypes.data,
                          M.values.ctypes.data, rvsp.DTYPE_FP32, rvsp.FORMAT_CSR, 0)


# ---------------------------------------------------------------------------
# host operands: refused before any device work (CPU)
# ---------------------------------------------------------------------------
# rvsp_csr_check, like rv-sparse's, does not compare row_ptr[rows] with the
# declared nnz; a smaller last pointer only leaves trailing entries unused
HOST_KINDS = [k for k in KINDS if k != "row_ptr_last"]


@pytest.mark.parametrize("kind", HOST_KINDS)
@pytest.mark.parametrize("side", ["A", "B"])
def test_host_work_estimation_refuses_non_canonical(kind, side):
    M = good()
    X = broken(M, kind)
    a = struct_nnz(X if side == "A" else M, M.nnz)
    b = struct_nnz(X if side == "B" else M, M.nnz)
    d = rvsp.Descr()
    st, _, _ = d.estimate(a, b)
    assert st == rvsp.ERROR_INVALID_CSR
    d.close()


@pytest.mark.parametrize("kind", ["col_too_big", "unsorted", "duplicate", "row_ptr_decreasing"])
def test_raw_oneshot_refuses_non_canonical(kind):
    M = good()
    X = broken(M, kind)
    with pytest.raises(rvsp.RvspError) as e:
        rvsp.spgemm_raw(X, M)
    assert e.value.status == rvsp.ERROR_INVALID_CSR


@pytest.mark.parametrize("kind", ["col_too_big", "unsorted", "duplicate"])
def test_host_oneshot_refuses_non_canonical(kind):
    M = good()
    st, _ = rvsp.spgemm_csr(broken(M, kind), M, raise_on_error=False)
    assert st == rvsp.ERROR_INVALID_CSR


# ---------------------------------------------------------------------------
# device operands (GPU)
# ---------------------------------------------------------------------------
def device(M, nnz):
    """DeviceCsr of M with exactly rows + 1 row pointers and nnz entries."""
    d = rvsp.DeviceCsr.__new__(rvsp.DeviceCsr)
    d.rows, d.cols, d.nnz = M.rows, M.cols, nnz
    d.rp = rvsp.DeviceBuffer.from_array(M.row_ptr)
    d.ci = rvsp.DeviceBuffer.from_array(M.col_idx[:nnz])
    d.va = rvsp.DeviceBuffer.from_array(M.values[:nnz])
    return d


def oneshot_status(dA, dB):
    L = rvsp.lib()
    a, b = dA.struct(), dB.struct()
    c = rvsp.CsrStruct()
    st = L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c), rvsp.ALGO_HIP)
    if st == 0:
        L.rvsp_csr_destroy(C.byref(c))
    return st


@pytest.mark.gpu
@pytest.mark.parametrize("kind", KINDS)
def test_device_operands_refused_before_the_analysis(kind):
    print(f"device operand check: {kind}", flush=True)
    M = good()
    X = broken(M, kind)
    dM, dX = device(M, M.nnz), device(X, M.nnz)
    assert oneshot_status(dX, dM) == rvsp.ERROR_INVALID_CSR  # bad A
    assert oneshot_status(dM, dX) == rvsp.ERROR_INVALID_CSR  # bad B
    assert oneshot_status(dX, dX) == rvsp.ERROR_INVALID_CSR  # C = X * X, one matrix
    # the library is still usable, and a good product is still exact
    got = rvsp.spgemm_device(dM, dM, algo=rvsp.ALGO_HIP)
    want = rvsp.spgemm_raw(M, M)
    np.testing.assert_array_equal(got.col_idx, want.col_idx)
    np.testing.assert_array_equal(got.values.view(np.uint32), want.values.view(np.uint32))


@pytest.mark.gpu
def test_device_descriptor_estimation_refuses_and_compute_checks_pointers():
    M = good()
    dM = device(M, M.nnz)
    dX = device(broken(M, "duplicate"), M.nnz)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    st, _, _ = d.estimate(dX.struct(), dM.struct())
    assert st == rvsp.ERROR_INVALID_CSR
    d.close()
    # compute with other structure arrays than the estimation's: refused
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    a = dM.struct()
    st, ws, nnz = d.estimate(a, a)
    assert st == 0
    crp = rvsp.DeviceBuffer(4 * (M.rows + 1))
    cci = rvsp.DeviceBuffer(4 * nnz)
    cva = rvsp.DeviceBuffer(4 * nnz)
    c = rvsp.CsrStruct(M.rows, M.cols, nnz, crp.ptr, cci.ptr, cva.ptr, rvsp.DTYPE_FP32, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    copy = device(M, M.nnz)  # same structure, other arrays
    assert d.compute(copy.struct(), a, c, wsb.ctypes.data) == rvsp.ERROR_INVALID_ARGUMENT
    assert d.compute(a, copy.struct(), c, wsb.ctypes.data) == rvsp.ERROR_INVALID_ARGUMENT
    # new values through the estimated arrays are fine
    assert d.compute(a, a, c, wsb.ctypes.data) == 0
    assert d.wait() == 0
    want = rvsp.spgemm_raw(M, M)
    np.testing.assert_array_equal(cva.to_array(np.float32, nnz).view(np.uint32),
                                  want.values.view(np.uint32))
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["col_too_big", "unsorted", "duplicate", "row_ptr_decreasing",
                                  "row_ptr_last"])
def test_batched_refuses_a_bad_problem(kind):
    """Problem 1 of 3 is broken. A column one past its own B's rows would
    still fall inside the stacked B (it is the next problem's first row):
    the stacking checks it against the problem's own column count."""
    print(f"batched operand check: {kind}", flush=True)
    Ms = [good(seed=s) for s in (11, 12, 13)]
    dA = [device(M, M.nnz) for M in Ms]
    dB = [device(M, M.nnz) for M in Ms]
    dA[1] = device(broken(Ms[1], kind), Ms[1].nnz)
    with pytest.raises(rvsp.RvspError) as e:
        rvsp.spgemm_batched(dA, dB)
    assert e.value.status == rvsp.ERROR_INVALID_CSR
    # and the same batch without the defect is exact
    dA[1] = device(Ms[1], Ms[1].nnz)
    got, _ = rvsp.spgemm_batched(dA, dB)
    for M, G in zip(Ms, got):
        want = rvsp.spgemm_raw(M, M)
        np.testing.assert_array_equal(G.values.view(np.uint32), want.values.view(np.uint32))
