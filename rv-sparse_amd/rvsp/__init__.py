"""ctypes binding of the MI355X rv_sparse C library (tests, bench, scripts).

The product is the C-ABI shared library lib/librvsparse_hip.so (rv_sparse.h,
rv_sparse_hip.h, rvsp_spgemm.h). This module only marshals numpy arrays in
and out of it; every product runs inside that library on the GPU. Nothing
here computes a SpGEMM, and a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
# RVSP_LIB names another build of the library (e.g. a variant being timed)
LIB_PATH = os.environ.get("RVSP_LIB") or os.path.join(LIB_DIR, "librvsparse_hip.so")
GEN_PATH = os.path.join(LIB_DIR, "librvsp_gen.so")

# rv_sparse_types.h
SUCCESS = 0
ERROR_NULL_POINTER = -1
ERROR_INVALID_ARGUMENT = -2
ERROR_INVALID_CSR = -3
ERROR_UNSUPPORTED_DTYPE = -4
ERROR_UNSUPPORTED_BACKEND = -5
ERROR_ALLOCATION_FAILED = -6
ERROR_SCRATCH_EXCEEDED = -7

DTYPE_INT8, DTYPE_INT32, DTYPE_BF16, DTYPE_FP32, DTYPE_FP64 = range(5)
FORMAT_CSR = 0
BACKEND_SCALAR, BACKEND_SCALAR_UNROLL4, BACKEND_RVV_INTRINSICS, BACKEND_HIP = range(4)
ALGO_DEFAULT, ALGO_RVV, ALGO_CONTIG, ALGO_ADAPTIVE, ALGO_HIP, ALGO_HIP_LEARNED = range(6)
MEMORY_HOST, MEMORY_DEVICE = 0, 1
OWNS_DEVICE = 2

CSR_OK, CSR_BAD_ROW_PTR, CSR_COL_OUT_OF_RANGE, CSR_UNSORTED, CSR_DUPLICATE = range(5)


class RvspError(RuntimeError):
    def __init__(self, where: str, status: int):
        super().__init__(f"{where}: status {status} ({status_string(status)})")
        self.status = status


class CsrStruct(C.Structure):
    """rvsp_csr_matrix_t, field for field."""

    _fields_ = [
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("nnz", C.c_int32),
        ("row_ptr", C.c_void_p),
        ("col_idx", C.c_void_p),
        ("values", C.c_void_p),
        ("dtype", C.c_int),
        ("format", C.c_int),
        ("owns_data", C.c_int),
    ]


class Options(C.Structure):
    _fields_ = [("backend", C.c_int), ("input_dtype", C.c_int), ("output_dtype", C.c_int)]


class Tuning(C.Structure):
    """rvsp_tuning_t (rv_sparse_hip.h)."""
    _fields_ = [("big_window_outputs", C.c_int32), ("big_window_chunks", C.c_int32),
                ("sym_big_threads", C.c_int32), ("seg_products", C.c_int32),
                ("hash_max_u", C.c_int32), ("fan_min_products", C.c_int64),
                ("fan_min_fill", C.c_int64)]

    FIELDS = ("big_window_outputs", "big_window_chunks", "sym_big_threads", "seg_products",
              "hash_max_u", "fan_min_products", "fan_min_fill")

    def as_dict(self):
        return {f: int(getattr(self, f)) for f in self.FIELDS}

    @staticmethod
    def of(d) -> "Tuning":
        return Tuning(*[int(d[f]) for f in Tuning.FIELDS])


TUNING_CLASSES = 8


class GenmatParams(C.Structure):
    """genmat_params_t (include/genmat.h)."""
    _fields_ = [("row_cnt", C.c_int), ("col_cnt", C.c_int), ("density", C.c_double),
                ("cv", C.c_double), ("min", C.c_int), ("imbalance", C.c_double),
                ("is_symmetric", C.c_int), ("is_column", C.c_int), ("low_bandwidth", C.c_int),
                ("up_bandwidth", C.c_int), ("random_seed", C.c_int)]


class GenmatCsr(C.Structure):
    """csr_matrix_t (include/genmat.h): int indices, double values."""
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int), ("nnz", C.c_longlong),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("values", C.c_void_p)]


_lib = None
_gen = None


def _declare(lib):
    P, I32, I64, SZ, V = C.POINTER, C.c_int32, C.c_int64, C.c_size_t, C.c_void_p
    CSRP = P(CsrStruct)
    sig = {
        "rvsp_get_version": (C.c_char_p, []),
        "rvsp_status_to_string": (C.c_char_p, [C.c_int]),
        "rvsp_csr_create": (C.c_int, [CSRP, I32, I32, I32, V, V, V, C.c_int]),
        "rvsp_csr_validate": (C.c_int, [CSRP]),
        "rvsp_csr_destroy": (None, [CSRP]),
        "rvsp_mtx_load": (C.c_int, [C.c_char_p, CSRP]),
        "genmat_default_params": (GenmatParams, [C.c_int, C.c_int]),
        "genmat_generate_csr": (GenmatCsr, [GenmatParams]),
        "genmat_free_csr": (None, [P(GenmatCsr)]),
        "genmat_save_csr_to_mtx": (C.c_int, [P(GenmatCsr), C.c_char_p]),
        "rvsp_genmat_bench_pair": (C.c_int, [I32, I32, C.c_double, I32, C.c_int, CSRP, CSRP]),
        "rvsp_spgemm_csr": (C.c_int, [CSRP, CSRP, CSRP, P(Options)]),
        "rvsp_spgemm_descr_create": (C.c_int, [P(V)]),
        "rvsp_spgemm_descr_destroy": (None, [V]),
        "rvsp_spgemm_set_algo": (C.c_int, [V, C.c_int]),
        "rvsp_spgemm_work_estimation": (C.c_int, [V, CSRP, CSRP, P(SZ), P(I32)]),
        "rvsp_spgemm_compute": (C.c_int, [V, CSRP, CSRP, CSRP, V]),
        "rvsp_spgemm_get_op_counts": (C.c_int, [V, V, I32]),
        "rvsp_spgemm_invalidate_structure": (None, [V]),
        "rvsp_spgemm_buffer_size": (C.c_int, [I32, P(SZ)]),
        "rvsp_csr_check": (C.c_int, [I32, I32, V, V, P(I32)]),
        "rvsp_csr_status_string": (C.c_char_p, [C.c_int]),
        "rvsp_spgemm_hip_f32": (
            C.c_int,
            [I32, I32, I32, V, V, V, V, V, V, V, SZ, P(V), P(V), P(V), P(I32), V],
        ),
        "rvsp_hip_available": (C.c_int, [P(C.c_int)]),
        "rvsp_hip_malloc": (C.c_int, [P(V), SZ]),
        "rvsp_hip_free": (C.c_int, [V]),
        "rvsp_hip_memcpy_h2d": (C.c_int, [V, V, SZ]),
        "rvsp_hip_memcpy_d2h": (C.c_int, [V, V, SZ]),
        "rvsp_hip_memcpy_d2d": (C.c_int, [V, V, SZ]),
        "rvsp_hip_memset": (C.c_int, [V, C.c_int, SZ]),
        "rvsp_hip_synchronize": (C.c_int, []),
        "rvsp_hip_event_create": (C.c_int, [P(V)]),
        "rvsp_hip_event_record": (C.c_int, [V, V]),
        "rvsp_hip_event_synchronize": (C.c_int, [V]),
        "rvsp_hip_event_elapsed_ms": (C.c_int, [V, V, P(C.c_float)]),
        "rvsp_hip_event_destroy": (None, [V]),
        "rvsp_spgemm_set_memory": (C.c_int, [V, C.c_int]),
        "rvsp_spgemm_set_stream": (C.c_int, [V, V]),
        "rvsp_spgemm_get_stream": (C.c_int, [V, P(V)]),
        "rvsp_spgemm_set_graph": (C.c_int, [V, C.c_int]),
        "rvsp_spgemm_wait": (C.c_int, [V]),
        "rvsp_spgemm_get_products": (C.c_int, [V, P(I64)]),
        "rvsp_spgemm_csr_device": (C.c_int, [CSRP, CSRP, CSRP, P(Options)]),
        "rvsp_spgemm_csr_device_algo": (C.c_int, [CSRP, CSRP, CSRP, C.c_int]),
        "rvsp_spgemm_csr_device_noalloc": (C.c_int, [CSRP, CSRP, CSRP, V]),
        "rvsp_hip_fail_next_allocs": (None, [C.c_int]),
        "rvsp_hip_set_timing": (None, [C.c_int]),
        "rvsp_hip_get_timing": (C.c_int, [P(C.c_double), P(C.c_longlong), C.c_int]),
        "rvsp_hip_reset_timing": (None, []),
        "rvsp_hip_alloc_async": (C.c_int, [P(V), SZ, V]),
        "rvsp_hip_free_async": (C.c_int, [V, V]),
        "rvsp_hip_stream_sync": (C.c_int, [V]),
        "rvsp_hip_copy_async": (C.c_int, [V, V, SZ, V]),
        "rvsp_spgemm_batched_device": (C.c_int, [I32, CSRP, CSRP, CSRP, C.c_int, P(V)]),
        "rvsp_spgemm_batch_chunks": (I32, [V]),
        "rvsp_spgemm_batch_destroy": (None, [V]),
        # autotuned launch parameters (rv_sparse_hip.h)
        "rvsp_tuning_class": (I32, [I64, I64, I32, I32]),
        "rvsp_tuning_defaults": (C.c_int, [P(Tuning)]),
        "rvsp_tuning_get": (C.c_int, [I32, P(Tuning)]),
        "rvsp_tuning_set": (C.c_int, [I32, P(Tuning)]),
        "rvsp_tuning_force": (C.c_int, [P(Tuning)]),
        "rvsp_tuning_load_file": (C.c_int, [C.c_char_p]),
        # learned selectors (rv_sparse_hip.h)
        "rvsp_selector_param_count": (I32, [I32]),
        "rvsp_selector_set": (C.c_int, [I32, V, I32, C.c_float]),
        "rvsp_selector_get": (C.c_int, [I32, V, I32, P(C.c_float)]),
        "rvsp_selector_load_file": (C.c_int, [I32, C.c_char_p]),
        "rvsp_selector_clear": (C.c_int, [I32]),
        "rvsp_selector_predict": (C.c_int, [I32, V, I32, V, V]),
        "rvsp_selector_grad": (C.c_int, [I32, V, V, I32, V, V, P(C.c_float)]),
        "rvsp_selector_train": (C.c_int, [I32, V, V, I32, I32, I32, C.c_float, C.c_uint64, V,
                                          P(C.c_float)]),
        # device plan (rvsp_hip_abi.h)
        "rvsp_hip_plan_create": (C.c_int, [P(V), V]),
        "rvsp_hip_plan_destroy": (None, [V]),
        "rvsp_hip_plan_stream": (V, [V]),
        "rvsp_hip_plan_analyse": (C.c_int, [V, I32, I32, I32, I32, V, V, I32, V, V, C.c_int,
                                            P(I64), P(I64)]),
        "rvsp_hip_plan_row_ptr": (V, [V]),
        "rvsp_hip_plan_fill": (C.c_int, [V, V]),
        "rvsp_hip_plan_numeric": (C.c_int, [V, V, V, V, V]),
        "rvsp_hip_plan_fill_numeric": (C.c_int, [V, V, V, V, V]),
        "rvsp_hip_plan_check": (C.c_int, [V]),
        "rvsp_hip_plan_prepare_reuse": (C.c_int, [V, V]),
        "rvsp_hip_plan_get_bins": (C.c_int, [V, V, V]),
        "rvsp_hip_plan_set_num_bins": (C.c_int, [V, V]),
        "rvsp_hip_plan_features": (C.c_int, [V, V, V]),
        "rvsp_hip_plan_selector_stats": (C.c_int, [V, P(I64), P(I64)]),
        "rvsp_hip_plan_time_bin": (C.c_int, [V, I32, V, I32, V, V, V, V, I32,
                                             P(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with __graft_entry__.build()")
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def status_string(status: int) -> str:
    return lib().rvsp_status_to_string(status).decode()


def check(where: str, status: int):
    if status != SUCCESS:
        raise RvspError(where, status)


def gpu_available() -> bool:
    n = C.c_int(0)
    return lib().rvsp_hip_available(C.byref(n)) == SUCCESS


# ---------------------------------------------------------------------------
# host CSR
# ---------------------------------------------------------------------------
@dataclass
class Csr:
    """Host CSR: int32 row_ptr/col_idx, float32 values."""

    rows: int
    cols: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1]) if len(self.row_ptr) else 0

    def struct(self) -> CsrStruct:
        for a, dt in ((self.row_ptr, np.int32), (self.col_idx, np.int32), (self.values, np.float32)):
            assert a.dtype == dt and a.flags.c_contiguous
        return CsrStruct(
            self.rows,
            self.cols,
            self.nnz,
            self.row_ptr.ctypes.data,
            self.col_idx.ctypes.data,
            self.values.ctypes.data,
            DTYPE_FP32,
            FORMAT_CSR,
            0,
        )

    @staticmethod
    def make(rows, cols, row_ptr, col_idx, values) -> "Csr":
        return Csr(
            int(rows),
            int(cols),
            np.ascontiguousarray(row_ptr, dtype=np.int32),
            np.ascontiguousarray(col_idx, dtype=np.int32),
            np.ascontiguousarray(values, dtype=np.float32),
        )


def _from_owned_host(c: CsrStruct) -> Csr:
    """Copies a library-allocated host C into numpy and frees it."""
    n = c.nnz
    rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (c.rows + 1,)).copy()
    if n > 0:
        ci = np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (n,)).copy()
        va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(C.c_float)), (n,)).copy()
    else:
        ci = np.zeros(0, np.int32)
        va = np.zeros(0, np.float32)
    out = Csr(c.rows, c.cols, rp, ci, va)
    lib().rvsp_csr_destroy(C.byref(c))
    return out


def spgemm_csr(A: Csr, B: Csr, backend: int = BACKEND_HIP, raise_on_error=True):
    """rvsp_spgemm_csr (one-shot, host memory). Returns (status, Csr|None)."""
    a, b = A.struct(), B.struct()
    c = CsrStruct()
    opt = Options(backend, DTYPE_FP32, DTYPE_FP32)
    st = lib().rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt))
    if st != SUCCESS:
        if raise_on_error:
            raise RvspError("rvsp_spgemm_csr", st)
        return st, None
    return st, _from_owned_host(c)


def spgemm_raw(A: Csr, B: Csr, want_ops=False, timed=False):
    """rvsp_spgemm_hip_f32, the raw one-shot kernel with callee-malloc'd C;
    timed=True also returns the seconds of the library call alone."""
    L = lib()
    ws_bytes = C.c_size_t(0)
    check("buffer_size", L.rvsp_spgemm_buffer_size(B.cols, C.byref(ws_bytes)))
    ws = np.zeros(ws_bytes.value, np.uint8)
    rp, ci, va = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int32(0)
    ops = np.zeros(A.rows, np.int64) if want_ops else None
    t0 = time.perf_counter()
    st = L.rvsp_spgemm_hip_f32(
        A.rows, A.cols, B.cols,
        A.row_ptr.ctypes.data, A.col_idx.ctypes.data, A.values.ctypes.data,
        B.row_ptr.ctypes.data, B.col_idx.ctypes.data, B.values.ctypes.data,
        ws.ctypes.data, ws_bytes.value,
        C.byref(rp), C.byref(ci), C.byref(va), C.byref(nnz),
        ops.ctypes.data if want_ops else None,
    )
    dt = time.perf_counter() - t0
    check("rvsp_spgemm_hip_f32", st)
    c = CsrStruct(A.rows, B.cols, nnz.value, rp.value, ci.value, va.value, DTYPE_FP32, FORMAT_CSR, 1)
    out = _from_owned_host(c)
    res = (out,) + ((ops,) if want_ops else ()) + ((dt,) if timed else ())
    return res if len(res) > 1 else out


class Descr:
    """rvsp_spgemm_descr_t wrapper (host or device memory mode)."""

    def __init__(self):
        h = C.c_void_p()
        check("descr_create", lib().rvsp_spgemm_descr_create(C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().rvsp_spgemm_descr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_algo(self, algo) -> int:
        return lib().rvsp_spgemm_set_algo(self.h, algo)

    def set_memory(self, mem) -> int:
        return lib().rvsp_spgemm_set_memory(self.h, mem)

    def set_graph(self, on) -> int:
        return lib().rvsp_spgemm_set_graph(self.h, int(on))

    def stream(self):
        s = C.c_void_p()
        check("get_stream", lib().rvsp_spgemm_get_stream(self.h, C.byref(s)))
        return s.value

    def estimate(self, a: CsrStruct, b: CsrStruct):
        ws = C.c_size_t(0)
        nnz = C.c_int32(0)
        st = lib().rvsp_spgemm_work_estimation(self.h, C.byref(a), C.byref(b), C.byref(ws), C.byref(nnz))
        return st, ws.value, nnz.value

    def compute(self, a: CsrStruct, b: CsrStruct, c: CsrStruct, ws) -> int:
        return lib().rvsp_spgemm_compute(self.h, C.byref(a), C.byref(b), C.byref(c), ws)

    def op_counts(self, n):
        out = np.zeros(max(n, 1), np.int64)
        st = lib().rvsp_spgemm_get_op_counts(self.h, out.ctypes.data, n)
        return st, out[:n]

    def products(self) -> int:
        v = C.c_int64(0)
        check("get_products", lib().rvsp_spgemm_get_products(self.h, C.byref(v)))
        return v.value

    def wait(self) -> int:
        return lib().rvsp_spgemm_wait(self.h)

    def invalidate(self):
        lib().rvsp_spgemm_invalidate_structure(self.h)


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check("hip_malloc", lib().rvsp_hip_malloc(C.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @staticmethod
    def from_array(a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        d = DeviceBuffer(a.nbytes)
        check("h2d", lib().rvsp_hip_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes))
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check("h2d", lib().rvsp_hip_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def to_array(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype)
        if count:
            check("d2h", lib().rvsp_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().rvsp_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceCsr:
    """A CSR whose three arrays live in device memory."""

    def __init__(self, host: Csr):
        self.rows, self.cols, self.nnz = host.rows, host.cols, host.nnz
        self.rp = DeviceBuffer.from_array(host.row_ptr)
        self.ci = DeviceBuffer.from_array(host.col_idx)
        self.va = DeviceBuffer.from_array(host.values)

    def struct(self) -> CsrStruct:
        return CsrStruct(self.rows, self.cols, self.nnz, self.rp.ptr, self.ci.ptr, self.va.ptr,
                         DTYPE_FP32, FORMAT_CSR, 0)


def mtx_load(path: str, raise_on_error=True):
    """rvsp_mtx_load: a Matrix Market file as a host Csr (FP32, canonical),
    or (status, None) with raise_on_error=False on failure."""
    c = CsrStruct()
    st = lib().rvsp_mtx_load(str(path).encode(), C.byref(c))
    if st != SUCCESS:
        if raise_on_error:
            check("mtx_load", st)
        return st, None
    n = c.nnz

    def host(ptr, ct, count, dt):
        if not count:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()

    rp = host(c.row_ptr, C.c_int32, c.rows + 1, np.int32)
    ci = host(c.col_idx, C.c_int32, n, np.int32)
    va = host(c.values, C.c_float, n, np.float32)
    out = Csr(c.rows, c.cols, rp, ci, va)
    lib().rvsp_csr_destroy(C.byref(c))
    return out if raise_on_error else (SUCCESS, out)


def genmat_params(rows, cols, **kw) -> GenmatParams:
    """genmat_default_params(rows, cols) with the named fields replaced
    (density, cv, min, imbalance, is_symmetric, is_column, low_bandwidth,
    up_bandwidth, random_seed)."""
    p = lib().genmat_default_params(rows, cols)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"genmat_params_t has no field {k}")
        setattr(p, k, v)
    return p


def genmat_arrays(m: GenmatCsr):
    """(rows, cols, row_ptr int32, col_idx int32, values float64) of a csr_matrix_t."""
    n = int(m.nnz)

    def host(ptr, ct, count, dt):
        if not count:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()

    return (m.nrows, m.ncols, host(m.row_ptr, C.c_int32, m.nrows + 1, np.int32),
            host(m.col_idx, C.c_int32, n, np.int32), host(m.values, C.c_double, n, np.float64))


def genmat(rows, cols, **kw):
    """genmat_generate_csr: (rows, cols, row_ptr, col_idx, values float64), or
    None when the generator reports an error (nrows == -1)."""
    m = lib().genmat_generate_csr(genmat_params(rows, cols, **kw))
    if m.nrows < 0:
        return None
    out = genmat_arrays(m)
    lib().genmat_free_csr(C.byref(m))
    return out


def gen_bench_pair(rows, cols, density, seed):
    """rv-sparse's bench inputs for `--gen R C DENSITY SEED` in FP32: host
    Csr A (rows x cols, seed) and B (cols x rows, seed + 1)."""
    a, b = CsrStruct(), CsrStruct()
    check("genmat_bench_pair", lib().rvsp_genmat_bench_pair(rows, cols, density, seed, DTYPE_FP32,
                                                            C.byref(a), C.byref(b)))
    out = []
    for c in (a, b):
        n = c.nnz
        rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), shape=(c.rows + 1,)).copy()
        ci = (np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), shape=(n,)).copy()
              if n else np.zeros(0, np.int32))
        va = (np.ctypeslib.as_array(C.cast(c.values, C.POINTER(C.c_float)), shape=(n,)).copy()
              if n else np.zeros(0, np.float32))
        out.append(Csr(c.rows, c.cols, rp, ci, va))
        lib().rvsp_csr_destroy(C.byref(c))
    return out[0], out[1]


def device_csr_to_host(c: CsrStruct) -> Csr:
    L = lib()
    rp = np.empty(c.rows + 1, np.int32)
    check("d2h", L.rvsp_hip_memcpy_d2h(rp.ctypes.data, c.row_ptr, rp.nbytes))
    ci = np.empty(c.nnz, np.int32)
    va = np.empty(c.nnz, np.float32)
    if c.nnz:
        check("d2h", L.rvsp_hip_memcpy_d2h(ci.ctypes.data, c.col_idx, ci.nbytes))
        check("d2h", L.rvsp_hip_memcpy_d2h(va.ctypes.data, c.values, va.nbytes))
    return Csr(c.rows, c.cols, rp, ci, va)


def spgemm_device(dA: DeviceCsr, dB: DeviceCsr, algo=ALGO_DEFAULT, keep=False):
    """Device one-shot; returns the host copy (and the device struct if keep)."""
    a, b = dA.struct(), dB.struct()
    c = CsrStruct()
    check("csr_device", lib().rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c), algo))
    host = device_csr_to_host(c)
    if keep:
        return host, c
    lib().rvsp_csr_destroy(C.byref(c))
    return host


def tuning_class(A: "Csr", B: "Csr") -> int:
    return int(lib().rvsp_tuning_class(A.nnz, B.nnz, A.cols, B.cols))


def tuning_defaults() -> dict:
    t = Tuning()
    check("tuning_defaults", lib().rvsp_tuning_defaults(C.byref(t)))
    return t.as_dict()


def tuning_get(cls: int) -> dict:
    t = Tuning()
    check("tuning_get", lib().rvsp_tuning_get(cls, C.byref(t)))
    return t.as_dict()


def tuning_force(d):
    """Use the parameters d for every product until tuning_force(None)."""
    if d is None:
        check("tuning_force", lib().rvsp_tuning_force(None))
    else:
        t = Tuning.of(d)
        check("tuning_force", lib().rvsp_tuning_force(C.byref(t)))


PHASES = ("analyse", "fill", "numeric")


def set_timing(on: bool):
    lib().rvsp_hip_set_timing(int(on))


def reset_timing():
    lib().rvsp_hip_reset_timing()


def get_timing():
    """{phase: (total_ms, calls)} since the last reset."""
    ms = (C.c_double * 4)()
    calls = (C.c_longlong * 4)()
    check("get_timing", lib().rvsp_hip_get_timing(ms, calls, 4))
    return {name: (ms[i], calls[i]) for i, name in enumerate(PHASES)}


def spgemm_batched(dAs, dBs, algo=ALGO_DEFAULT, fetch=True, with_nnz=False):
    """rvsp_spgemm_batched_device over lists of DeviceCsr. Returns
    (list of host Csr or None, number of chunks[, nnz per problem]); the
    batch is released."""
    n = len(dAs)
    A = (CsrStruct * n)(*[d.struct() for d in dAs])
    B = (CsrStruct * n)(*[d.struct() for d in dBs])
    Cs = (CsrStruct * n)()
    h = C.c_void_p()
    check("batched", lib().rvsp_spgemm_batched_device(n, A, B, Cs, algo, C.byref(h)))
    chunks = lib().rvsp_spgemm_batch_chunks(h)
    out = [device_csr_to_host(Cs[i]) for i in range(n)] if fetch else None
    nnzs = [Cs[i].nnz for i in range(n)]
    lib().rvsp_spgemm_batch_destroy(h)
    return (out, chunks, nnzs) if with_nnz else (out, chunks)


# ---------------------------------------------------------------------------
# device plan: the analysis / fill / numeric stages one product at a time
# ---------------------------------------------------------------------------
SELECT_HEURISTIC, SELECT_LEARNED = 0, 1
SELECTOR_SIZER, SELECTOR_CHOOSER = 0, 1
NB_HASH0, NB_DENSE0, NB_RANK0, NB_WIN, NB_COUNT = 1, 29, 53, 61, 62
SB_EMPTY, SB_H64, SB_BMS, SB_BIG = 0, 1, 8, 9


class Plan:
    """rvsp_hip_plan_t over device operands (kept referenced), with C's
    structure and values in device buffers once filled."""

    def __init__(self, dA: DeviceCsr, dB: DeviceCsr, select=SELECT_HEURISTIC):
        h = C.c_void_p()
        check("plan_create", lib().rvsp_hip_plan_create(C.byref(h), None))
        self.h = h.value
        self.dA, self.dB = dA, dB
        self.analyse(select)

    def analyse(self, select=SELECT_HEURISTIC):
        nnz, prod = C.c_int64(0), C.c_int64(0)
        dA, dB = self.dA, self.dB
        check("plan_analyse", lib().rvsp_hip_plan_analyse(
            self.h, dA.rows, dA.cols, dB.cols, dA.nnz, dA.rp.ptr, dA.ci.ptr, dB.nnz, dB.rp.ptr,
            dB.ci.ptr, select, C.byref(nnz), C.byref(prod)))
        self.nnz, self.products = nnz.value, prod.value
        self.cci = DeviceBuffer(4 * max(self.nnz, 1))
        self.cval = DeviceBuffer(4 * max(self.nnz, 1))
        self.filled = False

    def fill(self):
        check("plan_fill", lib().rvsp_hip_plan_fill(self.h, self.cci.ptr))
        check("plan_check", lib().rvsp_hip_plan_check(self.h))
        self.filled = True

    def fill_numeric(self):
        """Structure and values in one call (big rows' columns written by
        the numeric pass)."""
        check("plan_fill_numeric", lib().rvsp_hip_plan_fill_numeric(
            self.h, self.dA.va.ptr, self.dB.va.ptr, self.cci.ptr, self.cval.ptr))
        check("plan_check", lib().rvsp_hip_plan_check(self.h))
        self.filled = True

    def numeric(self, sync=True):
        if not self.filled:
            self.fill()
        check("plan_numeric", lib().rvsp_hip_plan_numeric(
            self.h, self.dA.va.ptr, self.dB.va.ptr, self.cci.ptr, self.cval.ptr))
        if sync:
            check("plan_check", lib().rvsp_hip_plan_check(self.h))

    def prepare_reuse(self):
        """Per-product output ranks for later numeric passes (once per analysis)."""
        if not self.filled:
            self.fill()
        check("plan_prepare_reuse", lib().rvsp_hip_plan_prepare_reuse(self.h, self.cci.ptr))

    def stream(self):
        return lib().rvsp_hip_plan_stream(self.h)

    def row_ptr(self) -> np.ndarray:
        rp = np.empty(self.dA.rows + 1, np.int32)
        check("d2h", lib().rvsp_hip_memcpy_d2h(rp.ctypes.data, lib().rvsp_hip_plan_row_ptr(self.h),
                                                rp.nbytes))
        return rp

    def result(self) -> Csr:
        return Csr(self.dA.rows, self.dB.cols, self.row_ptr(),
                   self.cci.to_array(np.int32, self.nnz), self.cval.to_array(np.float32, self.nnz))

    def bins(self):
        m = self.dA.rows
        sb, nb = np.empty(m, np.uint8), np.empty(m, np.uint8)
        check("plan_get_bins", lib().rvsp_hip_plan_get_bins(self.h, sb.ctypes.data, nb.ctypes.data))
        return sb, nb

    def set_num_bins(self, nb: np.ndarray) -> int:
        nb = np.ascontiguousarray(nb, np.uint8)
        return lib().rvsp_hip_plan_set_num_bins(self.h, nb.ctypes.data)

    def features(self):
        """(device feats m x 5, device sizer target m) as DeviceBuffers."""
        m = self.dA.rows
        f, r = DeviceBuffer(20 * m), DeviceBuffer(4 * m)
        check("plan_features", lib().rvsp_hip_plan_features(self.h, f.ptr, r.ptr))
        return f, r

    def selector_stats(self):
        a, b = C.c_int64(0), C.c_int64(0)
        check("selector_stats", lib().rvsp_hip_plan_selector_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_bin(self, b: int, rows_dev: DeviceBuffer, nrows: int, reps=3) -> float:
        if not self.filled:
            self.fill()
        ms = C.c_float(0)
        st = lib().rvsp_hip_plan_time_bin(self.h, int(b), rows_dev.ptr, int(nrows),
                                          self.dA.va.ptr, self.dB.va.ptr, self.cci.ptr,
                                          self.cval.ptr, int(reps), C.byref(ms))
        check("plan_time_bin", st)
        return ms.value

    def close(self):
        if self.h:
            lib().rvsp_hip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def selector_param_count(which: int) -> int:
    return lib().rvsp_selector_param_count(which)


def selector_set(which: int, params: np.ndarray, margin: float = 1.0):
    p = np.ascontiguousarray(params, np.float32)
    check("selector_set", lib().rvsp_selector_set(which, p.ctypes.data, p.size, margin))


def selector_get(which: int):
    n = selector_param_count(which)
    p = np.empty(n, np.float32)
    mg = C.c_float(0)
    st = lib().rvsp_selector_get(which, p.ctypes.data, n, C.byref(mg))
    return (p, mg.value) if st == SUCCESS else None


def selector_clear(which: int):
    check("selector_clear", lib().rvsp_selector_clear(which))


def selector_load(which: int, path: str):
    check("selector_load", lib().rvsp_selector_load_file(which, path.encode()))


def selector_predict(which: int, feats_dev: DeviceBuffer, n: int, params: np.ndarray, nout: int):
    p = np.ascontiguousarray(params, np.float32)
    out = DeviceBuffer(4 * max(n, 1) * nout)
    check("selector_predict", lib().rvsp_selector_predict(which, feats_dev.ptr, n, p.ctypes.data,
                                                          out.ptr))
    return out.to_array(np.float32, n * nout).reshape(n, nout)


def selector_grad(which: int, feats_dev: DeviceBuffer, target_dev: DeviceBuffer, n: int,
                  params: np.ndarray):
    p = np.ascontiguousarray(params, np.float32)
    g = np.empty(p.size, np.float32)
    loss = C.c_float(0)
    check("selector_grad", lib().rvsp_selector_grad(which, feats_dev.ptr, target_dev.ptr, n,
                                                    p.ctypes.data, g.ctypes.data, C.byref(loss)))
    return g, loss.value


def selector_train(which: int, feats_dev: DeviceBuffer, target_dev: DeviceBuffer, n: int,
                   params: np.ndarray, epochs=20, batch=4096, lr=3e-3, seed=1):
    p = np.array(params, np.float32)
    loss = C.c_float(0)
    check("selector_train", lib().rvsp_selector_train(which, feats_dev.ptr, target_dev.ptr, n,
                                                      epochs, batch, lr, seed, p.ctypes.data,
                                                      C.byref(loss)))
    return p, loss.value


class Event:
    def __init__(self):
        e = C.c_void_p()
        check("event_create", lib().rvsp_hip_event_create(C.byref(e)))
        self.e = e.value

    def record(self, stream=None):
        check("event_record", lib().rvsp_hip_event_record(self.e, stream))

    def sync(self):
        check("event_sync", lib().rvsp_hip_event_synchronize(self.e))

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float(0)
        check("elapsed", lib().rvsp_hip_event_elapsed_ms(self.e, later.e, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().rvsp_hip_event_destroy(self.e)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# synthetic workloads (librvsp_gen.so)
# ---------------------------------------------------------------------------
def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} missing: build with __graft_entry__.build()")
        g = C.CDLL(GEN_PATH)
        P, I32, I64, U64, V, D = C.POINTER, C.c_int32, C.c_int64, C.c_uint64, C.c_void_p, C.c_double
        outs = [P(V), P(V), P(V)]
        g.rvsp_gen_random_rows.restype = I64
        g.rvsp_gen_random_rows.argtypes = [I32, I32, I32, I32, U64, C.c_int] + outs
        g.rvsp_gen_rmat.restype = I64
        g.rvsp_gen_rmat.argtypes = [I32, I64, D, D, D, U64, C.c_int] + outs
        g.rvsp_gen_stencil27.restype = I64
        g.rvsp_gen_stencil27.argtypes = [I32, U64, C.c_int] + outs
        g.rvsp_gen_zipf_rows.restype = I64
        g.rvsp_gen_zipf_rows.argtypes = [I32, I32, D, D, I32, U64, C.c_int] + outs
        g.rvsp_gen_banded.restype = I64
        g.rvsp_gen_banded.argtypes = [I32, I32, I32, I32, I32, U64, C.c_int] + outs
        g.rvsp_gen_fill_values.restype = None
        g.rvsp_gen_fill_values.argtypes = [V, I64, U64, C.c_int]
        g.rvsp_gen_loguniform_sizes.restype = None
        g.rvsp_gen_loguniform_sizes.argtypes = [I32, I32, I32, U64, V]
        g.rvsp_gen_free.restype = None
        g.rvsp_gen_free.argtypes = [V]
        _gen = g
    return _gen


def _take(rows, cols, nnz, rp, ci, va) -> Csr:
    g = gen()
    if nnz < 0:
        raise MemoryError("generator allocation failed")
    out = Csr(
        rows,
        cols,
        np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (rows + 1,)).copy(),
        np.ctypeslib.as_array(C.cast(ci, C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy(),
        np.ctypeslib.as_array(C.cast(va, C.POINTER(C.c_float)), (max(nnz, 1),))[:nnz].copy(),
    )
    for p in (rp, ci, va):
        g.rvsp_gen_free(p)
    return out


def _outs():
    return C.c_void_p(), C.c_void_p(), C.c_void_p()


def gen_random_rows(m, n, lo, hi, seed, normal=False) -> Csr:
    rp, ci, va = _outs()
    nnz = gen().rvsp_gen_random_rows(m, n, lo, hi, seed, int(normal), C.byref(rp), C.byref(ci), C.byref(va))
    return _take(m, n, nnz, rp.value, ci.value, va.value)


def gen_rmat(scale, edges, a, b, c, seed, normal=False) -> Csr:
    rp, ci, va = _outs()
    nnz = gen().rvsp_gen_rmat(scale, edges, a, b, c, seed, int(normal), C.byref(rp), C.byref(ci), C.byref(va))
    n = 1 << scale
    return _take(n, n, nnz, rp.value, ci.value, va.value)


def gen_stencil27(g_, seed, normal=True) -> Csr:
    rp, ci, va = _outs()
    nnz = gen().rvsp_gen_stencil27(g_, seed, int(normal), C.byref(rp), C.byref(ci), C.byref(va))
    n = g_ ** 3
    return _take(n, n, nnz, rp.value, ci.value, va.value)


def gen_zipf_rows(m, n, s, mean, maxlen, seed, normal=True) -> Csr:
    rp, ci, va = _outs()
    nnz = gen().rvsp_gen_zipf_rows(m, n, s, mean, maxlen, seed, int(normal), C.byref(rp), C.byref(ci), C.byref(va))
    return _take(m, n, nnz, rp.value, ci.value, va.value)


def gen_banded(m, n, per_row, stride, half, seed, normal=True) -> Csr:
    rp, ci, va = _outs()
    nnz = gen().rvsp_gen_banded(m, n, per_row, stride, half, seed, int(normal), C.byref(rp), C.byref(ci), C.byref(va))
    return _take(m, n, nnz, rp.value, ci.value, va.value)


def fill_values(count, seed, normal=False) -> np.ndarray:
    out = np.empty(count, np.float32)
    gen().rvsp_gen_fill_values(out.ctypes.data, count, seed, int(normal))
    return out


def loguniform_sizes(count, lo, hi, seed) -> np.ndarray:
    out = np.empty(count, np.int32)
    gen().rvsp_gen_loguniform_sizes(count, lo, hi, seed, out.ctypes.data)
    return out
