"""The synthetic generator (include/genmat.h, rv-sparse_amd/src/core/genmat.c)
against rv-sparse's own generator built from tools/src/genmat.c into
oracle/_ref (bit for bit: structure and double values), plus properties the
generator's definition fixes on its own: the value stream is glibc's rand()
after srand(seed), square lines hold their diagonal and stay inside their
band, column-wise generation is the transpose of row-wise generation of the
transposed shape, and the Matrix Market writer round-trips through
rvsp_mtx_load. No GPU."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

import rvsp
from oracle import genmat as og
from oracle import mtx as omtx

# run_bench.sh:185-190, the reference bench's gen sweep (R C DENSITY SEED)
GEN_CASES = [(512, 512, 0.02, 42), (1024, 1024, 0.01, 42), (2048, 2048, 0.005, 42),
             (4096, 4096, 0.002, 42), (64, 64, 0.05, 1)]


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def same(got, want):
    if got is None or want is None:
        return got is None and want is None
    return all(np.array_equal(bits(g), bits(w)) for g, w in zip(got, want))


def need_ref():
    if not og.ref_available():
        pytest.skip("rv-sparse reference generator not built (oracle/build_ref.sh)")


def random_params(rng):
    r = int(rng.integers(1, 400))
    c = r if rng.random() < 0.5 else int(rng.integers(1, 400))
    kw = dict(density=float(rng.choice([0.001, 0.004, 0.02, 0.1, 0.6])),
              random_seed=int(rng.integers(-10, 1 << 20)),
              cv=float(rng.choice([0.0, 0.25, 0.5, 1.5])), min=int(rng.integers(1, 4)))
    if rng.random() < 0.3:
        kw["imbalance"] = float(rng.choice([0.5, 3.0]))
    if rng.random() < 0.3:
        kw["is_column"] = 1
    if r == c and rng.random() < 0.6:
        kw["low_bandwidth"] = int(rng.integers(1, r + 1))
        kw["up_bandwidth"] = int(rng.integers(1, r + 1))
    return r, c, kw


@pytest.mark.parametrize("R,Cc,D,S", GEN_CASES)
def test_bench_sweep_matches_reference_generator(R, Cc, D, S):
    need_ref()
    for rows, cols, seed in ((R, Cc, S), (Cc, R, S + 1)):  # bench.c:677-696
        kw = dict(density=D, random_seed=seed, cv=0.5, min=1)
        got = rvsp.genmat(rows, cols, **kw)
        assert got is not None and got[3].size > 0
        assert same(got, og.ref_generate(rvsp.genmat_params(rows, cols, **kw)))


def test_random_parameters_match_reference_generator():
    """Square (both banding cases) and rectangular shapes, row and column
    lines, normal (cv < 1/3) and log-normal degrees, negative-mean seeds,
    imbalance caps."""
    need_ref()
    rng = np.random.default_rng(2024)
    for _ in range(200):
        r, c, kw = random_params(rng)
        got = rvsp.genmat(r, c, **kw)
        assert same(got, og.ref_generate(rvsp.genmat_params(r, c, **kw))), (r, c, kw)


def test_default_params_match_reference():
    need_ref()
    for r, c in ((10, 10), (7, 300)):
        a, b = rvsp.genmat_params(r, c), og.ref_default_params(r, c)
        for name, _ in rvsp.GenmatParams._fields_:
            assert getattr(a, name) == getattr(b, name), name


def test_values_are_glibc_rand_stream():
    """values[t] = (rand() % 9 + 1.0) / 10 for the t-th entry in row order,
    after srand(seed), drawn through libc's own process-wide rand()."""
    libc = C.CDLL(ctypes.util.find_library("c"))
    for seed in (1, 42, 123457):
        _, _, _, _, va = rvsp.genmat(300, 200, density=0.05, random_seed=seed)
        libc.srand(seed)
        want = np.array([(libc.rand() % 9 + 1.0) / 10 for _ in range(va.size)])
        np.testing.assert_array_equal(bits(va), bits(want))


def test_square_lines_hold_diagonal_and_stay_in_band():
    rng = np.random.default_rng(5)
    for _ in range(40):
        n = int(rng.integers(2, 300))
        low, up = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        _, _, rp, ci, _ = rvsp.genmat(n, n, density=0.05, random_seed=int(rng.integers(1, 999)),
                                      low_bandwidth=low, up_bandwidth=up)
        for j in range(n):
            row = ci[rp[j]:rp[j + 1]]
            assert j in row
            assert np.all(np.diff(row) > 0)
            assert row[0] >= max(0, j - low) and row[-1] < min(n, j + up)


def transpose(rows, cols, rp, ci, va):
    order = np.lexsort((np.repeat(np.arange(rows), np.diff(rp)), ci))
    t_rp = np.zeros(cols + 1, np.int32)
    np.add.at(t_rp, ci + 1, 1)
    return cols, rows, np.cumsum(t_rp).astype(np.int32), np.repeat(np.arange(rows), np.diff(rp))[order].astype(
        np.int32), va[order]


@pytest.mark.parametrize("r,c", [(150, 90), (64, 64), (40, 500)])
def test_column_lines_are_transpose_of_row_lines(r, c):
    """is_column draws line j of an r x c matrix as row-wise generation
    draws row j of the c x r one (same seeds and windows), values in line
    order, so the two are transposes of each other."""
    kw = dict(density=0.05, random_seed=77)
    by_col = rvsp.genmat(r, c, is_column=1, **kw)
    by_row = rvsp.genmat(c, r, **kw)
    assert same(by_col, transpose(*by_row))


def test_errors_and_symmetric():
    assert rvsp.genmat(10, 12, is_symmetric=1) is None  # not square (the reference's error too)
    assert rvsp.genmat(10, 10, is_symmetric=1) is None  # symmetric generation is not built
    assert rvsp.genmat(-1, 5) is None
    assert rvsp.genmat(10, 10, low_bandwidth=-1) is None
    # a band with an empty window for row 0 while draws are due
    assert rvsp.genmat(10, 10, density=0.5, up_bandwidth=0, low_bandwidth=0) is None
    if og.ref_available():
        assert og.ref_generate(rvsp.genmat_params(10, 12, is_symmetric=1)) is None
    empty = rvsp.genmat(0, 0)
    assert empty is not None and empty[2].tolist() == [0] and empty[3].size == 0


def test_save_to_mtx_round_trip_and_reference_text(tmp_path):
    p = rvsp.genmat_params(120, 80, density=0.05, random_seed=9)
    m = rvsp.lib().genmat_generate_csr(p)
    path = tmp_path / "g.mtx"
    assert rvsp.lib().genmat_save_csr_to_mtx(C.byref(m), str(path).encode()) == 0
    rows, cols, rp, ci, va = rvsp.genmat_arrays(m)
    rvsp.lib().genmat_free_csr(C.byref(m))
    back = rvsp.mtx_load(str(path))
    # the reader infers columns as the largest index + 1
    assert back.rows == rows and back.cols == int(ci.max()) + 1
    np.testing.assert_array_equal(back.row_ptr, rp)
    np.testing.assert_array_equal(back.col_idx, ci)
    np.testing.assert_array_equal(back.values, va.astype(np.float32))
    M, N, orp, oci, ova = omtx.load(str(path))
    np.testing.assert_array_equal(orp, rp)
    if og.ref_available():
        lib = og._ref()
        lib.genmat_save_csr_to_mtx.restype = C.c_int
        lib.genmat_save_csr_to_mtx.argtypes = [C.POINTER(rvsp.GenmatCsr), C.c_char_p]
        rm = lib.genmat_generate_csr(p)
        rpath = tmp_path / "r.mtx"
        assert lib.genmat_save_csr_to_mtx(C.byref(rm), str(rpath).encode()) == 0
        lib.genmat_free_csr(C.byref(rm))
        assert path.read_bytes() == rpath.read_bytes()


@pytest.mark.parametrize("dtype", [rvsp.DTYPE_FP32, rvsp.DTYPE_FP64, rvsp.DTYPE_INT8])
def test_bench_pair_conversions(dtype):
    """rvsp_genmat_bench_pair = bench.c's --gen inputs: A (R x C, seed),
    B (C x R, seed + 1), values converted as bench.c:391-450."""
    R, Cc, D, S = 300, 200, 0.03, 42
    a, b = rvsp.CsrStruct(), rvsp.CsrStruct()
    L = rvsp.lib()
    assert L.rvsp_genmat_bench_pair(R, Cc, D, S, dtype, C.byref(a), C.byref(b)) == 0
    npt = {rvsp.DTYPE_FP32: (C.c_float, np.float32), rvsp.DTYPE_FP64: (C.c_double, np.float64),
           rvsp.DTYPE_INT8: (C.c_int8, np.int8)}[dtype]
    for c, (rows, cols, seed) in ((a, (R, Cc, S)), (b, (Cc, R, S + 1))):
        g = rvsp.genmat(rows, cols, density=D, random_seed=seed, cv=0.5, min=1)
        assert (c.rows, c.cols, c.nnz, c.dtype, c.owns_data) == (rows, cols, g[3].size, dtype, 1)
        rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (rows + 1,))
        ci = np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (c.nnz,))
        va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(npt[0])), (c.nnz,))
        np.testing.assert_array_equal(rp, g[2])
        np.testing.assert_array_equal(ci, g[3])
        if dtype == rvsp.DTYPE_INT8:
            np.testing.assert_array_equal(va, np.ones(c.nnz, np.int8))  # (int)0.x % 127 = 0 -> 1
        else:
            np.testing.assert_array_equal(va, g[4].astype(npt[1]))
        L.rvsp_csr_destroy(C.byref(c))
    assert L.rvsp_genmat_bench_pair(R, Cc, D, S, rvsp.DTYPE_INT32, C.byref(a), C.byref(b)) == \
        rvsp.ERROR_UNSUPPORTED_DTYPE
    assert L.rvsp_genmat_bench_pair(R, Cc, D, S, dtype, None, C.byref(b)) == rvsp.ERROR_NULL_POINTER
