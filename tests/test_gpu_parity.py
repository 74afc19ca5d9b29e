"""GPU parity: the HIP path, through the C ABI, against the CPU oracle.

Bar: row_ptr and col_idx bit-exact; values bit-exact (the kernels keep the
reference's ascending-k, separately rounded accumulation), which implies the
1e-5 relative bound; FP64 error stats are printed per case. Covers
rv-sparse's fixtures and descriptor contract (tests/test_spgemm_*_f32.c),
crafted order/rounding cases, every workload at test scale, the uniform
workload at full size, edge cases (empty rows, empty products, overflow),
device-resident and graph-replayed reuse, and determinism.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import rvsp
from rvsp import workloads as W
from oracle import pyoracle
from helpers import assert_bitexact, csr_of, load_fixtures, rel_err_stats

pytestmark = pytest.mark.gpu
FX = load_fixtures()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_product(A, B, C_, what):
    rp, ci, va, vd = pyoracle.spgemm(A, B, f64=True)
    assert_bitexact(C_, rp, ci, va, what)
    mx, mean = rel_err_stats(C_.values, vd)
    say(f"[{what}] nnzC={len(ci)} max_rel_err_vs_f64={mx:.3e} mean={mean:.3e}")


# ---------------------------------------------------------------------------
# rv-sparse fixtures, every strategy the build provides (test_spgemm_reference_f32.c)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("pair", FX["pairs"], ids=[p["name"] for p in FX["pairs"]])
@pytest.mark.parametrize("algo", [rvsp.ALGO_DEFAULT, rvsp.ALGO_HIP, rvsp.ALGO_RVV, rvsp.ALGO_CONTIG,
                                  rvsp.ALGO_ADAPTIVE])
def test_fixture_every_strategy(pair, algo):
    say(f"fixture {pair['name']} algo {algo}")
    A, B, Cw = csr_of(pair["a"]), csr_of(pair["b"]), csr_of(pair["c"])
    d = rvsp.Descr()
    if algo in (rvsp.ALGO_RVV, rvsp.ALGO_CONTIG, rvsp.ALGO_ADAPTIVE):
        # the RISC-V Vector strategies: refused, as by an rv-sparse build without V
        # (rv-sparse's test_spgemm_reference_f32.c skips them on that status)
        assert d.set_algo(algo) == rvsp.ERROR_UNSUPPORTED_BACKEND
        d.close()
        return
    assert d.set_algo(algo) == rvsp.SUCCESS
    a, b = A.struct(), B.struct()
    st, ws, nnz = d.estimate(a, b)
    assert st == 0 and nnz == Cw.nnz
    rp = np.zeros(A.rows + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    va = np.zeros(max(nnz, 1), np.float32)
    wsb = np.zeros(ws, np.uint8)
    c = rvsp.CsrStruct(A.rows, B.cols, nnz, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    assert d.compute(a, b, c, wsb.ctypes.data) == 0
    np.testing.assert_array_equal(rp, Cw.row_ptr)
    np.testing.assert_array_equal(ci[:nnz], Cw.col_idx)
    # integers: exact; the reference's own tolerance is 1e-5 + 1e-4|want|
    np.testing.assert_array_equal(va[:nnz], Cw.values)
    d.close()


def test_legacy_oneshot_2x2_scalar_backend_and_null_options():
    d = FX["legacy_oneshot_2x2"]
    A, B, Cw = csr_of(d["a"]), csr_of(d["b"]), csr_of(d["c"])
    for be in (rvsp.BACKEND_SCALAR, rvsp.BACKEND_SCALAR_UNROLL4, rvsp.BACKEND_HIP):
        _, C_ = rvsp.spgemm_csr(A, B, backend=be)
        assert (C_.rows, C_.cols, C_.nnz) == (2, 2, 3)
        np.testing.assert_array_equal(C_.row_ptr, Cw.row_ptr)
        np.testing.assert_array_equal(C_.col_idx, Cw.col_idx)
        np.testing.assert_array_equal(C_.values, Cw.values)
    # options == NULL
    L = rvsp.lib()
    a, b = A.struct(), B.struct()
    c = rvsp.CsrStruct()
    assert L.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), None) == 0
    assert c.owns_data == 1 and c.nnz == 3
    L.rvsp_csr_destroy(C.byref(c))


# ---------------------------------------------------------------------------
# descriptor contract (test_spgemm_descr_f32.c:45-240) on the GPU
# ---------------------------------------------------------------------------
def test_descriptor_contract_full_sequence():
    say("descriptor contract")
    L = rvsp.lib()
    pair = FX["pairs"][0]
    A, B = csr_of(pair["a"]), csr_of(pair["b"])
    a_vals = A.values.copy()
    A = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, a_vals)
    a, b = A.struct(), B.struct()
    d = rvsp.Descr()
    assert d.set_algo(rvsp.ALGO_DEFAULT) == 0
    st, ws, nnz = d.estimate(a, b)
    assert st == 0 and nnz == 3 and ws > 0
    assert d.set_algo(rvsp.ALGO_DEFAULT) == rvsp.ERROR_INVALID_ARGUMENT
    st, ops = d.op_counts(2)
    assert st == 0 and ops.tolist() == FX["descr_contract"]["op_counts"]
    st, _ = d.op_counts(1)
    assert st == rvsp.ERROR_INVALID_ARGUMENT
    rp = np.zeros(3, np.int32)
    ci = np.zeros(3, np.int32)
    va = np.zeros(3, np.float32)
    wsb = np.zeros(ws, np.uint8)
    c = rvsp.CsrStruct(2, 2, nnz, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    assert d.compute(a, b, c, None) == rvsp.ERROR_NULL_POINTER
    small = rvsp.CsrStruct(2, 2, nnz - 1, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    assert d.compute(a, b, small, wsb.ctypes.data) == rvsp.ERROR_INVALID_ARGUMENT
    wrong = rvsp.CsrStruct(3, 2, nnz, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    assert d.compute(a, b, wrong, wsb.ctypes.data) == rvsp.ERROR_INVALID_ARGUMENT
    bad_a = A.struct()
    bad_a.nnz = 2
    assert d.compute(bad_a, b, c, wsb.ctypes.data) == rvsp.ERROR_INVALID_ARGUMENT
    assert d.compute(a, b, c, wsb.ctypes.data) == 0
    assert c.nnz == 3
    np.testing.assert_array_equal(ci, [0, 1, 1])
    np.testing.assert_array_equal(va, [4, 10, 15])
    # reuse: values of A doubled in place, no re-analysis, same C
    a_vals *= 2
    ci[:] = [7, 7, 7]  # a stale-structure sentinel: must stay untouched on reuse
    assert d.compute(a, b, c, wsb.ctypes.data) == 0
    np.testing.assert_array_equal(va, [8, 20, 30])
    np.testing.assert_array_equal(ci, [7, 7, 7])
    d.invalidate()
    assert d.compute(a, b, c, wsb.ctypes.data) == 0
    np.testing.assert_array_equal(ci, [0, 1, 1])
    d.close()


# ---------------------------------------------------------------------------
# crafted cases with exact FP32 answers
# ---------------------------------------------------------------------------
def test_gpu_sums_in_ascending_k_order_and_rounds_separately():
    A = rvsp.Csr.make(1, 3, [0, 3], [0, 1, 2], [1, 1, 1])
    B = rvsp.Csr.make(3, 1, [0, 1, 2, 3], [0, 0, 0], [1.0, 1e8, -1e8])
    assert rvsp.spgemm_raw(A, B).values.tolist() == [0.0]
    x = np.float32(1 + 2.0 ** -12)
    A2 = rvsp.Csr.make(1, 2, [0, 2], [0, 1], [-1.0, x])
    B2 = rvsp.Csr.make(2, 1, [0, 1, 2], [0, 0], [1.0, x])
    assert rvsp.spgemm_raw(A2, B2).values[0] == np.float32(2.0 ** -11)
    A3 = rvsp.Csr.make(1, 1, [0, 1], [0], [-0.0])
    B3 = rvsp.Csr.make(1, 1, [0, 1], [0], [1.0])
    assert rvsp.spgemm_raw(A3, B3).values.view(np.uint32)[0] == 0


def test_gpu_long_cancellation_chain_every_bin():
    """Rows whose entries cancel to exactly 0.0 after many ordered adds, in
    rows long enough to reach the hash, dense and window accumulators."""
    rng = np.random.default_rng(5)
    for ncols in (8, 300, 3000, 9000):
        k = 40
        # B row j: all ncols columns with value v_j; A row: k ones
        vals = rng.uniform(-1, 1, k).astype(np.float32)
        Bc = np.tile(np.arange(ncols, dtype=np.int32), k)
        Bv = np.repeat(vals, ncols)
        B = rvsp.Csr.make(k, ncols, np.arange(k + 1) * ncols, Bc, Bv)
        A = rvsp.Csr.make(2, k, [0, k, 2 * k], np.tile(np.arange(k), 2), np.ones(2 * k))
        C_ = rvsp.spgemm_raw(A, B)
        check_product(A, B, C_, f"chain-{ncols}")


# ---------------------------------------------------------------------------
# workloads at test scale, and uniform at full size
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["uniform", "rmat", "stencil", "mixed"])
def test_workload_small_oneshot(name):
    say(f"workload {name}")
    A, B = W.SMALL[name]()
    C_, ops = rvsp.spgemm_raw(A, B, want_ops=True)
    check_product(A, B, C_, name)
    _, oops, _ = pyoracle.symbolic(A, B)
    np.testing.assert_array_equal(ops, oops)


def test_workload_small_batched_each():
    for i, (A, B) in enumerate(W.SMALL["batched"]()):
        C_ = rvsp.spgemm_raw(A, B)
        check_product(A, B, C_, f"batched[{i}] n={A.rows}")


def test_uniform_full_size():
    say("uniform 4096^2 x 16/row, C = A*A")
    A, B = W.uniform()
    _, C_ = rvsp.spgemm_csr(A, B)
    check_product(A, B, C_, "uniform-4096")
    assert int(pyoracle.symbolic(A, B)[1].sum()) == 4096 * 16 * 16


def test_stencil_closed_form_on_gpu():
    g = 16
    A, _ = W.stencil(g=g, seed=9)
    d = rvsp.Descr()
    st, _, nnz = d.estimate(A.struct(), A.struct())
    assert st == 0 and nnz == (5 * g - 6) ** 3
    assert d.products() == (9 * g - 10) ** 3
    d.close()


def test_rmat_wide_rows_exercise_windows():
    """Scale 15 R-MAT: rows spanning more than 16384 columns with thousands
    of distinct ones (column windows with cursors) next to tiny ones (hash)."""
    say("rmat scale 15")
    A, B = W.rmat(scale=15, seed=31)
    C_ = rvsp.spgemm_raw(A, B)
    check_product(A, B, C_, "rmat-15")


@pytest.mark.parametrize("alen", [10, 700])
def test_wide_random_rows_windows(alen):
    """Rows of 10 or 700 A entries over B rows of ~300 random columns in
    [0, 100000): spans far above one window, so every window, cursor and
    empty-window path runs; 700 entries per row stresses per-entry cursors."""
    say(f"wide rows alen={alen}")
    B = rvsp.gen_random_rows(3000, 100000, 250, 350, 77)
    A = rvsp.gen_random_rows(40, 3000, alen, alen, 78)
    check_product(A, B, rvsp.spgemm_raw(A, B), f"wide-{alen}")


@pytest.mark.parametrize("threads", ["256", "512", "1024"])
@pytest.mark.parametrize("shape", ["2048x16", "4096x32", "2048x32", "1024x16", "1024x8"])
def test_big_rows_every_window_shape_and_count_block(shape, threads, monkeypatch):
    """Big rows under every launchable numeric window shape and every block
    size of the symbolic count: windows that split rows at many places
    (segments starting mid-row, pieces ending inside windows, groups of
    entries with nothing in a window)."""
    monkeypatch.setenv("RVSP_BIG_WINDOW", shape)
    monkeypatch.setenv("RVSP_SYM_BIG_THREADS", threads)
    say(f"big rows shape={shape} count threads={threads}")
    A, B = W.rmat(scale=15, seed=31)
    check_product(A, B, rvsp.spgemm_raw(A, B), f"rmat-15 {shape} threads={threads}")
    B2 = rvsp.gen_random_rows(3000, 100000, 250, 350, 77)
    A2 = rvsp.gen_random_rows(40, 3000, 700, 700, 78)
    check_product(A2, B2, rvsp.spgemm_raw(A2, B2), f"wide-700 {shape} threads={threads}")


@pytest.mark.parametrize("shape", ["2048x16", "4096x32", "1024x8"])
def test_fill_numeric_matches_fill_then_numeric(shape, monkeypatch):
    """The combined call (big rows' columns written by the numeric pass)
    leaves exactly what fill followed by numeric leaves, column indices and
    value bits, including over a C buffer holding stale data."""
    monkeypatch.setenv("RVSP_BIG_WINDOW", shape)
    for name, (A, B) in [("rmat-15", W.rmat(scale=15, seed=31)),
                         ("wide-700", (rvsp.gen_random_rows(40, 3000, 700, 700, 78),
                                       rvsp.gen_random_rows(3000, 100000, 250, 350, 77)))]:
        say(f"fill+numeric {name} {shape}")
        dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
        sep = rvsp.Plan(dA, dB)
        sep.fill()
        sep.numeric()
        ref = sep.result()
        one = rvsp.Plan(dA, dB)
        one.cci.upload(np.full(max(one.nnz, 1), -7, np.int32))
        one.cval.upload(np.full(max(one.nnz, 1), np.nan, np.float32))
        one.fill_numeric()
        got = one.result()
        np.testing.assert_array_equal(got.col_idx, ref.col_idx)
        np.testing.assert_array_equal(got.values.view(np.uint32), ref.values.view(np.uint32))
        check_product(A, B, got, f"{name} fill+numeric {shape}")
        sep.close()
        one.close()


# ---------------------------------------------------------------------------
# edge cases
# ---------------------------------------------------------------------------
def test_empty_product_and_empty_rows():
    A = rvsp.Csr.make(3, 4, [0, 0, 0, 0], [], [])
    B = rvsp.Csr.make(4, 5, [0, 1, 2, 2, 2], [0, 4], [1, 2])
    C_ = rvsp.spgemm_raw(A, B)
    assert C_.nnz == 0 and C_.row_ptr.tolist() == [0, 0, 0, 0]
    A2 = rvsp.Csr.make(3, 4, [0, 1, 1, 2], [2, 1], [1, 3])
    C2 = rvsp.spgemm_raw(A2, B)  # row 0 hits an empty B row
    check_product(A2, B, C2, "empty-B-row")


def test_single_dense_row_and_column():
    n = 5000
    A = rvsp.Csr.make(1, n, [0, n], np.arange(n), np.linspace(-1, 1, n))
    B = rvsp.Csr.make(n, 3, np.arange(n + 1), np.arange(n) % 3, np.linspace(1, -1, n))
    check_product(A, B, rvsp.spgemm_raw(A, B), "dense-row")
    At = rvsp.Csr.make(n, 1, np.arange(n + 1), np.zeros(n), np.linspace(-1, 1, n))
    Bt = rvsp.Csr.make(1, n, [0, n], np.arange(n), np.linspace(1, -1, n))
    check_product(At, Bt, rvsp.spgemm_raw(At, Bt), "outer-product")


def test_nnz_overflow_reports_allocation_failed():
    n = 1 << 16
    rows = (1 << 15) + 1
    A = rvsp.Csr.make(rows, 1, np.arange(rows + 1), np.zeros(rows), np.ones(rows))
    B = rvsp.Csr.make(1, n, [0, n], np.arange(n), np.ones(n))
    d = rvsp.Descr()
    st, _, _ = d.estimate(A.struct(), B.struct())
    assert st == rvsp.ERROR_ALLOCATION_FAILED
    d.close()


def test_allocation_failure_cleans_up():
    A, B = W.SMALL["uniform"]()
    L = rvsp.lib()
    for n in (1, 2, 3, 5):
        L.rvsp_hip_fail_next_allocs(n)
        st, C_ = rvsp.spgemm_csr(A, B, raise_on_error=False)
        L.rvsp_hip_fail_next_allocs(0)
        assert st == rvsp.ERROR_ALLOCATION_FAILED and C_ is None
    _, C_ = rvsp.spgemm_csr(A, B)  # and the library still works
    check_product(A, B, C_, "after-alloc-failures")


# ---------------------------------------------------------------------------
# device-resident paths, reuse, graphs, determinism
# ---------------------------------------------------------------------------
def test_device_oneshot_matches_host_oneshot():
    A, B = W.SMALL["mixed"]()
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    C_ = rvsp.spgemm_device(dA, dB)
    check_product(A, B, C_, "device-oneshot")


def test_device_oneshots_reuse_freed_blocks_without_aliasing_live_ones():
    """Device one-shots recycle blocks freed on the library stream: a C that
    is still alive is never handed out again, and results stay exact when a
    product lands in blocks another product used."""
    L = rvsp.lib()
    prods = [W.SMALL["mixed"](), W.SMALL["rmat"](), W.SMALL["uniform"]()]
    dev = [(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)) for A, B in prods]

    def run(i):
        a, b = dev[i][0].struct(), dev[i][1].struct()
        c = rvsp.CsrStruct()
        rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c),
                                                               rvsp.ALGO_HIP))
        return c

    def spans(c):
        return [(c.row_ptr, c.row_ptr + 4 * (c.rows + 1)), (c.col_idx, c.col_idx + 4 * max(c.nnz, 1)),
                (c.values, c.values + 4 * max(c.nnz, 1))]

    def check(i, c, tag):
        A, B = prods[i]
        check_product(A, B, rvsp.device_csr_to_host(c), tag)

    for rnd in range(2):
        say(f"block reuse round {rnd}")
        live = [run(i) for i in range(3)]
        ranges = [r for c in live for r in spans(c)]
        for x in range(len(ranges)):
            for y in range(x + 1, len(ranges)):
                assert ranges[x][1] <= ranges[y][0] or ranges[y][1] <= ranges[x][0], "live blocks overlap"
        for i, c in enumerate(live):
            check(i, c, f"round {rnd} product {i}")
        L.rvsp_csr_destroy(C.byref(live[1]))
        again = run(0)  # may take the freed blocks of the other product
        check(0, again, f"round {rnd} product 0 in recycled blocks")
        for c in (live[0], live[2], again):
            L.rvsp_csr_destroy(C.byref(c))


@pytest.mark.parametrize("graph", [False, True])
def test_device_descriptor_reuse_with_redrawn_values(graph):
    say(f"device reuse graph={graph}")
    A, _ = W.stencil(g=10, seed=41)
    dA = rvsp.DeviceCsr(A)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_graph(graph) == 0
    a = dA.struct()
    st, ws, nnz = d.estimate(a, a)
    assert st == 0
    crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
    cci = rvsp.DeviceBuffer(4 * nnz)
    cva = rvsp.DeviceBuffer(4 * nnz)
    c = rvsp.CsrStruct(A.rows, A.cols, nnz, crp.ptr, cci.ptr, cva.ptr, 3, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    for p in range(4):
        vals = rvsp.fill_values(A.nnz, 1000 + p, normal=True)
        dA.va.upload(vals)
        assert d.compute(a, a, c, wsb.ctypes.data) == 0
        assert d.wait() == 0
        got = rvsp.Csr(A.rows, A.cols, crp.to_array(np.int32, A.rows + 1), cci.to_array(np.int32, nnz),
                       cva.to_array(np.float32, nnz))
        Ah = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, vals)
        check_product(Ah, Ah, got, f"reuse pass {p}")
    d.close()


def test_numeric_is_deterministic():
    A, B = W.SMALL["rmat"]()
    c1 = rvsp.spgemm_raw(A, B)
    c2 = rvsp.spgemm_raw(A, B)
    np.testing.assert_array_equal(c1.values.view(np.uint32), c2.values.view(np.uint32))


# ---------------------------------------------------------------------------
# batched products: one stacked launch set, identical to separate calls
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("chunk_products", [None, "20000"])
def test_batched_matches_separate_calls(chunk_products, monkeypatch):
    if chunk_products:
        monkeypatch.setenv("RVSP_BATCH_CHUNK_PRODUCTS", chunk_products)
    pairs = W.batched(count=9, lo=64, hi=1500, seed=21)
    dA = [rvsp.DeviceCsr(A) for A, _ in pairs]
    dB = [rvsp.DeviceCsr(B) for _, B in pairs]
    got, chunks = rvsp.spgemm_batched(dA, dB)
    say(f"batched: {len(pairs)} problems in {chunks} chunk(s)")
    if chunk_products:
        assert chunks > 1
    for i, ((A, B), Ci) in enumerate(zip(pairs, got)):
        single = rvsp.spgemm_device(dA[i], dB[i])
        np.testing.assert_array_equal(Ci.row_ptr, single.row_ptr)
        np.testing.assert_array_equal(Ci.col_idx, single.col_idx)
        np.testing.assert_array_equal(Ci.values.view(np.uint32), single.values.view(np.uint32))
        check_product(A, B, Ci, f"batched[{i}]")
