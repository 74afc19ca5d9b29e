"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is independent of the oracle's own code:
  * rv-sparse's hand-computed fixtures (tests/csr_fixtures.h) and legacy 2x2
    case, and its contract op counts {2, 1};
  * set-union brute force for structure and op counts;
  * a float64 dense product for values (error bounded by the FP32 model);
  * closed forms: the 27-point stencil's C structure and product count;
    accumulation-order and no-FMA cases whose FP32 answers are exact;
  * rv-sparse's own scalar library compiled from its sources
    (oracle/_ref, when built) and the vectors it produced
    (tests/golden/ref_vectors.npz).
"""
import os

import numpy as np
import pytest

import rvsp
from oracle import pyoracle
from helpers import csr_of, dense_of, load_fixtures, structure_bruteforce

FX = load_fixtures()


@pytest.mark.parametrize("pair", FX["pairs"], ids=[p["name"] for p in FX["pairs"]])
def test_oracle_matches_reference_fixtures(pair):
    A, B, Cw = csr_of(pair["a"]), csr_of(pair["b"]), csr_of(pair["c"])
    rp, ci, va = pyoracle.spgemm(A, B)
    np.testing.assert_array_equal(rp, Cw.row_ptr)
    np.testing.assert_array_equal(ci, Cw.col_idx)
    # small integers: exact in binary32, so exact equality is the bar
    np.testing.assert_array_equal(va, Cw.values)


def test_oracle_legacy_2x2_and_op_counts():
    d = FX["legacy_oneshot_2x2"]
    A, B, Cw = csr_of(d["a"]), csr_of(d["b"]), csr_of(d["c"])
    rp, ci, va = pyoracle.spgemm(A, B)
    np.testing.assert_array_equal(va, Cw.values)
    _, ops, tot = pyoracle.symbolic(A, B)
    assert ops.tolist() == FX["descr_contract"]["op_counts"]
    assert tot == 3


def _random_pair(seed, m=60, k=50, n=70, dens=0.08):
    rng = np.random.default_rng(seed)

    def mk(r, c):
        mask = rng.random((r, c)) < dens
        if r > 3:
            mask[rng.integers(r)] = False  # an empty row
        rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
        ci = np.nonzero(mask)[1].astype(np.int32)
        va = rng.uniform(-1, 1, ci.size).astype(np.float32)
        return rvsp.Csr.make(r, c, rp, ci, va)

    return mk(m, k), mk(k, n)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_structure_vs_set_union(seed):
    A, B = _random_pair(seed)
    rp, ci, va = pyoracle.spgemm(A, B)
    brp, bci, bops = structure_bruteforce(A, B)
    np.testing.assert_array_equal(rp, brp)
    np.testing.assert_array_equal(ci, bci)
    _, ops, _ = pyoracle.symbolic(A, B)
    np.testing.assert_array_equal(ops, bops)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_values_vs_dense_f64(seed):
    A, B = _random_pair(seed)
    rp, ci, va, vd = pyoracle.spgemm(A, B, f64=True)
    D = dense_of(A) @ dense_of(B)
    absprod = np.abs(dense_of(A)) @ np.abs(dense_of(B))
    rows = np.repeat(np.arange(A.rows), np.diff(rp))
    # float64 oracle is the textbook product
    np.testing.assert_allclose(vd, D[rows, ci], rtol=1e-12, atol=1e-12)
    # FP32 sequential sum: |err| <= n_terms * 2^-23 * sum|a||b| (loose model)
    nterms = max(int((np.diff(A.row_ptr)).max()), 1)
    bound = (nterms + 1) * 2.0 ** -23 * absprod[rows, ci] + 1e-30
    assert np.all(np.abs(va.astype(np.float64) - D[rows, ci]) <= bound)


def test_oracle_sums_in_ascending_k_order():
    # C[0,0] = 1*1 + 1*1e8 + 1*(-1e8) in k order: fp32 (0+1)+1e8 = 1e8,
    # then -1e8 -> 0. Any other association gives 1.
    A = rvsp.Csr.make(1, 3, [0, 3], [0, 1, 2], [1, 1, 1])
    B = rvsp.Csr.make(3, 1, [0, 1, 2, 3], [0, 0, 0], [1.0, 1e8, -1e8])
    _, _, va = pyoracle.spgemm(A, B)
    assert va.tolist() == [0.0]
    B2 = rvsp.Csr.make(3, 1, [0, 1, 2, 3], [0, 0, 0], [1e8, -1e8, 1.0])
    _, _, va2 = pyoracle.spgemm(A, B2)
    assert va2.tolist() == [1.0]


def test_oracle_rounds_product_then_sum():
    # (-1) + fl((1+2^-12)^2): the product 1+2^-11+2^-24 rounds (tie, even)
    # to 1+2^-11, so the sum is exactly 2^-11; a fused multiply-add would
    # give 2^-11 + 2^-24.
    x = np.float32(1 + 2.0 ** -12)
    A = rvsp.Csr.make(1, 2, [0, 2], [0, 1], [-1.0, x])
    B = rvsp.Csr.make(2, 1, [0, 1, 2], [0, 0], [1.0, x])
    _, _, va = pyoracle.spgemm(A, B)
    assert va[0] == np.float32(2.0 ** -11)
    assert va[0] != np.float32(2.0 ** -11 + 2.0 ** -24)


def test_oracle_zero_sign_follows_plus_zero_start():
    # acc starts at +0.0f: a single -0.0 product yields +0.0
    A = rvsp.Csr.make(1, 1, [0, 1], [0], [-0.0])
    B = rvsp.Csr.make(1, 1, [0, 1], [0], [1.0])
    _, _, va = pyoracle.spgemm(A, B)
    assert va.view(np.uint32)[0] == 0


@pytest.mark.parametrize("g", [3, 5, 12])
def test_oracle_stencil_closed_form(g):
    A = rvsp.gen_stencil27(g, 7)
    rp, ops, tot = pyoracle.symbolic(A, A)
    assert tot == (5 * g - 6) ** 3
    assert int(ops.sum()) == (9 * g - 10) ** 3


def test_oracle_overflow_guard_small_model():
    # nnz(C) = rows * 2^16 > INT32_MAX -> status -6, like rvsp_symbolic.c:69-72
    n = 1 << 16
    rows = (1 << 15) + 1
    A = rvsp.Csr.make(rows, 1, np.arange(rows + 1), np.zeros(rows), np.ones(rows))
    B = rvsp.Csr.make(1, n, [0, n], np.arange(n), np.ones(n))
    with pytest.raises(OverflowError):
        pyoracle.symbolic(A, B)


def test_oracle_csr_check_reasons():
    o = pyoracle.oracle()

    def chk(rows, cols, rp, ci):
        rp = np.asarray(rp, np.int32)
        ci = np.asarray(ci if len(ci) else [0], np.int32)
        import ctypes as C
        bad = C.c_int32(99)
        rc = o.orc_csr_check(rows, cols, rp.ctypes.data, ci.ctypes.data, C.byref(bad))
        return rc, bad.value

    assert chk(2, 3, [0, 1, 2], [0, 2]) == (0, -1)
    assert chk(2, 3, [1, 1, 2], [0, 2])[0] == 1
    assert chk(2, 3, [0, 2, 1], [0, 2]) == (1, 1)
    assert chk(2, 3, [0, 1, 2], [0, 3]) == (2, 1)
    assert chk(1, 3, [0, 2], [2, 1]) == (3, 0)
    assert chk(1, 3, [0, 2], [1, 1]) == (4, 0)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


def _load_ref_vectors():
    z = np.load(GOLD, allow_pickle=False)
    cases = sorted({k.split("__")[0] for k in z.files})
    out = []
    for c in cases:
        g = lambda s: z[f"{c}__{s}"]
        A = rvsp.Csr.make(int(g("a_shape")[0]), int(g("a_shape")[1]), g("a_rp"), g("a_ci"), g("a_va"))
        B = rvsp.Csr.make(int(g("b_shape")[0]), int(g("b_shape")[1]), g("b_rp"), g("b_ci"), g("b_va"))
        out.append((c, A, B, g("c_rp"), g("c_ci"), g("c_va"), g("ops")))
    return out


@pytest.mark.skipif(not os.path.exists(GOLD), reason="golden vectors not generated")
def test_oracle_bitexact_vs_reference_vectors():
    for name, A, B, crp, cci, cva, ops in _load_ref_vectors():
        rp, ci, va = pyoracle.spgemm(A, B)
        _, o, _ = pyoracle.symbolic(A, B)
        np.testing.assert_array_equal(rp, crp, err_msg=name)
        np.testing.assert_array_equal(ci, cci, err_msg=name)
        np.testing.assert_array_equal(va.view(np.uint32), cva.view(np.uint32), err_msg=name)
        np.testing.assert_array_equal(o, ops, err_msg=name)


@pytest.mark.skipif(not pyoracle.ref_available(), reason="reference library not built")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_bitexact_vs_reference_library(seed):
    A, B = _random_pair(100 + seed, m=300, k=250, n=280, dens=0.05)
    rrp, rci, rva, rops = pyoracle.ref_spgemm(A, B, want_ops=True)
    rp, ci, va = pyoracle.spgemm(A, B)
    _, ops, _ = pyoracle.symbolic(A, B)
    np.testing.assert_array_equal(rp, rrp)
    np.testing.assert_array_equal(ci, rci)
    np.testing.assert_array_equal(va.view(np.uint32), rva.view(np.uint32))
    np.testing.assert_array_equal(ops, rops)


@pytest.mark.skipif(not pyoracle.ref_available(), reason="reference library not built")
def test_oracle_bitexact_vs_reference_library_workloads():
    from rvsp import workloads as W
    for name in ("uniform", "stencil", "rmat"):
        A, B = W.SMALL[name]()
        rrp, rci, rva = pyoracle.ref_spgemm(A, B)
        rp, ci, va = pyoracle.spgemm(A, B)
        np.testing.assert_array_equal(rp, rrp, err_msg=name)
        np.testing.assert_array_equal(ci, rci, err_msg=name)
        np.testing.assert_array_equal(va.view(np.uint32), rva.view(np.uint32), err_msg=name)


def test_planted_mistake_is_caught():
    """A mistake planted in a copy of the oracle's output (one value nudged
    by one ulp, one column swapped) must fail the bit-exact comparison."""
    from helpers import assert_bitexact
    A, B = _random_pair(3)
    rp, ci, va = pyoracle.spgemm(A, B)
    bad_va = va.copy()
    bad_va[len(bad_va) // 2] = np.nextafter(bad_va[len(bad_va) // 2], np.float32(np.inf))
    with pytest.raises(AssertionError):
        assert_bitexact(rvsp.Csr(A.rows, B.cols, rp, ci, bad_va), rp, ci, va)
    bad_ci = ci.copy()
    j = int(np.argmax(np.diff(rp)))
    s = rp[j]
    bad_ci[s], bad_ci[s + 1] = bad_ci[s + 1], bad_ci[s]
    with pytest.raises(AssertionError):
        assert_bitexact(rvsp.Csr(A.rows, B.cols, rp, bad_ci, va), rp, ci, va)


@pytest.mark.parametrize("wl", ["rmat", "mixed", "stencil"])
def test_row_parallel_oracle_equals_sequential(wl):
    """The row-parallel count and fill/numeric (used at full benchmark
    sizes) give the sequential passes' outputs bit for bit, and the FP64
    pass equals orc_numeric_f64."""
    from rvsp import workloads as W

    A, B = W.SMALL[wl]()
    rp, ops, tot = pyoracle.symbolic(A, B)
    prp, pops, ptot = pyoracle.symbolic(A, B, threads=4)
    assert tot == ptot
    np.testing.assert_array_equal(rp, prp)
    np.testing.assert_array_equal(ops, pops)
    want = pyoracle.spgemm(A, B, f64=True)
    got = pyoracle.spgemm_par(A, B, threads=4)
    for w, g in zip(want, got):
        np.testing.assert_array_equal(np.asarray(w).view(np.uint8), np.asarray(g).view(np.uint8))


def test_row_parallel_count_reports_overflow():
    # 2^15 rows, each with 2^16 distinct columns: nnz(C) = 2^31 > INT32_MAX
    m, n = 1 << 15, 1 << 16
    B = rvsp.Csr.make(1, n, [0, n], np.arange(n, dtype=np.int32), np.ones(n, np.float32))
    A = rvsp.Csr.make(m, 1, np.arange(m + 1, dtype=np.int32), np.zeros(m, np.int32), np.ones(m, np.float32))
    with pytest.raises(OverflowError):
        pyoracle.symbolic(A, B, threads=4)
