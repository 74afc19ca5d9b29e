"""Special FP32 values on the GPU, against the x86 oracle.

Contract: C's structure is bit-exact; every C value that is not NaN is
bit-identical to the reference's (subnormal products and sums included,
signed zeros and infinities included); a C value is NaN exactly where the
reference's is NaN. NaN bits themselves are not compared: on x86 an
invalid operation (Inf * 0, Inf - Inf) gives the negative default NaN
0xFFC00000 and an add of two NaNs returns its first operand's, which
depends on how the compiler orders `acc + a * b`; the GPU's multiplies
and adds follow their own NaN encoding. No caller can rely on NaN payloads.

Values are drawn so that products and running sums are subnormal
(2^-70 * 2^-70 scales, cancellations into the subnormal range), or hit
+-Inf, NaN, Inf * 0 and Inf - Inf, and the rows are routed to every
accumulator: the default bins, each forced numeric bin (LDS hash tables,
dense and rank windows at 8..64 lanes), every big-row window shape, and
the ranked reuse pass.
"""
import numpy as np
import pytest

import rvsp
from rvsp import selector as S
from rvsp import workloads as W
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def subnormal_values(n, seed):
    """Magnitudes 2^-60..2^-80 with random signs: products 2^-120..2^-160,
    most of them subnormal or rounding to a signed zero, and sums that
    cancel into the subnormal range."""
    rng = np.random.default_rng(seed)
    e = rng.integers(-80, -59, n)
    m = rng.uniform(1.0, 2.0, n)
    s = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    return (s * np.ldexp(m, e)).astype(np.float32)


def special_values(n, seed):
    """Ordinary values with +-Inf, NaN, 0, -0 and huge values sprinkled in
    (products overflow to Inf; Inf * 0 and Inf - Inf give NaN)."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, n).astype(np.float32)
    pick = rng.random(n)
    v[pick < 0.01] = np.inf
    v[(pick >= 0.01) & (pick < 0.02)] = -np.inf
    v[(pick >= 0.02) & (pick < 0.025)] = np.nan
    v[(pick >= 0.025) & (pick < 0.05)] = 0.0
    v[(pick >= 0.05) & (pick < 0.07)] = -0.0
    v[(pick >= 0.07) & (pick < 0.09)] = np.float32(3e38)
    return v


def with_values(M, v):
    return rvsp.Csr(M.rows, M.cols, M.row_ptr, M.col_idx, v)


def assert_contract(got, want_rp, want_ci, want_va, what):
    np.testing.assert_array_equal(got.row_ptr, want_rp, err_msg=f"{what}: row_ptr")
    np.testing.assert_array_equal(got.col_idx, want_ci, err_msg=f"{what}: col_idx")
    g, w = got.values, np.asarray(want_va, np.float32)
    gn, wn = np.isnan(g), np.isnan(w)
    bad = np.nonzero(gn != wn)[0]
    assert bad.size == 0, f"{what}: NaN where the reference has none (or the reverse) at {bad[:5]}"
    ok = ~wn
    diff = np.nonzero(g[ok].view(np.uint32) != w[ok].view(np.uint32))[0]
    assert diff.size == 0, (f"{what}: {diff.size} values differ, first "
                            f"{g[ok][diff[:3]].tolist()} vs {w[ok][diff[:3]].tolist()}")


def summary(va):
    va = np.asarray(va, np.float32)
    tiny = np.finfo(np.float32).tiny
    sub = np.count_nonzero((va != 0) & (np.abs(va) < tiny))
    return (f"subnormal={sub} inf={np.count_nonzero(np.isinf(va))} "
            f"nan={np.count_nonzero(np.isnan(va))} zero={np.count_nonzero(va == 0)}")


CASES = {
    "uniform": lambda: W.SMALL["uniform"](),
    "rmat14": lambda: W.rmat(scale=14, seed=41),
    "mixed": lambda: W.SMALL["mixed"](),
    "stencil": lambda: W.SMALL["stencil"](),
}
KINDS = {"subnormal": subnormal_values, "special": special_values}


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("case", list(CASES))
def test_special_values_default_bins(case, kind):
    A0, B0 = CASES[case]()
    A = with_values(A0, KINDS[kind](A0.nnz, 1))
    B = A if B0 is A0 else with_values(B0, KINDS[kind](B0.nnz, 2))
    rp, ci, va = pyoracle.spgemm(A, B)
    say(f"{case} {kind}: reference {summary(va)}")
    got = rvsp.spgemm_raw(A, B)
    assert_contract(got, rp, ci, va, f"{case} {kind}")


@pytest.mark.parametrize("kind", list(KINDS))
def test_special_values_every_numeric_bin(kind):
    """Each chooser output (hash / window accumulators, 8..64 lanes) forced
    on every row that admits it."""
    A, _ = W.SMALL["uniform"]()
    A = with_values(A, KINDS[kind](A.nnz, 3))
    rp, ci, va = pyoracle.spgemm(A, A)
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    plan.numeric()
    assert_contract(plan.result(), rp, ci, va, f"default bins {kind}")
    _, nb0 = plan.bins()
    ops, alen, span, nz = S.row_stats(A, A, plan.row_ptr())
    cb = S.chooser_bins(nz, span)
    for o in range(8):
        nb = nb0.copy()
        ok = (cb[:, o] >= 0) & (nb0 != 0) & (nb0 != rvsp.NB_WIN)
        nb[ok] = cb[ok, o]
        assert plan.set_num_bins(nb) == 0
        plan.numeric()
        assert_contract(plan.result(), rp, ci, va, f"chooser output {o} {kind}")
    # the ranked reuse pass on the same values
    plan.prepare_reuse()
    plan.cval.upload(np.full(max(plan.nnz, 1), 1.0, np.float32))
    plan.numeric()
    assert_contract(plan.result(), rp, ci, va, f"ranked pass {kind}")
    plan.close()


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("shape", ["2048x16", "4096x32", "1024x8"])
def test_special_values_big_rows_every_window_shape(kind, shape, monkeypatch):
    monkeypatch.setenv("RVSP_BIG_WINDOW", shape)
    A2 = rvsp.gen_random_rows(40, 3000, 700, 700, 78)
    B2 = rvsp.gen_random_rows(3000, 100000, 250, 350, 77)
    A2 = with_values(A2, KINDS[kind](A2.nnz, 4))
    B2 = with_values(B2, KINDS[kind](B2.nnz, 5))
    rp, ci, va = pyoracle.spgemm(A2, B2)
    say(f"wide rows {shape} {kind}: reference {summary(va)}")
    assert_contract(rvsp.spgemm_raw(A2, B2), rp, ci, va, f"wide {shape} {kind}")
