"""Seeded randomized parity: forty products whose structure comes from the
library's genmat with parameters drawn per case (square and rectangular,
banded, column-wise degrees, log-normal and normal degree spreads, imbalance
caps), signed U[-1,1) FP32 values so that sums cancel, run through the host
one-shot (heuristic bins) and the device one-shot (learned bins) and
compared bit for bit with the oracle. The cases print their parameters, so
a failure names its case."""
import numpy as np
import pytest

import rvsp
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def gen_csr(rows, cols, rng, **kw):
    got = rvsp.genmat(rows, cols, **kw)
    assert got is not None, (rows, cols, kw)
    r, c, rp, ci, _ = got
    va = rng.uniform(-1, 1, ci.size).astype(np.float32)
    return rvsp.Csr(r, c, rp, ci, va)


def draw_params(rng, rows, cols):
    kw = dict(density=float(rng.choice([0.002, 0.01, 0.05, 0.2])),
              cv=float(rng.choice([0.1, 0.5, 1.0])), min=int(rng.integers(1, 3)),
              random_seed=int(rng.integers(1, 1 << 30)))
    if rng.random() < 0.3:
        kw["imbalance"] = float(rng.choice([1.0, 4.0, 50.0]))
    if rng.random() < 0.25:
        kw["is_column"] = 1
    if rows == cols and rng.random() < 0.4:
        kw["low_bandwidth"] = int(rng.integers(1, rows + 1))
        kw["up_bandwidth"] = int(rng.integers(1, rows + 1))
    return kw


def assert_same(C, want, what):
    rp, ci, va = want[:3]
    np.testing.assert_array_equal(C.row_ptr, rp, err_msg=what)
    np.testing.assert_array_equal(C.col_idx, ci, err_msg=what)
    np.testing.assert_array_equal(C.values.view(np.uint32), np.asarray(va, np.float32).view(np.uint32),
                                  err_msg=what)


@pytest.mark.parametrize("case", range(40))
def test_random_generated_products(case):
    rng = np.random.default_rng(1000 + case)
    m = int(rng.integers(1, 3000))
    k = m if rng.random() < 0.5 else int(rng.integers(1, 3000))
    n = int(rng.integers(1, 6000))
    # at most 4e7 products, so the oracle answers in about a second: denser
    # draws are thinned until the pair fits
    for _ in range(8):
        ka, kb = draw_params(rng, m, k), draw_params(rng, k, n)
        A, B = gen_csr(m, k, rng, **ka), gen_csr(k, n, rng, **kb)
        if int(np.asarray(pyoracle.symbolic(A, B)[1]).sum()) <= 4e7:  # products (>= nnz)
            break
        m, n = max(1, m // 2), max(1, n // 2)
    what = f"case {case}: A {m}x{k} {ka}, B {k}x{n} {kb}"
    print(what, flush=True)
    want = pyoracle.spgemm(A, B)
    st, C1 = rvsp.spgemm_csr(A, B, raise_on_error=False)
    assert st == rvsp.SUCCESS, what
    assert_same(C1, want, what + " (host one-shot)")
    C2 = rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B), algo=rvsp.ALGO_HIP_LEARNED)
    assert_same(C2, want, what + " (device one-shot, learned)")
