"""Every tuning-table entry, and the extremes of every parameter, give C
bit-identical to the oracle: the parameters only move rows between bins and
change how the work is scheduled (tuning/table.txt, rvsp_tuning_t)."""
import numpy as np
import pytest

import rvsp
from rvsp import workloads as W
from oracle import pyoracle
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu


def products():
    """Products that reach every row path: hash and bitmap rows, big rows
    over wide B, exact cancellations."""
    A1, _ = W.rmat(scale=15, seed=91)
    A2 = rvsp.gen_random_rows(40, 3000, 700, 700, 92)
    B2 = rvsp.gen_random_rows(3000, 100000, 250, 350, 93)
    A3, B3 = W.mixed(m=1 << 13, k=1 << 9, n=1 << 15, maxlen=256, seed_a=94, seed_b=95)
    A4, _ = W.stencil(g=12, seed=96)
    A5, _ = W.uniform(n=1024, per_row=16, seed=97)
    return [("rmat-15", A1, A1), ("wide", A2, B2), ("mixed", A3, B3), ("stencil", A4, A4),
            ("uniform", A5, A5)]


@pytest.fixture(scope="module")
def cases():
    out = []
    for name, A, B in products():
        out.append((name, A, B, pyoracle.spgemm(A, B)))
    return out


def run_all(cases, entry, what):
    rvsp.tuning_force(entry)
    try:
        for name, A, B, (rp, ci, va) in cases:
            got = rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B), algo=rvsp.ALGO_HIP)
            assert_bitexact(got, rp, ci, va, f"{what}: {name}")
    finally:
        rvsp.tuning_force(None)


@pytest.mark.parametrize("cls", range(rvsp.TUNING_CLASSES))
def test_every_table_entry_is_bit_exact(cases, cls):
    entry = rvsp.tuning_get(cls)
    print(f"class {cls}: {entry}", flush=True)
    run_all(cases, entry, f"class {cls}")


EXTREMES = [
    {"hash_max_u": 64},
    {"hash_max_u": 256, "seg_products": 1024},
    {"big_window_outputs": 1024, "big_window_chunks": 8, "sym_big_threads": 256},
    {"big_window_outputs": 4096, "big_window_chunks": 32, "sym_big_threads": 1024,
     "seg_products": 1 << 22},
    {"fan_min_products": 0, "fan_min_fill": 0},
    {"fan_min_products": 1 << 62, "fan_min_fill": 1 << 62},
]


@pytest.mark.parametrize("change", EXTREMES, ids=[str(e) for e in EXTREMES])
def test_parameter_extremes_are_bit_exact(cases, change):
    run_all(cases, dict(rvsp.tuning_defaults(), **change), str(change))


def test_table_entry_applies_to_its_class(cases):
    """A class entry set through the API is what the analysis uses for a
    product of that class (a non-default entry still gives the same C)."""
    name, A, B, (rp, ci, va) = cases[0]
    cls = rvsp.tuning_class(A, B)
    before = rvsp.tuning_get(cls)
    try:
        t = rvsp.Tuning.of(dict(before, hash_max_u=128, seg_products=2048))
        assert rvsp.lib().rvsp_tuning_set(cls, t) == 0
        got = rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B), algo=rvsp.ALGO_HIP)
        assert_bitexact(got, rp, ci, va, f"class {cls} entry set")
    finally:
        assert rvsp.lib().rvsp_tuning_set(cls, rvsp.Tuning.of(before)) == 0


def test_narrow_band_rows_over_the_hash_limit():
    """Rows of more candidate columns than the hash limit, too wide for the
    one-pass bitmap (16384 columns) but spanning at most 64 columns a
    candidate, stay in the hash tables (interior rows of a 27-point stencil
    64^3 times the whole stencil: 729 candidates over about 16.6 k
    columns); C is the oracle's, bit for bit, with the limit below and
    above their candidate count."""
    S, _ = W.stencil(g=64, seed=98)
    # interior rows of the z = 32 plane (x, y in 2..61: all 125 neighbours)
    rows = np.array([32 * 4096 + y * 64 + x for y in range(2, 10) for x in range(2, 62)])
    parts = [np.arange(S.row_ptr[r], S.row_ptr[r + 1]) for r in rows]
    at = np.concatenate(parts)
    rptr = np.zeros(len(rows) + 1, np.int32)
    rptr[1:] = np.cumsum([len(q) for q in parts])
    A = rvsp.Csr(len(rows), S.cols, rptr, np.ascontiguousarray(S.col_idx[at]),
                 np.ascontiguousarray(S.values[at]))
    # the rows reach the span test: u = min(products, span) > 512,
    # 16384 < span <= 64 u
    blen = np.diff(S.row_ptr)
    bfirst, blast = S.col_idx[S.row_ptr[:-1]], S.col_idx[S.row_ptr[1:] - 1]
    for r in range(A.rows):
        ks = A.col_idx[A.row_ptr[r]:A.row_ptr[r + 1]]
        ops, span = int(blen[ks].sum()), int(blast[ks].max() - bfirst[ks].min() + 1)
        u = min(ops, span)
        assert u > 512 and 16384 < span <= 64 * u, (r, ops, span)
    rp, ci, va = pyoracle.spgemm(A, S)[:3]
    for hmax in (256, 512, 2048):
        print(f"stencil band rows, hash limit {hmax}", flush=True)
        run_all([("stencil band", A, S, (rp, ci, va))], dict(rvsp.tuning_defaults(), hash_max_u=hmax),
                f"hash limit {hmax}")
