"""Numeric passes under a reused structure with per-product output ranks
(rvsp_hip_plan_prepare_reuse): bit-identical to the plain kernels and to
the CPU oracle, under heuristic bins, a random learned chooser, every forced
accumulator, a graph-replayed descriptor, and values rewritten between
passes."""
import ctypes as C
import sys

import numpy as np
import pytest

import rvsp
from rvsp import selector as S
from rvsp import workloads as W
from oracle import pyoracle
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def no_selectors():
    rvsp.selector_clear(rvsp.SELECTOR_SIZER)
    rvsp.selector_clear(rvsp.SELECTOR_CHOOSER)
    yield
    rvsp.selector_clear(rvsp.SELECTOR_SIZER)
    rvsp.selector_clear(rvsp.SELECTOR_CHOOSER)


def pair(wl):
    if wl == "batched0":
        return W.SMALL["batched"]()[0]
    return W.SMALL[wl]()


def bits(x):
    return np.asarray(x, np.float32).view(np.uint32)


@pytest.mark.parametrize("wl", ["stencil", "mixed", "uniform", "rmat", "batched0"])
def test_ranked_pass_matches_plain_and_oracle(wl):
    A, B = pair(wl)
    say(f"ranked pass {wl} m={A.rows}")
    dA = rvsp.DeviceCsr(A)
    dB = dA if B is A else rvsp.DeviceCsr(B)
    plan = rvsp.Plan(dA, dB)
    plan.numeric()
    plain = plan.result()
    plan.prepare_reuse()
    plan.cval.upload(np.full(max(plan.nnz, 1), np.nan, np.float32))  # no stale values
    plan.numeric()
    ranked = plan.result()
    np.testing.assert_array_equal(bits(ranked.values), bits(plain.values))
    rp, ci, va = pyoracle.spgemm(A, B)[:3]
    assert_bitexact(ranked, rp, ci, va, f"{wl} ranked")
    plan.close()


def test_ranked_pass_new_values_each_time():
    """The ranks depend on the structure only: passes with values redrawn
    on the device match the oracle on the new values."""
    A, _ = W.SMALL["stencil"]()
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    plan.numeric()
    plan.prepare_reuse()
    rng = np.random.default_rng(5)
    for it in range(3):
        v = rng.standard_normal(A.nnz).astype(np.float32)
        dA.va.upload(v)
        plan.numeric()
        A2 = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, v)
        rp, ci, va = pyoracle.spgemm(A2, A2)[:3]
        assert_bitexact(plan.result(), rp, ci, va, f"stencil pass {it}")
    plan.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_ranked_pass_with_random_chooser(seed):
    A, B = W.SMALL["mixed"]()
    rvsp.selector_set(rvsp.SELECTOR_CHOOSER, S.init_params(rvsp.SELECTOR_CHOOSER, seed), 1.0)
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    plan = rvsp.Plan(dA, dB, rvsp.SELECT_LEARNED)
    plan.prepare_reuse()
    plan.numeric()
    rp, ci, va = pyoracle.spgemm(A, B)[:3]
    assert_bitexact(plan.result(), rp, ci, va, f"ranked, random chooser {seed}")
    plan.close()


def test_ranked_pass_every_accumulator():
    A, _ = W.SMALL["uniform"]()
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    plan.numeric()
    ref = plan.result()
    plan.prepare_reuse()
    _, nb0 = plan.bins()
    ops, alen, span, nz = S.row_stats(A, A, plan.row_ptr())
    cb = S.chooser_bins(nz, span)
    for o in range(8):
        nb = nb0.copy()
        ok = (cb[:, o] >= 0) & (nb0 != 0) & (nb0 != rvsp.NB_WIN)
        nb[ok] = cb[ok, o]
        assert plan.set_num_bins(nb) == 0
        plan.numeric()
        say(f"ranked output {o}: {int(ok.sum())} rows moved")
        np.testing.assert_array_equal(bits(plan.result().values), bits(ref.values))
    plan.close()


@pytest.mark.parametrize("graph", [0, 1])
def test_descriptor_reuse_uses_ranks_and_stays_exact(graph):
    """Device-memory descriptor: the second compute prepares the ranks; the
    third (graph-replayed or not) must still be bit-exact."""
    A, _ = W.SMALL["stencil"]()
    L = rvsp.lib()
    dA = rvsp.DeviceCsr(A)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_graph(graph) == 0
    a = dA.struct()
    st, ws, nnz = d.estimate(a, a)
    assert st == 0
    crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
    cci = rvsp.DeviceBuffer(4 * max(nnz, 1))
    cva = rvsp.DeviceBuffer(4 * max(nnz, 1))
    c = rvsp.CsrStruct(A.rows, A.cols, nnz, crp.ptr, cci.ptr, cva.ptr, rvsp.DTYPE_FP32, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    rng = np.random.default_rng(9)
    for it in range(4):
        v = rng.standard_normal(A.nnz).astype(np.float32)
        dA.va.upload(v)
        assert d.compute(a, a, c, wsb.ctypes.data) == 0
        assert d.wait() == 0
        got = rvsp.device_csr_to_host(c)
        A2 = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, v)
        rp, ci, va = pyoracle.spgemm(A2, A2)[:3]
        assert_bitexact(got, rp, ci, va, f"descriptor reuse graph={graph} pass {it}")
    d.close()
