"""Learned selectors on the GPU: the forward/backward/Adam kernels against
the CPU oracle (oracle/mlp.py), determinism of training, and that learned
table sizing and learned accumulator choice never change a result (bit
for bit against the heuristic path and the CPU oracle), including a sizer
planted to underestimate every row, which forces the exact-count
fallback."""
import os
import sys

import numpy as np
import pytest

import rvsp
from rvsp import selector as S
from rvsp import workloads as W
from oracle import mlp, pyoracle
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu
SEL_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "rv-sparse_amd", "selector")


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


def rand_case(nf, no, n, seed, nan_frac=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, mlp.param_count(nf, no)).astype(np.float32)
    f = rng.uniform(0, 1, (n, 5)).astype(np.float32)
    t = rng.normal(0, 1, (n, no)).astype(np.float32)
    t[rng.uniform(size=t.shape) < nan_frac] = np.nan
    return p, f, t


@pytest.mark.parametrize("which", [rvsp.SELECTOR_SIZER, rvsp.SELECTOR_CHOOSER])
@pytest.mark.parametrize("n", [1, 127, 1000])
def test_gradient_matches_oracle(which, n):
    nf, no = S.NF_NO[which]
    say(f"grad which={which} n={n}")
    p, f, t = rand_case(nf, no, n, seed=10 * which + n)
    fd, td = rvsp.DeviceBuffer.from_array(f), rvsp.DeviceBuffer.from_array(t)
    g, L = rvsp.selector_grad(which, fd, td, n, p)
    Lr, gr = mlp.loss_grad(p.astype(np.float64), f, t, nf, no)
    assert L == pytest.approx(Lr, rel=1e-4)
    scale = np.abs(gr).max()
    np.testing.assert_allclose(g, gr, rtol=1e-3, atol=1e-5 * scale)


@pytest.mark.parametrize("which", [rvsp.SELECTOR_SIZER, rvsp.SELECTOR_CHOOSER])
def test_predict_matches_oracle(which):
    nf, no = S.NF_NO[which]
    p, f, _ = rand_case(nf, no, 777, seed=3)
    y = rvsp.selector_predict(which, rvsp.DeviceBuffer.from_array(f), 777, p, no)
    yr, _ = mlp.forward(p.astype(np.float64), f, nf, no)
    np.testing.assert_allclose(y, yr, rtol=1e-4, atol=1e-5)


def test_one_adam_step_matches_oracle():
    """epochs=1 with one minibatch holding every sample = exactly one Adam
    step on the full-batch gradient."""
    nf, no = mlp.CHOOSER
    n = 256
    p, f, t = rand_case(nf, no, n, seed=21)
    fd, td = rvsp.DeviceBuffer.from_array(f), rvsp.DeviceBuffer.from_array(t)
    p1, _ = rvsp.selector_train(rvsp.SELECTOR_CHOOSER, fd, td, n, p, epochs=1, batch=n, lr=1e-2,
                                seed=4)
    _, g = mlp.loss_grad(p.astype(np.float64), f, t, nf, no)
    pr, _, _ = mlp.adam(p.astype(np.float64), g, 1e-2, 0, np.zeros(p.size), np.zeros(p.size))
    # parameters whose gradient is tiny move by lr*g/(|g|+eps): compare where |g| is not
    big = np.abs(g) > 1e-4 * np.abs(g).max()
    np.testing.assert_allclose(p1[big], pr[big], rtol=1e-4, atol=1e-5)


def test_training_deterministic_and_learns():
    """Training twice from one seed gives identical parameters, and fits a
    known target function far better than the starting point."""
    rng = np.random.default_rng(8)
    n = 5000
    f = rng.uniform(0, 1, (n, 5)).astype(np.float32)
    t = (-(f[:, 1] - f[:, 2]) * 4 + 0.5 * f[:, 0]).astype(np.float32)[:, None]
    fd, td = rvsp.DeviceBuffer.from_array(f), rvsp.DeviceBuffer.from_array(t)
    p0 = S.init_params(rvsp.SELECTOR_SIZER, 1)
    _, L0 = rvsp.selector_grad(rvsp.SELECTOR_SIZER, fd, td, n, p0)
    pa, La = rvsp.selector_train(rvsp.SELECTOR_SIZER, fd, td, n, p0, epochs=30, batch=512, lr=3e-3)
    pb, Lb = rvsp.selector_train(rvsp.SELECTOR_SIZER, fd, td, n, p0, epochs=30, batch=512, lr=3e-3)
    np.testing.assert_array_equal(pa, pb)
    _, L1 = rvsp.selector_grad(rvsp.SELECTOR_SIZER, fd, td, n, pa)
    say(f"loss {L0:.4f} -> {L1:.5f}")
    assert L1 < 0.05 * L0


def learned_vs_oracle(A, B, what):
    dA = rvsp.DeviceCsr(A)
    dB = dA if B is A else rvsp.DeviceCsr(B)
    C_ = rvsp.spgemm_device(dA, dB, algo=rvsp.ALGO_HIP_LEARNED)
    rp, ci, va = pyoracle.spgemm(A, B)[:3]
    assert_bitexact(C_, rp, ci, va, what)


def sizer_params(bias):
    """A sizer that predicts log2(nnz/u) = bias for every row."""
    p = np.zeros(369, np.float32)
    p[-1] = bias
    return p


@pytest.mark.parametrize("wl", ["rmat", "mixed", "uniform", "stencil"])
def test_underestimating_sizer_falls_back_exactly(wl):
    """Every row's table sized for 1/64 of its bound: rows with more than
    that many columns overflow, are recounted in their heuristic bin, and
    the product still matches the oracle bit for bit."""
    A, B = W.SMALL[wl]()
    say(f"underestimating sizer {wl} m={A.rows}")
    rvsp.selector_set(rvsp.SELECTOR_SIZER, sizer_params(-6.0), 1.0)
    dA = rvsp.DeviceCsr(A)
    dB = dA if B is A else rvsp.DeviceCsr(B)
    plan = rvsp.Plan(dA, dB, rvsp.SELECT_LEARNED)
    moved, back = plan.selector_stats()
    say(f"  rows kept in learned tables {moved}, recounted {back}")
    assert back > 0
    plan.numeric()
    rp, ci, va = pyoracle.spgemm(A, B)[:3]
    assert_bitexact(plan.result(), rp, ci, va, f"{wl} fallback")
    plan.close()


@pytest.mark.parametrize("wl", ["rmat", "mixed", "batched0", "uniform"])
def test_exact_sizer_keeps_every_row(wl):
    """A margin of 1 with a sizer that predicts 0 (nnz = u, the bound)
    never overflows; the BMS rows it moves into small tables fill right."""
    A, B = W.SMALL[wl]() if wl != "batched0" else W.SMALL["batched"]()[0]
    rvsp.selector_set(rvsp.SELECTOR_SIZER, sizer_params(0.0), 1.0)
    dA = rvsp.DeviceCsr(A)
    dB = dA if B is A else rvsp.DeviceCsr(B)
    plan = rvsp.Plan(dA, dB, rvsp.SELECT_LEARNED)
    moved, back = plan.selector_stats()
    say(f"exact sizer {wl}: moved {moved}, recounted {back}")
    assert back == 0
    plan.numeric()
    rp, ci, va = pyoracle.spgemm(A, B)[:3]
    assert_bitexact(plan.result(), rp, ci, va, f"{wl} exact sizer")
    plan.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_chooser_never_changes_values(seed):
    """Whatever accumulator a (random) chooser picks, values are the same."""
    A, B = W.SMALL["mixed"]()
    rvsp.selector_set(rvsp.SELECTOR_CHOOSER, S.init_params(rvsp.SELECTOR_CHOOSER, seed), 1.0)
    learned_vs_oracle(A, B, f"random chooser {seed}")


def test_every_accumulator_gives_identical_values():
    """Force each of the 8 chooser outputs on every row that admits it."""
    A, B = W.SMALL["uniform"]()
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    plan.numeric()
    ref = plan.result()
    _, nb0 = plan.bins()
    ops, alen, span, nz = S.row_stats(A, A, plan.row_ptr())
    cb = S.chooser_bins(nz, span)
    for o in range(8):
        nb = nb0.copy()
        ok = (cb[:, o] >= 0) & (nb0 != 0) & (nb0 != rvsp.NB_WIN)
        nb[ok] = cb[ok, o]
        assert plan.set_num_bins(nb) == 0
        plan.numeric()
        got = plan.result()
        say(f"output {o}: {int(ok.sum())} rows moved")
        np.testing.assert_array_equal(got.values.view(np.uint32), ref.values.view(np.uint32))
    plan.close()


def test_set_num_bins_rejects_bad_bins():
    A, B = W.SMALL["rmat"]()
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    sb, nb = plan.bins()
    bad = nb.copy()
    bad[nb == 0] = rvsp.NB_HASH0  # an empty row given a kernel
    if (nb == 0).any():
        assert plan.set_num_bins(bad) == rvsp.ERROR_INVALID_ARGUMENT
    bad = nb.copy()
    bad[nb != 0] = 200
    assert plan.set_num_bins(bad) == rvsp.ERROR_INVALID_ARGUMENT
    plan.close()


def test_time_bin_rejects_rows_that_do_not_fit():
    """A 64-slot table for rows with hundreds of columns raises the
    consistency flag; the call reports it and the plan stays usable."""
    A, B = W.SMALL["uniform"]()
    dA = rvsp.DeviceCsr(A)
    plan = rvsp.Plan(dA, dA)
    plan.fill()
    rows = rvsp.DeviceBuffer.from_array(np.arange(64, dtype=np.int32))
    with pytest.raises(rvsp.RvspError):
        plan.time_bin(rvsp.NB_HASH0, rows, 64)
    plan.numeric()
    rp, ci, va = pyoracle.spgemm(A, A)[:3]
    assert_bitexact(plan.result(), rp, ci, va, "after rejected timing")
    plan.close()


@pytest.mark.parametrize("name", ["sizer.txt", "chooser.txt"])
def test_shipped_weights_load_and_stay_exact(name):
    path = os.path.join(SEL_DIR, name)
    if not os.path.exists(path):
        pytest.skip("not trained yet")
    which = rvsp.SELECTOR_SIZER if name == "sizer.txt" else rvsp.SELECTOR_CHOOSER
    rvsp.selector_load(which, path)
    for wl in ("rmat", "mixed"):
        A, B = W.SMALL[wl]()
        learned_vs_oracle(A, B, f"shipped {name} {wl}")


# products the shipped chooser never saw: the same families as its training
# set (rv-sparse's generator, the mixed power-law x banded product) at seeds
# that neither the training products (rvsp/selector.py) nor the benchmark use
HELD_OUT = {
    "gen-4096-d0.002-s7": lambda: rvsp.gen_bench_pair(4096, 4096, 0.002, 7),
    "gen-2048-d0.005-s8": lambda: rvsp.gen_bench_pair(2048, 2048, 0.005, 8),
    "mixed-2^17-s90": lambda: W.mixed(m=1 << 17, k=1 << 13, n=1 << 17, seed_a=90, seed_b=91),
}


def numeric_ms(dA, dB, algos, runs=15):
    """Median numeric-phase time (HIP events) of `runs` device one-shots
    per algo, the algos alternating call by call (so clock or cache drift
    during the measurement affects both alike), after one untimed call
    each."""
    for algo in algos:
        rvsp.spgemm_device(dA, dB, algo=algo)
    ts = {algo: [] for algo in algos}
    for _ in range(runs):
        for algo in algos:
            rvsp.reset_timing()
            rvsp.set_timing(True)
            rvsp.spgemm_device(dA, dB, algo=algo)
            rvsp.set_timing(False)
            ms, calls = rvsp.get_timing()["numeric"]
            assert calls == 1
            ts[algo].append(ms)
    return [float(np.median(ts[algo])) for algo in algos]


@pytest.mark.parametrize("name", list(HELD_OUT))
def test_shipped_chooser_gain_on_held_out_products(name):
    """What the trained chooser gains over the heuristic bins it replaces:
    the numeric pass of held-out products with the shipped weights against
    the same pass with heuristic bins. The chooser may not lose beyond timing
    noise (at most 5 % slower: where both pick nearly the same bins, as on
    the mixed product, medians of a 0.5 ms pass differ by up to about 2 %
    from run to run) and may not change a bit of C."""
    path = os.path.join(SEL_DIR, "chooser.txt")
    if not os.path.exists(path):
        pytest.skip("no chooser weights shipped")
    rvsp.selector_load(rvsp.SELECTOR_CHOOSER, path)
    A, B = HELD_OUT[name]()
    dA = rvsp.DeviceCsr(A)
    dB = dA if B is A else rvsp.DeviceCsr(B)
    plain = rvsp.spgemm_device(dA, dB, algo=rvsp.ALGO_HIP)
    learned = rvsp.spgemm_device(dA, dB, algo=rvsp.ALGO_HIP_LEARNED)
    np.testing.assert_array_equal(learned.row_ptr, plain.row_ptr)
    np.testing.assert_array_equal(learned.col_idx, plain.col_idx)
    np.testing.assert_array_equal(learned.values.view(np.uint32), plain.values.view(np.uint32))
    t_plain, t_learned = numeric_ms(dA, dB, [rvsp.ALGO_HIP, rvsp.ALGO_HIP_LEARNED])
    say(f"[gain] {name}: numeric plain {t_plain:.4f} ms, learned {t_learned:.4f} ms, "
        f"learned speedup {t_plain / t_learned:.3f}x")
    assert t_learned <= 1.05 * t_plain, f"{name}: learned {t_learned:.4f} ms > plain {t_plain:.4f} ms"
