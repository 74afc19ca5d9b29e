"""Learned selectors, CPU side: the network oracle pinned by finite
differences and closed forms, the host-side candidate table, the parameter
file format, and the library's selector entry points that need no GPU."""
import ctypes as C
import os

import numpy as np
import pytest

import rvsp
from rvsp import selector as S
from oracle import mlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEL_DIR = os.path.join(ROOT, "rv-sparse_amd", "selector")


def rand_case(nf, no, n=37, seed=0, nan_frac=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, mlp.param_count(nf, no))
    f = rng.uniform(0, 1, (n, 5))
    t = rng.normal(0, 1, (n, no))
    t[rng.uniform(size=t.shape) < nan_frac] = np.nan
    return p, f, t


@pytest.mark.parametrize("shape", [mlp.SIZER, mlp.CHOOSER])
def test_oracle_gradient_matches_finite_differences(shape):
    nf, no = shape
    p, f, t = rand_case(nf, no, seed=nf)
    L, g = mlp.loss_grad(p, f, t, nf, no)
    assert L == pytest.approx(mlp.loss(p, f, t, nf, no), rel=1e-12)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp.loss(p + e, f, t, nf, no) - mlp.loss(p - e, f, t, nf, no)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_oracle_gradient_planted_mistake_is_caught():
    """A dropped ReLU mask in the backward pass fails the check above."""
    nf, no = mlp.CHOOSER
    p, f, t = rand_case(nf, no, seed=5)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, nf, no)
    y, (ff, z1, h1, z2, h2) = mlp.forward(p, f, nf, no)
    m = ~np.isnan(t)
    e = np.where(m, y - np.where(m, t, 0), 0)
    dy = 2 * e / y.shape[0]
    d2_wrong = dy @ W3  # mask (z2 > 0) dropped
    gW2_wrong = d2_wrong.T @ h1
    _, g = mlp.loss_grad(p, f, t, nf, no)
    off = mlp.H1 * nf + mlp.H1
    assert not np.allclose(g[off:off + 256], gW2_wrong.ravel(), rtol=1e-5, atol=1e-7)


def test_oracle_forward_closed_forms():
    nf, no = mlp.SIZER
    p = np.zeros(mlp.param_count(nf, no))
    p[-1] = 0.75  # only the output bias: y == b3 everywhere
    y, _ = mlp.forward(p, np.random.default_rng(1).uniform(size=(9, 5)), nf, no)
    np.testing.assert_array_equal(y, 0.75)
    # identity path: W1[0][0] = 1, W2[0][0] = 1, W3[0][0] = 1 -> y = relu(f0)
    p[:] = 0
    p[0] = 1.0
    off2 = mlp.H1 * nf + mlp.H1
    p[off2] = 1.0
    off3 = off2 + 256 + mlp.H2
    p[off3] = 1.0
    f = np.array([[0.3, 0, 0, 0, 0], [-0.2, 0, 0, 0, 0]])
    y, _ = mlp.forward(p, f, nf, no)
    np.testing.assert_allclose(y[:, 0], [0.3, 0.0])


def test_oracle_adam_first_step_is_signed_lr():
    """First Adam step with bias correction moves each parameter by
    lr * g / (|g| + eps), i.e. about lr * sign(g)."""
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m1, m2 = mlp.adam(p, g, 0.01, 0, np.zeros(3), np.zeros(3))
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_features_match_definition():
    f = mlp.features([0, 15, 1023], [0, 3, 1], [0, 7, 65535], [0, 7, 31])
    np.testing.assert_allclose(f[1], [2 / 16, 4 / 32, 3 / 32, np.log2(6) / 16, 3 / 32])
    np.testing.assert_allclose(f[2], [1 / 16, 10 / 32, 16 / 32, 10 / 16, 5 / 32])


def test_param_counts():
    assert mlp.param_count(*mlp.SIZER) == 369
    assert mlp.param_count(*mlp.CHOOSER) == 504
    assert rvsp.selector_param_count(rvsp.SELECTOR_SIZER) == 369
    assert rvsp.selector_param_count(rvsp.SELECTOR_CHOOSER) == 504
    assert rvsp.selector_param_count(7) == -1


def brute_chooser_bin(o, nz, span):
    G = 8 << (o & 3)
    if o < 4:
        if nz > 2730:
            return -1
        H = 64
        while H < (3 * nz + 1) // 2:
            H *= 2
        if G < (32 if H >= 4096 else 16 if H >= 2048 else 8):
            return -1
        return 1 + 4 * (int(np.log2(H)) - 6) + (o & 3)
    C = 512
    while C < nz:
        C *= 2
    rank = 53 + 2 * (int(np.log2(C)) - 9) + (1 if G == 64 else 0)
    rank_ok = span <= 16384 and nz <= 4096 and G >= 32
    dense = -1
    if span <= 16384:
        W = 512
        while W < span:
            W *= 2
        if G >= (64 if W >= 16384 else 32 if W >= 8192 else 16 if W >= 4096 else 8):
            dense = 29 + 4 * (int(np.log2(W)) - 9) + (o & 3)
    # the window with the smaller LDS footprint: rank 512 words x 6 bytes + 4 per
    # output slot, dense 4 per column (the rank window on a tie)
    if rank_ok and (dense < 0 or 512 * 6 + 4 * C <= 4 * W):
        return rank
    return dense


def test_chooser_candidates_match_brute_force():
    rng = np.random.default_rng(3)
    nz = np.concatenate([[1, 42, 43, 682, 683, 1365, 1366, 2730, 2731], rng.integers(1, 4000, 300)])
    span = np.concatenate([[1, 512, 513, 4096, 4097, 8192, 8193, 16384, 16385],
                           rng.integers(1, 20000, 300)])
    got = S.chooser_bins(nz, span)
    for i in range(nz.size):
        for o in range(8):
            assert got[i, o] == brute_chooser_bin(o, int(nz[i]), int(span[i])), (nz[i], span[i], o)


def test_row_stats_brute_force():
    A = rvsp.gen_random_rows(50, 70, 0, 9, 5)
    B = rvsp.gen_random_rows(70, 90, 0, 6, 6)
    ops, alen, span, nz = S.row_stats(A, B, np.zeros(51, np.int32))
    for i in range(A.rows):
        ks = A.col_idx[A.row_ptr[i]:A.row_ptr[i + 1]]
        cols = np.concatenate([B.col_idx[B.row_ptr[k]:B.row_ptr[k + 1]] for k in ks] or [[]])
        assert ops[i] == cols.size and alen[i] == ks.size
        assert span[i] == (cols.max() - cols.min() + 1 if cols.size else 0)


def test_param_file_round_trip(tmp_path):
    p = np.linspace(-1, 1, 369).astype(np.float32)
    path = tmp_path / "s.txt"
    S.write_params(str(path), p, 1.25)
    vals = np.array(open(path).read().split(), np.float64)
    assert vals[0] == 1.25 and vals.size == 370
    np.testing.assert_array_equal(vals[1:].astype(np.float32), p)


def test_load_file_rejects_malformed_without_gpu(tmp_path):
    L = rvsp.lib()
    bad = tmp_path / "bad.txt"
    bad.write_text("1.0 2 3\n")  # too few numbers
    assert L.rvsp_selector_load_file(0, str(bad).encode()) == rvsp.ERROR_INVALID_ARGUMENT
    p = np.zeros(370)
    p[0] = -1.0  # margin must be > 0
    bad.write_text(" ".join(map(str, p)))
    assert L.rvsp_selector_load_file(0, str(bad).encode()) == rvsp.ERROR_INVALID_ARGUMENT
    p[0] = 1.0
    bad.write_text(" ".join(map(str, p)) + " 5")  # one too many
    assert L.rvsp_selector_load_file(0, str(bad).encode()) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_selector_load_file(0, b"/nonexistent/file") == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_selector_load_file(9, str(bad).encode()) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_selector_load_file(0, None) == rvsp.ERROR_NULL_POINTER
    assert L.rvsp_selector_set(0, None, 369, C.c_float(1.0)) == rvsp.ERROR_NULL_POINTER
    q = np.zeros(369, np.float32)
    assert L.rvsp_selector_set(0, q.ctypes.data, 368, C.c_float(1.0)) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_selector_set(1, q.ctypes.data, 369, C.c_float(1.0)) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_selector_set(0, q.ctypes.data, 369, C.c_float(0.0)) == rvsp.ERROR_INVALID_ARGUMENT


@pytest.mark.parametrize("name,count", [("sizer.txt", 369), ("chooser.txt", 504)])
def test_shipped_weights_parse(name, count):
    path = os.path.join(SEL_DIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not trained yet")
    vals = np.array(open(path).read().split(), np.float64)
    assert vals.size == count + 1 and vals[0] > 0 and np.all(np.isfinite(vals))
