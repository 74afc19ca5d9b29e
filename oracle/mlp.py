"""CPU restatement of the learned row selectors' network (float64 numpy).

TEST INFRASTRUCTURE ONLY: the GPU forward, backward and Adam kernels of
rv-sparse_amd/src/kernels/rvsp_selector.hip are checked against this;
nothing in the product imports it. The selectors have no counterpart in
rv-sparse, so there is no reference file to follow; this module is pinned
by central finite differences of its own loss (tests/test_selector_cpu.py).

Network: f (nf) -> 16 ReLU -> 16 ReLU -> no, parameters stored layer by
layer W1[16][nf], b1[16], W2[16][16], b2[16], W3[no][16], b3[no].
Loss: L = (1/n) * sum over samples and measured outputs of (y - t)^2, a
NaN target being unmeasured.
"""
from __future__ import annotations

import numpy as np

H1 = H2 = 16
SIZER = (4, 1)
CHOOSER = (5, 8)


def param_count(nf: int, no: int) -> int:
    return H1 * nf + H1 + H2 * H1 + H2 + no * H2 + no


def unpack(p, nf, no):
    p = np.asarray(p, np.float64)
    i = 0

    def take(k, shape):
        nonlocal i
        v = p[i:i + k].reshape(shape)
        i += k
        return v

    W1 = take(H1 * nf, (H1, nf))
    b1 = take(H1, (H1,))
    W2 = take(H2 * H1, (H2, H1))
    b2 = take(H2, (H2,))
    W3 = take(no * H2, (no, H2))
    b3 = take(no, (no,))
    assert i == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, feats, nf, no):
    W1, b1, W2, b2, W3, b3 = unpack(p, nf, no)
    f = np.asarray(feats, np.float64)[:, :nf]
    z1 = f @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    return y, (f, z1, h1, z2, h2)


def loss(p, feats, target, nf, no):
    y, _ = forward(p, feats, nf, no)
    t = np.asarray(target, np.float64).reshape(y.shape)
    m = ~np.isnan(t)
    e = np.where(m, y - np.where(m, t, 0.0), 0.0)
    return float((e * e).sum() / y.shape[0])


def loss_grad(p, feats, target, nf, no):
    """(L, dL/dp) by backpropagation."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nf, no)
    y, (f, z1, h1, z2, h2) = forward(p, feats, nf, no)
    n = y.shape[0]
    t = np.asarray(target, np.float64).reshape(y.shape)
    m = ~np.isnan(t)
    e = np.where(m, y - np.where(m, t, 0.0), 0.0)
    L = float((e * e).sum() / n)
    dy = 2.0 * e / n
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    d2 = (dy @ W3) * (z2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ f
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def adam(p, grads, lr, steps_done, m1, m2):
    """One Adam step (beta 0.9 / 0.999, eps 1e-8) as the GPU applies it;
    steps_done counts earlier steps. Returns the new (p, m1, m2)."""
    b1, b2, eps = 0.9, 0.999, 1e-8
    t = steps_done + 1
    m1 = b1 * m1 + (1 - b1) * grads
    m2 = b2 * m2 + (1 - b2) * grads * grads
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2


def features(ops, alen, span, nz):
    """The 5 per-row features (float64)."""
    ops = np.asarray(ops, np.float64)
    alen = np.asarray(alen, np.float64)
    span = np.asarray(span, np.float64)
    nz = np.asarray(nz, np.float64)
    return np.stack([
        np.log2(1 + alen) / 16,
        np.log2(1 + ops) / 32,
        np.log2(1 + span) / 32,
        np.log2(1 + ops / np.maximum(alen, 1)) / 16,
        np.log2(1 + nz) / 32,
    ], axis=1)
