"""Training and evaluation of the learned row selectors.

The networks, their forward/backward/Adam kernels and the selection logic
live in the C library (rvsp_selector.hip); this module is the host-side
loop that feeds them:

  sizer    samples = rows of training products, target = the exact
           log2(nnz(C_i) / min(ops_i, span_i)) the symbolic pass produced;
           the margin is then picked by timing the learned analysis.
  chooser  rows are grouped by their features (log2 of A-row length,
           products, span and nnz(C_i), in half-octave steps, plus the
           table and window sizes the row needs); every feasible numeric
           accumulator is timed on each group with HIP events
           (rvsp_hip_plan_time_bin), and the samples' targets are the
           measured log2(ns per row), unmeasured ones NaN; the margin is
           then picked by timing whole numeric passes.

Training products are generated with seeds and sizes distinct from the
benchmark workloads. `python -m rvsp.selector train --out DIR` writes
DIR/sizer.txt and DIR/chooser.txt (the margin, then the parameters), the
format rvsp_selector_load_file reads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import (ALGO_HIP, DeviceBuffer, DeviceCsr, NB_DENSE0, NB_HASH0, NB_RANK0, NB_WIN, Plan,
               SELECT_HEURISTIC, SELECT_LEARNED, SELECTOR_CHOOSER, SELECTOR_SIZER, Csr,
               gen_banded, gen_random_rows, gen_rmat, gen_stencil27, gen_zipf_rows,
               get_timing, reset_timing, selector_clear, selector_param_count,
               selector_predict, selector_set, selector_train, set_timing)

NF_NO = {SELECTOR_SIZER: (4, 1), SELECTOR_CHOOSER: (5, 8)}
HASH_MAX_NNZ = 2730
DENSE_MAX_COLS = 16384
RANK_MAX_COLS = 16384
RANK_MAX_NNZ = 4096
RANK_WORDS = RANK_MAX_COLS // 32


def training_products():
    """(name, A, B) host products: the benchmark families and rv-sparse's
    generated inputs, at other seeds and sizes."""
    out = []
    for scale, ef, seed in ((15, 16, 101), (16, 8, 102), (14, 32, 103)):
        A = gen_rmat(scale, ef << scale, 0.57, 0.19, 0.19, seed)
        out.append((f"rmat{scale}x{ef}", A, A))
    for m, k, n, seed in ((1 << 17, 1 << 13, 1 << 17, 111), (1 << 16, 1 << 14, 1 << 18, 113)):
        A = gen_zipf_rows(m, k, 1.2, 8.0, 4096, seed, normal=True)
        B = gen_banded(k, n, 12, max(1, n // k), 16, seed + 1, normal=True)
        out.append((f"mixed{m}", A, B))
    for n, lo, hi, seed in ((8192, 4, 64, 121), (65536, 4, 64, 123), (2048, 16, 128, 125),
                            (16384, 1, 16, 127)):
        A = gen_random_rows(n, n, lo, hi, seed)
        B = gen_random_rows(n, n, lo, hi, seed + 1)
        out.append((f"rows{n}_{lo}_{hi}", A, B))
    A = gen_stencil27(32, 131, normal=True)
    out.append(("stencil32", A, A))
    # rv-sparse's own generated inputs (log-normal degrees, near-dense rows),
    # at other seeds than its bench sweep
    from . import gen_bench_pair
    for r, d, seed in ((4096, 0.002, 141), (2048, 0.005, 143), (1024, 0.01, 145), (512, 0.02, 147)):
        A, B = gen_bench_pair(r, r, d, seed)
        out.append((f"genmat{r}", A, B))
    return out


def init_params(which: int, seed: int) -> np.ndarray:
    """He-initialised weights, zero biases (deterministic)."""
    nf, no = NF_NO[which]
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / nf), 16 * nf), np.zeros(16),
             rng.normal(0, np.sqrt(2.0 / 16), 256), np.zeros(16),
             rng.normal(0, np.sqrt(1.0 / 16), no * 16), np.zeros(no)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == selector_param_count(which)
    return p


def row_stats(A: Csr, B: Csr, crp: np.ndarray):
    """Per-row (ops, alen, span, nnz) on the host (numpy)."""
    arp, aci = A.row_ptr.astype(np.int64), A.col_idx
    brp, bci = B.row_ptr.astype(np.int64), B.col_idx
    alen = np.diff(arp)
    blen = np.diff(brp)
    m = A.rows
    ops = np.zeros(m, np.int64)
    lo = np.full(m, np.iinfo(np.int64).max, np.int64)
    hi = np.full(m, -1, np.int64)
    if A.nnz:
        rows = np.repeat(np.arange(m), alen)
        bl = blen[aci]
        np.add.at(ops, rows, bl)
        has = bl > 0
        first = np.where(has, bci[np.minimum(brp[aci], max(B.nnz - 1, 0))], lo[0])
        last = np.where(has, bci[np.maximum(brp[aci + 1] - 1, 0)], -1)
        np.minimum.at(lo, rows[has], first[has])
        np.maximum.at(hi, rows[has], last[has])
    span = np.where(ops > 0, hi - lo + 1, 0)
    nz = np.diff(crp.astype(np.int64))
    return ops, alen, span, nz


def _next_pow2(v):
    v = np.maximum(np.asarray(v, np.int64), 1)
    return np.left_shift(1, np.ceil(np.log2(v)).astype(np.int64))


def chooser_bins(nz, span):
    """(m, 8) numeric bin of each chooser output per row, -1 if infeasible;
    the host image of chooser_bin in rvsp_mlp.hpp (outputs 0-3 hash table,
    4-7 window: a rank window where it fits with 32/64 lanes and its LDS
    footprint is no larger than a dense window's, else dense)."""
    nz = np.asarray(nz, np.int64)
    span = np.asarray(span, np.int64)
    H = np.maximum(64, _next_pow2((3 * nz + 1) // 2))
    W = np.maximum(512, _next_pow2(np.minimum(span, 1 << 30)))
    lh = np.log2(H).astype(np.int64)
    lw = np.log2(W).astype(np.int64)
    out = np.full((nz.size, 8), -1, np.int64)
    for o in range(8):
        G = 8 << (o & 3)
        if o < 4:
            gmin = np.where(H >= 4096, 32, np.where(H >= 2048, 16, 8))
            ok = (nz <= HASH_MAX_NNZ) & (G >= gmin)
            out[:, o] = np.where(ok, NB_HASH0 + 4 * (lh - 6) + (o & 3), -1)
        else:
            gmin = np.where(W >= 16384, 64, np.where(W >= 8192, 32, np.where(W >= 4096, 16, 8)))
            ok = (span <= DENSE_MAX_COLS) & (G >= gmin)
            dense = np.where(ok, NB_DENSE0 + 4 * (lw - 9) + (o & 3), -1)
            C = np.maximum(512, _next_pow2(nz))
            lc = np.log2(C).astype(np.int64)
            rank_ok = (span <= RANK_MAX_COLS) & (nz <= RANK_MAX_NNZ) & (G >= 32)
            # the smaller LDS footprint, as the heuristic decides it
            rank_first = rank_ok & (~ok | (RANK_WORDS * 6 + 4 * C <= 4 * W))
            out[:, o] = np.where(rank_first, NB_RANK0 + 2 * (lc - 9) + (1 if G == 64 else 0), dense)
    return out


class Product:
    """Device operands, a plan, and per-row statistics of one product."""

    def __init__(self, name, A, B, select=SELECT_HEURISTIC):
        self.name, self.A, self.B = name, A, B
        self.dA = DeviceCsr(A)
        self.dB = self.dA if B is A else DeviceCsr(B)
        self.plan = Plan(self.dA, self.dB, select)
        self.plan.fill()
        crp = self.plan.row_ptr()
        self.ops, self.alen, self.span, self.nz = row_stats(A, B, crp)

    def close(self):
        self.plan.close()


# ---------------------------------------------------------------------------
# sizer
# ---------------------------------------------------------------------------
def sizer_samples(prods, max_per_product=200_000, seed=7):
    rng = np.random.default_rng(seed)
    F, T = [], []
    for p in prods:
        fd, rd = p.plan.features()
        m = p.A.rows
        f = fd.to_array(np.float32, 5 * m).reshape(m, 5)
        r = rd.to_array(np.float32, m)
        keep = np.nonzero(~np.isnan(r))[0]
        if keep.size > max_per_product:
            keep = np.sort(rng.choice(keep, max_per_product, replace=False))
        F.append(f[keep])
        T.append(r[keep])
    return np.concatenate(F), np.concatenate(T)


def train_net(which, F, T, epochs, batch, lr, seed):
    n = F.shape[0]
    fd = DeviceBuffer.from_array(np.ascontiguousarray(F, np.float32))
    td = DeviceBuffer.from_array(np.ascontiguousarray(T, np.float32))
    p0 = init_params(which, seed)
    p, loss = selector_train(which, fd, td, n, p0, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return p, loss


def structure_ms(prods, select, reps=3):
    """Summed milliseconds (HIP events) of analysis plus structure fill
    over the products, best of reps each."""
    set_timing(True)
    total = 0.0
    for p in prods:
        best = None
        for _ in range(reps):
            reset_timing()
            p.plan.analyse(select)
            p.plan.fill()
            t = get_timing()
            ms = t["analyse"][0] + t["fill"][0]
            best = ms if best is None else min(best, ms)
        total += best
    set_timing(False)
    for p in prods:
        p.plan.analyse(SELECT_HEURISTIC)
        p.plan.fill()
    return total


def passes_ms(prods, select, reps=5):
    """Milliseconds (HIP events) of each product's numeric pass with
    heuristic or learned numeric bins, best of reps each."""
    for p in prods:
        p.plan.analyse(select)
        p.plan.fill()
    set_timing(True)
    out = []
    for p in prods:
        best = None
        for _ in range(reps):
            reset_timing()
            p.plan.numeric()
            ms = get_timing()["numeric"][0]
            best = ms if best is None else min(best, ms)
        out.append(best)
    set_timing(False)
    return np.array(out)


# ---------------------------------------------------------------------------
# chooser
# ---------------------------------------------------------------------------
def buckets(p: Product, min_rows=4):
    """Groups of non-big rows with products: {key: row array}."""
    sb, nb = p.plan.bins()
    ok = (nb != 0) & (nb != NB_WIN)
    idx = np.nonzero(ok)[0]
    q = lambda v: np.floor(2 * np.log2(1 + v[idx].astype(np.float64))).astype(np.int64)
    cb = chooser_bins(p.nz[idx], p.span[idx])
    hk = np.where(cb[:, 3] >= 0, cb[:, 3], -1)
    wk = np.where(cb[:, 7] >= 0, cb[:, 7], -1)
    keys = np.stack([q(p.alen), q(p.ops), q(p.span), q(p.nz), hk, wk], axis=1)
    uk, inv = np.unique(keys, axis=0, return_inverse=True)
    inv = inv.ravel()
    order = np.argsort(inv, kind="stable")
    bounds = np.searchsorted(inv[order], np.arange(uk.shape[0] + 1))
    out = {}
    for j in range(uk.shape[0]):
        rows = idx[order[bounds[j]:bounds[j + 1]]]
        if rows.size >= min_rows:
            out[tuple(uk[j])] = rows
    return out, nb


def time_bucket(p: Product, rows, cands, target_products=32e6, max_rows=65536, reps=3):
    """ns per row of each candidate bin over the bucket's rows (repeated
    to fill the GPU): {bin: ns}."""
    mean_ops = max(float(p.ops[rows].mean()), 1.0)
    n = int(min(max_rows, max(rows.size, 4096, np.ceil(target_products / mean_ops))))
    n = max(n, rows.size)
    rep = np.resize(rows.astype(np.int32), n)
    d = DeviceBuffer.from_array(rep)
    out = {}
    for b in cands:
        ms = p.plan.time_bin(int(b), d, n, reps=reps)
        out[int(b)] = ms * 1e6 / n
    d.free()
    return out


def chooser_samples(prods, per_bucket=256, seed=11, log=None):
    """Features, targets (n x 8, NaN unmeasured) and the bucket table."""
    rng = np.random.default_rng(seed)
    F, T, gid, table = [], [], [], []
    for p in prods:
        bk, nb = buckets(p)
        fd, _ = p.plan.features()
        m = p.A.rows
        f = fd.to_array(np.float32, 5 * m).reshape(m, 5)
        for key, rows in bk.items():
            cb = chooser_bins(p.nz[rows[:1]], p.span[rows[:1]])[0]
            cands = sorted({int(b) for b in cb if b >= 0})
            if not cands:
                continue
            t = time_bucket(p, rows, cands)
            tg = np.full(8, np.nan, np.float32)
            for o in range(8):
                if cb[o] >= 0:
                    tg[o] = np.log2(t[int(cb[o])])
            pick = rows if rows.size <= per_bucket else rng.choice(rows, per_bucket, replace=False)
            F.append(f[pick])
            T.append(np.repeat(tg[None, :], pick.size, axis=0))
            gid.append(np.full(pick.size, len(table), np.int64))
            heur = np.bincount(nb[rows], minlength=64)
            table.append({"product": p.name, "key": [int(x) for x in key], "rows": int(rows.size),
                          "ns": {str(b): v for b, v in t.items()},
                          "bins": [int(b) for b in cb],
                          "heuristic_bin": int(np.argmax(heur))})
        if log:
            log(f"chooser timings: {p.name}: {len(bk)} groups")
    return np.concatenate(F), np.concatenate(T), np.concatenate(gid), table


def group_totals(table, F, gid, params):
    """Measured microseconds over the timed row groups if every row took
    the heuristic's bin, the chooser's pick (majority over the group's
    samples) or the group's fastest bin."""
    n = F.shape[0]
    pred = selector_predict(SELECTOR_CHOOSER, DeviceBuffer.from_array(np.ascontiguousarray(F)), n,
                            params, 8)
    tot = {"heuristic": 0.0, "chooser": 0.0, "best": 0.0}
    for j, g in enumerate(table):
        ns = {int(k): v for k, v in g["ns"].items()}
        bins = np.array(g["bins"])
        y = np.where(bins[None, :] >= 0, pred[gid == j], np.inf)
        pick = int(bins[np.bincount(np.argmin(y, axis=1), minlength=8).argmax()])
        h = g["heuristic_bin"]
        if h not in ns:
            continue
        tot["heuristic"] += ns[h] * g["rows"] / 1e3
        tot["chooser"] += ns[pick] * g["rows"] / 1e3
        tot["best"] += min(ns.values()) * g["rows"] / 1e3
    return tot


def write_params(path, params, margin):
    with open(path, "w") as fh:
        fh.write(f"{margin:.6g}\n")
        for i in range(0, params.size, 8):
            fh.write(" ".join(f"{v:.9g}" for v in params[i:i + 8]) + "\n")


def train(out_dir, epochs=40, chooser_epochs=400, seed=1, log=print):
    os.makedirs(out_dir, exist_ok=True)
    t0 = time.time()
    selector_clear(SELECTOR_SIZER)
    selector_clear(SELECTOR_CHOOSER)
    prods = [Product(n, A, B) for n, A, B in training_products()]
    log(f"products ready ({time.time() - t0:.1f}s)")
    report = {"products": [p.name for p in prods]}

    # sizer on exact counts, margin on analysis timings
    F, T = sizer_samples(prods)
    ps, loss = train_net(SELECTOR_SIZER, F, T, epochs, 4096, 3e-3, seed)
    log(f"sizer: {F.shape[0]} samples, loss {loss:.4f} ({time.time() - t0:.1f}s)")
    base = structure_ms(prods, SELECT_HEURISTIC)
    margins = {}
    for mg in (1.0, 1.25, 1.5, 2.0, 3.0):
        selector_set(SELECTOR_SIZER, ps, mg)
        margins[mg] = structure_ms(prods, SELECT_LEARNED)
    best_mg = min(margins, key=margins.get)
    ship = margins[best_mg] < base
    log(f"sizer margins (analysis+fill ms, heuristic {base:.3f}): {margins} -> {best_mg}"
        f"{'' if ship else ' (slower than the heuristic: not shipped)'}")
    path = os.path.join(out_dir, "sizer.txt")
    if ship:
        write_params(path, ps, best_mg)
    elif os.path.exists(path):
        os.remove(path)
    report["sizer"] = {"samples": int(F.shape[0]), "loss": loss, "heuristic_ms": base,
                       "margin_ms": {str(k): v for k, v in margins.items()}, "margin": best_mg,
                       "shipped": bool(ship)}
    selector_clear(SELECTOR_SIZER)

    # chooser on kernel timings
    F, T, gid, table = chooser_samples(prods, log=log)
    pc, loss = train_net(SELECTOR_CHOOSER, F, T, chooser_epochs, 4096, 3e-3, seed + 1)
    tot = group_totals(table, F, gid, pc)
    log(f"chooser: {F.shape[0]} samples, {len(table)} groups, loss {loss:.4f}, timed groups "
        f"(us) {tot} ({time.time() - t0:.1f}s)")
    # margin on whole numeric passes: the chooser overrides the heuristic's
    # bin only where it predicts a gain of more than the margin. Each
    # margin is scored by the learned/heuristic time ratio of every
    # training product's pass: the smallest geometric mean among margins
    # whose worst product is at most 2 % slower (else the margin whose
    # worst product loses least); shipped only if that mean is below 1.
    base = passes_ms(prods, SELECT_HEURISTIC)
    margins, scores = {}, {}
    for mg in (1.0, 1.1, 1.25, 1.5, 2.0, 3.0):
        selector_set(SELECTOR_CHOOSER, pc, mg)
        r = passes_ms(prods, SELECT_LEARNED) / np.maximum(base, 1e-6)
        margins[mg] = {"geomean": float(np.exp(np.mean(np.log(r)))), "worst": float(r.max()),
                       "worst_product": prods[int(np.argmax(r))].name}
        scores[mg] = (0, margins[mg]["geomean"]) if r.max() <= 1.02 else (1, r.max())
    selector_clear(SELECTOR_CHOOSER)
    best_mg = min(scores, key=scores.get)
    ship = margins[best_mg]["geomean"] < 1.0
    log(f"chooser margins (learned/heuristic numeric time over {len(prods)} products, "
        f"heuristic total {base.sum():.3f} ms): {margins} -> {best_mg}"
        f"{'' if ship else ' (not faster than the heuristic: not shipped)'}")
    path = os.path.join(out_dir, "chooser.txt")
    if ship:
        write_params(path, pc, best_mg)
    elif os.path.exists(path):
        os.remove(path)
    report["chooser"] = {"samples": int(F.shape[0]), "groups": len(table), "loss": loss,
                         "timed_groups_us": tot, "heuristic_ms": [float(v) for v in base],
                         "margin_ms": {str(k): v for k, v in margins.items()}, "margin": best_mg,
                         "shipped": bool(ship)}
    with open(os.path.join(out_dir, "train_report.json"), "w") as fh:
        json.dump({"report": report, "groups": table}, fh, indent=1)
    for p in prods:
        p.close()
    return report


# ---------------------------------------------------------------------------
# evaluation: heuristic vs learned vs exhaustive-best per row group
# ---------------------------------------------------------------------------
def numeric_ms(p: Product, reps=5):
    set_timing(True)
    best = None
    for _ in range(reps):
        reset_timing()
        p.plan.numeric()
        ms = get_timing()["numeric"][0]
        best = ms if best is None else min(best, ms)
    set_timing(False)
    return best


def evaluate_chooser(name, A, B, log=print):
    """Numeric-pass milliseconds of one product under the heuristic bins,
    the chooser's bins and the measured-best bin of every row group; the
    values of all three are compared bit for bit."""
    selector_clear(SELECTOR_SIZER)
    p = Product(name, A, B)
    heur_ms = numeric_ms(p)
    ref = p.plan.result()
    _, nb_heur = p.plan.bins()
    # exhaustive: every feasible accumulator timed on every row group
    bk, _ = buckets(p, min_rows=1)
    nb_best = nb_heur.copy()
    for key, rows in bk.items():
        cb = chooser_bins(p.nz[rows[:1]], p.span[rows[:1]])[0]
        cands = sorted({int(b) for b in cb if b >= 0})
        t = time_bucket(p, rows, cands)
        nb_best[rows] = min(t, key=t.get)
    assert p.plan.set_num_bins(nb_best) == 0
    best_ms = numeric_ms(p)
    best = p.plan.result()
    # learned: the analysis with the chooser loaded
    p.plan.analyse(SELECT_LEARNED)
    p.plan.fill()
    learned_ms = numeric_ms(p)
    lrn = p.plan.result()
    _, nb_lrn = p.plan.bins()
    same = all(np.array_equal(getattr(ref, f), getattr(x, f)) for x in (best, lrn)
               for f in ("row_ptr", "col_idx")) and \
        all(np.array_equal(ref.values.view(np.uint32), x.values.view(np.uint32)) for x in (best, lrn))
    agree = float((nb_lrn == nb_best).mean())
    p.close()
    out = {"product": name, "numeric_ms": {"heuristic": heur_ms, "learned": learned_ms,
                                           "exhaustive_best": best_ms},
           "bins_agree_with_best": agree, "bit_identical": bool(same)}
    log(json.dumps(out))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--chooser-epochs", type=int, default=400)
    t.add_argument("--seed", type=int, default=1)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        rep = train(args.out, epochs=args.epochs, chooser_epochs=args.chooser_epochs, seed=args.seed,
                    log=lambda s: print(s, flush=True))
        print(json.dumps(rep))
    return 0


if __name__ == "__main__":
    sys.exit(main())
