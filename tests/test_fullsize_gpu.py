"""GPU parity at the benchmark workloads' full sizes, through the C ABI.

  R-MAT 2^18 (C = A*A, 2 929 408 248 products, nnz(C) = 1 277 297 235):
      both ALGO_HIP (device one-shot) and ALGO_HIP_LEARNED (device-memory
      descriptor); row_ptr and the int64 op counts equal the oracle's; every
      row's columns in range and strictly ascending; 2048 seeded rows,
      including the 256 with the most products, bit-exact against the oracle.
  stencil 64^3: nnz(C) = 314^3 and products = 566^3 (closed forms); four
      numeric passes with redrawn values through a device-memory descriptor,
      per-call and hipGraph-replayed, every pass fully bit-exact.
  mixed 2^20 x 2^16 x 2^20: fully bit-exact, both algos.
  batched 512 uneven: all 512 C_i equal to separate device one-shots; 64
      problems (the 8 largest and 56 seeded others) bit-exact vs the oracle.

Every case prints its max, mean and 99.9th-percentile relative error
against the FP64 Gustavson oracle and its largest error relative to
sum |a_ik * b_kj| (`[fp64] ...` lines). The oracle runs row-parallel on the
host's cores (same results as its sequential passes).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import rvsp
from rvsp import workloads as W
from oracle import pyoracle
from helpers import assert_bitexact, rel_err_stats

pytestmark = pytest.mark.gpu

THREADS = pyoracle.threads_default()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def abs_sums(A, B, sym=None):
    """Per-entry sum of |a_ik * b_kj| (FP64 oracle on |A|, |B|): the scale
    an FP32 accumulation's error is bounded against."""
    Aa = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, np.abs(A.values))
    Ba = Aa if B is A else rvsp.Csr(B.rows, B.cols, B.row_ptr, B.col_idx, np.abs(B.values))
    return pyoracle.spgemm_par(Aa, Ba, threads=THREADS, sym=sym)[3]


def fp64_line(config, got_vals, vd, absd=None):
    """Prints max / mean / 99.9th-percentile relative error against the FP64
    oracle (entries whose FP64 value is 0 are left out) and, with absd, the
    largest error relative to sum |a*b|."""
    mx, mean = rel_err_stats(got_vals, vd)
    got = np.asarray(got_vals, np.float64)
    nz = np.abs(vd) > 0
    p999 = float(np.quantile(np.abs(got[nz] - vd[nz]) / np.abs(vd[nz]), 0.999)) if nz.any() else 0.0
    line = (f"[fp64] {config}: entries={len(vd)} max_rel_err={mx:.3e} mean_rel_err={mean:.3e} "
            f"p99.9_rel_err={p999:.3e}")
    if absd is not None:
        pos = absd > 0
        scaled = float((np.abs(got[pos] - vd[pos]) / absd[pos]).max()) if pos.any() else 0.0
        line += f" max_err_over_sum_abs={scaled:.3e}"
    say(line)
    return mx, mean


def d2h(ptr, dtype, count):
    out = np.empty(count, dtype)
    if count:
        rvsp.check("d2h", rvsp.lib().rvsp_hip_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
    return out


def check_columns_ascending(rp, ci_ptr, ncols, what, chunk_rows=1 << 14):
    """Every row's columns lie in [0, ncols) and strictly ascend; fetched
    from the device a block of rows at a time."""
    m = len(rp) - 1
    for r0 in range(0, m, chunk_rows):
        r1 = min(m, r0 + chunk_rows)
        s, e = int(rp[r0]), int(rp[r1])
        if e == s:
            continue
        ci = d2h(ci_ptr + 4 * s, np.int32, e - s)
        assert ci.min() >= 0 and ci.max() < ncols, f"{what}: column out of range in rows {r0}..{r1}"
        up = np.diff(ci) > 0
        # positions where a new row starts are exempt
        starts = rp[r0 + 1:r1].astype(np.int64) - s - 1
        starts = starts[(starts >= 0) & (starts < len(up))]
        up[starts] = True
        assert up.all(), f"{what}: columns not strictly ascending near entry {s + int(np.argmin(up))}"


def row_subset(A, rows):
    rows = np.asarray(rows, np.int64)
    lens = (A.row_ptr[rows + 1] - A.row_ptr[rows]).astype(np.int64)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = (np.concatenate([np.arange(A.row_ptr[r], A.row_ptr[r + 1]) for r in rows])
           if len(rows) else np.zeros(0, np.int64))
    return rvsp.Csr.make(len(rows), A.cols, rp, A.col_idx[idx], A.values[idx])


def check_sample_rows(A, B, rp, ci_ptr, va_ptr, ops, what, nrows=2048, heavy=256, seed=2024):
    """Bit-exact comparison of seeded rows (the `heavy` rows with the most
    products among them) with the oracle; returns the FP64 error stats."""
    order = np.argsort(ops, kind="stable")[::-1]
    top = order[:heavy]
    rng = np.random.default_rng(seed)
    nonempty = np.nonzero(np.diff(A.row_ptr))[0]
    rest = rng.choice(np.setdiff1d(nonempty, top), size=nrows - heavy, replace=False)
    rows = np.sort(np.concatenate([top, rest]))
    S = row_subset(A, rows)
    orp, oci, ova, ovd = pyoracle.spgemm_par(S, B, threads=THREADS)
    got_c, got_v = [], []
    for r in rows:
        s, e = int(rp[r]), int(rp[r + 1])
        got_c.append(d2h(ci_ptr + 4 * s, np.int32, e - s))
        got_v.append(d2h(va_ptr + 4 * s, np.float32, e - s))
    got = rvsp.Csr(len(rows), B.cols, orp.copy(), np.concatenate(got_c), np.concatenate(got_v))
    lens = np.array([len(c) for c in got_c], np.int64)
    got.row_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    assert_bitexact(got, orp, oci, ova, f"{what}: {len(rows)} sampled rows")
    say(f"{what}: {len(rows)} rows bit-exact ({int(ops[rows].sum())} products, "
        f"heaviest row {int(ops[top[0]])} products)")
    return fp64_line(what, got.values, ovd, abs_sums(S, B))


# ---------------------------------------------------------------------------
# R-MAT 2^18
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def rmat():
    A, _ = W.rmat()
    rp, ops, tot = pyoracle.symbolic(A, A, threads=THREADS)
    assert tot == 1_277_297_235 and int(ops.sum()) == 2_929_408_248
    return A, rp, ops


def test_rmat_full_size_plain_device_oneshot(rmat):
    A, orp, oops = rmat
    say("R-MAT 2^18, ALGO_HIP, device one-shot")
    dA = rvsp.DeviceCsr(A)
    L = rvsp.lib()
    # op counts from the descriptor's analysis of the same operands
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_algo(rvsp.ALGO_HIP) == 0
    st, _, nnz = d.estimate(dA.struct(), dA.struct())
    assert st == 0 and nnz == 1_277_297_235
    st, ops = d.op_counts(A.rows)
    assert st == 0
    np.testing.assert_array_equal(ops, oops)
    assert d.products() == 2_929_408_248
    d.close()
    a = dA.struct()
    c = rvsp.CsrStruct()
    rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(a), C.byref(c),
                                                           rvsp.ALGO_HIP))
    try:
        assert c.nnz == 1_277_297_235
        rp = d2h(c.row_ptr, np.int32, A.rows + 1)
        np.testing.assert_array_equal(rp, orp)
        check_columns_ascending(rp, c.col_idx, A.cols, "rmat plain")
        check_sample_rows(A, A, rp, c.col_idx, c.values, oops, "rmat-2^18 ALGO_HIP")
    finally:
        L.rvsp_csr_destroy(C.byref(c))


def test_rmat_full_size_learned_descriptor(rmat):
    A, orp, oops = rmat
    say("R-MAT 2^18, ALGO_HIP_LEARNED, device-memory descriptor")
    dA = rvsp.DeviceCsr(A)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_algo(rvsp.ALGO_HIP_LEARNED) == 0
    a = dA.struct()
    st, ws, nnz = d.estimate(a, a)
    assert st == 0 and nnz == 1_277_297_235
    st, ops = d.op_counts(A.rows)
    assert st == 0
    np.testing.assert_array_equal(ops, oops)
    assert d.products() == 2_929_408_248
    crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
    cci = rvsp.DeviceBuffer(4 * nnz)
    cva = rvsp.DeviceBuffer(4 * nnz)
    c = rvsp.CsrStruct(A.rows, A.cols, nnz, crp.ptr, cci.ptr, cva.ptr, rvsp.DTYPE_FP32, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    assert d.compute(a, a, c, wsb.ctypes.data) == 0
    assert d.wait() == 0
    rp = crp.to_array(np.int32, A.rows + 1)
    np.testing.assert_array_equal(rp, orp)
    check_columns_ascending(rp, cci.ptr, A.cols, "rmat learned")
    check_sample_rows(A, A, rp, cci.ptr, cva.ptr, oops, "rmat-2^18 ALGO_HIP_LEARNED")
    d.close()


# ---------------------------------------------------------------------------
# stencil 64^3: analysis once, numeric passes with redrawn values
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("graph", [False, True], ids=["per-call", "graph"])
def test_stencil_64_reuse_full_size(graph):
    g = 64
    A, _ = W.stencil(g=g)
    say(f"stencil 64^3 reuse, graph={graph}")
    dA = rvsp.DeviceCsr(A)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_graph(graph) == 0
    a = dA.struct()
    st, ws, nnz = d.estimate(a, a)
    assert st == 0
    assert nnz == (5 * g - 6) ** 3 == 30_959_144
    assert d.products() == (9 * g - 10) ** 3 == 181_321_496
    crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
    cci = rvsp.DeviceBuffer(4 * nnz)
    cva = rvsp.DeviceBuffer(4 * nnz)
    c = rvsp.CsrStruct(A.rows, A.cols, nnz, crp.ptr, cci.ptr, cva.ptr, rvsp.DTYPE_FP32, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    sym = pyoracle.symbolic(A, A, threads=THREADS)
    worst = (0.0, 0.0)
    for p in range(4):
        vals = rvsp.fill_values(A.nnz, 7000 + p, normal=True)
        dA.va.upload(vals)
        assert d.compute(a, a, c, wsb.ctypes.data) == 0
        assert d.wait() == 0
        Ah = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx, vals)
        orp, oci, ova, ovd = pyoracle.spgemm_par(Ah, Ah, threads=THREADS, sym=sym)
        got = rvsp.Csr(A.rows, A.cols, crp.to_array(np.int32, A.rows + 1),
                       cci.to_array(np.int32, nnz), cva.to_array(np.float32, nnz))
        assert_bitexact(got, orp, oci, ova, f"stencil pass {p}")
        mx, mean = fp64_line(f"stencil-64^3 pass {p} graph={graph}", got.values, ovd,
                             abs_sums(Ah, Ah, sym=sym))
        worst = (max(worst[0], mx), max(worst[1], mean))
    say(f"stencil 64^3 graph={graph}: 4 passes bit-exact, worst max/mean rel err {worst[0]:.3e} / "
        f"{worst[1]:.3e}")
    d.close()


# ---------------------------------------------------------------------------
# mixed 2^20 x 2^16 x 2^20
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def mixed():
    A, B = W.mixed()
    ref = pyoracle.spgemm_par(A, B, threads=THREADS)
    return A, B, ref


@pytest.mark.parametrize("algo", [rvsp.ALGO_HIP, rvsp.ALGO_HIP_LEARNED], ids=["plain", "learned"])
def test_mixed_full_size(mixed, algo):
    A, B, (orp, oci, ova, ovd) = mixed
    say(f"mixed 2^20 x 2^16 x 2^20, algo {algo}")
    got = rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B), algo=algo)
    assert_bitexact(got, orp, oci, ova, f"mixed algo {algo}")
    fp64_line(f"mixed-2^20 algo={algo}", got.values, ovd, abs_sums(A, B))


# ---------------------------------------------------------------------------
# batched 512 uneven problems
# ---------------------------------------------------------------------------
def test_batched_512_full_size():
    pairs = W.batched()
    n = len(pairs)
    assert n == 512
    say("batched 512 uneven, one stacked product per chunk vs separate calls")
    L = rvsp.lib()
    dA = [rvsp.DeviceCsr(A) for A, _ in pairs]
    dB = [rvsp.DeviceCsr(B) for _, B in pairs]
    As = (rvsp.CsrStruct * n)(*[d.struct() for d in dA])
    Bs = (rvsp.CsrStruct * n)(*[d.struct() for d in dB])
    Cs = (rvsp.CsrStruct * n)()
    h = C.c_void_p()
    rvsp.check("batched", L.rvsp_spgemm_batched_device(n, As, Bs, Cs, rvsp.ALGO_DEFAULT, C.byref(h)))
    try:
        say(f"batched: {L.rvsp_spgemm_batch_chunks(h)} chunk(s), total nnz(C) "
            f"{sum(Cs[i].nnz for i in range(n))}")
        prods = np.array([int(np.diff(B.row_ptr)[A.col_idx].sum()) for A, B in pairs], np.int64)
        largest = set(np.argsort(prods, kind="stable")[::-1][:8].tolist())
        rng = np.random.default_rng(512)
        others = rng.choice(sorted(set(range(n)) - largest), size=56, replace=False).tolist()
        to_oracle = largest | set(others)
        errs = []
        for i in range(n):
            Ci = rvsp.device_csr_to_host(Cs[i])
            single = rvsp.spgemm_device(dA[i], dB[i])
            np.testing.assert_array_equal(Ci.row_ptr, single.row_ptr, err_msg=f"problem {i}")
            np.testing.assert_array_equal(Ci.col_idx, single.col_idx, err_msg=f"problem {i}")
            np.testing.assert_array_equal(Ci.values.view(np.uint32), single.values.view(np.uint32),
                                          err_msg=f"problem {i}")
            if i in to_oracle:
                A, B = pairs[i]
                orp, oci, ova, ovd = pyoracle.spgemm_par(A, B, threads=THREADS)
                assert_bitexact(Ci, orp, oci, ova, f"batched[{i}] n={A.rows}")
                absd = abs_sums(A, B)
                got = Ci.values.astype(np.float64)
                pos = absd > 0
                scaled = float((np.abs(got[pos] - ovd[pos]) / absd[pos]).max()) if pos.any() else 0.0
                nzc = int((np.abs(ovd) > 0).sum())
                errs.append((rel_err_stats(Ci.values, ovd), nzc, scaled))
        mx = max(e[0][0] for e in errs)
        mean = sum(e[0][1] * e[1] for e in errs) / max(1, sum(e[1] for e in errs))
        scaled = max(e[2] for e in errs)
        say(f"batched: 512 problems equal to separate calls; {len(to_oracle)} bit-exact vs oracle "
            f"({int(prods[sorted(to_oracle)].sum())} products)")
        say(f"[fp64] batched-512 (64 problems): max_rel_err={mx:.3e} mean_rel_err={mean:.3e} "
            f"max_err_over_sum_abs={scaled:.3e}")
    finally:
        L.rvsp_spgemm_batch_destroy(h)
