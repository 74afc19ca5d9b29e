#!/usr/bin/env python3
"""SpGEMM benchmark on one MI355X: GFLOP/s (2 x intermediate products / s)
and fraction of the HBM roofline, with the rv-sparse scalar path timed on the
host's cores beside it.

A step is one full C = A*B on device-resident inputs (symbolic analysis,
C allocation, structure fill, numeric pass), as rv-sparse's bench times its
one-shot kernel. Default workload: R-MAT 2^18 x 2^18 (a,b,c,d = .57,.19,.19,
.05), 16 x 2^18 edge draws, C = A*A.

Prints one JSON line. Options: --steps K --warmup W --gpus N (1 only)
--workload {rmat,uniform,mixed,stencil,batched,mtx:PATH,gen:R:C:D:SEED}
--algo {hip,learned} --no-cpu.
  rmat, uniform, mixed  step = one device one-shot product
  mtx:PATH, gen:...     same, on a local Matrix Market file (A*A) or on
                        rv-sparse's generated pair (A*B)
  stencil               analysis once; step = one numeric pass with new
                        values (hipGraph replay; per-call launches reported)
  batched               512 uneven products; step = the whole batch in one
                        stacked launch set (512 separate calls reported)
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "rv-sparse_amd")]

import rvsp  # noqa: E402
from rvsp import workloads as W  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "SpGEMM GFLOP/s (2×products/s) + % HBM roofline, learned/batched vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_workload(name):
    if name == "rmat":
        A, B = W.rmat()
        return A, B, True, "R-MAT 2^18 x 2^18 (.57,.19,.19,.05), 16*2^18 edge draws, C=A*A, U[-1,1) FP32"
    if name == "uniform":
        A, B = W.uniform()
        return A, B, True, "uniform 4096 x 4096, 16 nnz/row, C=A*A, U[-1,1) FP32"
    if name == "mixed":
        A, B = W.mixed()
        return A, B, False, ("A 2^20 x 2^16 power-law rows (s=1.2, mean 8, max 4096) x B 2^16 x 2^20 "
                             "banded (12 of 32), N(0,1) FP32")
    if name == "stencil":
        A, B = W.stencil()
        return A, B, True, "27-point stencil 64^3, C=A*A, N(0,1) FP32"
    if name.startswith("mtx:"):
        # a local Matrix Market file, C = A*A as rv-sparse's bench runs its
        # real matrices (read by the library's rvsp_mtx_load)
        path = name[4:]
        A = rvsp.mtx_load(path)
        if A.rows != A.cols:
            raise SystemExit(f"{path}: {A.rows} x {A.cols} (inferred) is not square, no A*A")
        return A, A, True, f"{os.path.basename(path)} ({A.rows} x {A.cols}, nnz {A.nnz}), C=A*A, file values FP32"
    if name.startswith("gen:"):
        # rv-sparse's bench --gen R C DENSITY SEED (bench.c:677-707): A = R x C,
        # B = C x R (seed + 1), from the library's genmat generator
        R, Cc, D, S = gen_args(name)
        A, B = rvsp.gen_bench_pair(R, Cc, D, S)
        return A, B, False, (f"genmat {R} x {Cc} * {Cc} x {R}, density {D}, seed {S} "
                             f"(nnz {A.nnz} + {B.nnz}), C=A*B, values 0.1..0.9 FP32")
    raise SystemExit(f"unknown workload {name}")


def gen_args(name):
    try:
        R, Cc, D, S = name[4:].split(":")
        return int(R), int(Cc), float(D), int(S)
    except ValueError:
        raise SystemExit(f"{name}: expected gen:ROWS:COLS:DENSITY:SEED")


def default_csv_label(workload):
    """The label rv-sparse's run_bench.sh gives a case (run_bench.sh:1432
    for generated inputs: gen_<R>x<C>_d<DENSITY>; a matrix file's stem)."""
    if workload.startswith("gen:"):
        R, Cc, _, _ = gen_args(workload)
        return f"gen_{R}x{Cc}_d{workload.split(':')[3]}"
    if workload.startswith("mtx:"):
        return os.path.splitext(os.path.basename(workload[4:]))[0]
    return workload


def device_rows(c: rvsp.CsrStruct, rows):
    """Fetches selected rows of a device-resident C to the host."""
    L = rvsp.lib()
    rp = np.empty(c.rows + 1, np.int32)
    rvsp.check("d2h", L.rvsp_hip_memcpy_d2h(rp.ctypes.data, c.row_ptr, rp.nbytes))
    out = {}
    for r in rows:
        s, e = int(rp[r]), int(rp[r + 1])
        ci = np.empty(e - s, np.int32)
        va = np.empty(e - s, np.float32)
        if e > s:
            rvsp.check("d2h", L.rvsp_hip_memcpy_d2h(ci.ctypes.data, c.col_idx + 4 * s, ci.nbytes))
            rvsp.check("d2h", L.rvsp_hip_memcpy_d2h(va.ctypes.data, c.values + 4 * s, va.nbytes))
        out[r] = (ci, va)
    return rp, out


def row_subset(A: rvsp.Csr, rows) -> rvsp.Csr:
    rows = np.asarray(rows, np.int64)
    lens = A.row_ptr[rows + 1] - A.row_ptr[rows]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(A.row_ptr[r], A.row_ptr[r + 1]) for r in rows]) if len(rows) else \
        np.zeros(0, np.int64)
    return rvsp.Csr.make(len(rows), A.cols, rp, A.col_idx[idx], A.values[idx])


def spot_check(A, B, c, seed=7, nrows=64):
    """Bit-exact comparison of a seeded sample of C's rows with the oracle;
    returns the sample's parity record, with the max and mean relative
    error of those rows against the FP64 Gustavson oracle (entries whose
    FP64 value is 0 left out)."""
    from oracle import pyoracle

    rng = np.random.default_rng(seed)
    nonempty = np.nonzero(np.diff(A.row_ptr))[0]
    rows = np.sort(rng.choice(nonempty, size=min(nrows, len(nonempty)), replace=False))
    _, got = device_rows(c, rows)
    S = row_subset(A, rows)
    rp, ci, va, vd = pyoracle.spgemm(S, B, f64=True)
    for j, r in enumerate(rows):
        gci, gva = got[r]
        if not (np.array_equal(gci, ci[rp[j]:rp[j + 1]]) and
                np.array_equal(gva.view(np.uint32), va[rp[j]:rp[j + 1]].view(np.uint32))):
            raise SystemExit(f"spot check failed at row {r}")
    return parity_record(len(rows), va, vd, "seeded rows of C, bit-exact vs the oracle")


def parity_record(rows, va, vd, what):
    va64 = np.asarray(va, np.float64)
    nz = np.abs(vd) > 0
    rel = np.abs(va64[nz] - vd[nz]) / np.abs(vd[nz])
    return {"checked": what, "rows": int(rows), "entries": int(len(vd)), "bit_exact": True,
            "max_rel_err_f64": float(rel.max()) if rel.size else 0.0,
            "mean_rel_err_f64": float(rel.mean()) if rel.size else 0.0}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_cpu(run, warmups=3, runs=10):
    """BASELINE.md's CPU protocol: pinned to one host core, `warmups` untimed
    runs, then the median of `runs` timed runs. Returns (median s, core)."""
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        for _ in range(warmups):
            run()
        times = [run() for _ in range(runs)]
    finally:
        os.sched_setaffinity(0, old)
    return float(np.median(times)), core


CPU_BUILD = ("rv-sparse's sources, gcc -std=c11 -O3 -flto -march=x86-64-v3 (rv-sparse's Makefile "
             "uses -march=native; the library is built in another container than the host it runs on)")


def cpu_baseline(A, B, target_s=1.0, seed=11):
    """rv-sparse's scalar one-shot (compiled from its sources, or this
    project's restatement if that build is absent) on a seeded row sample of
    the same workload sized to about target_s per run: 1 thread pinned to a
    core, 3 warm-ups, median of 10 runs."""
    from oracle import pyoracle

    kind = "reference" if pyoracle.ref_available() else "port"

    def run(S):
        if kind == "reference":
            return pyoracle.ref_spgemm(S, B, timed=True)[-1]
        return pyoracle.time_oneshot(S, B)

    rng = np.random.default_rng(seed)
    perm = rng.permutation(A.rows)
    # calibrate on a small sample, then size the timed sample to ~target_s
    frac = 0.002
    S = row_subset(A, np.sort(perm[: max(1, int(A.rows * frac))]))
    t = run(S)
    want = min(1.0, frac * (target_s / max(t, 1e-6)))
    S = row_subset(A, np.sort(perm[: max(1, int(A.rows * want))]))
    prods = int(pyoracle.symbolic(S, B)[1].sum())
    t, core = timed_cpu(lambda: run(S))
    return {
        "value": round(2 * prods / t / 1e9, 4),
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": kind,
        "cpu_model": cpu_model(),
        "build": CPU_BUILD if kind == "reference" else "oracle/oracle.c, gcc -O3 -march=x86-64-v3",
        "protocol": f"pinned to core {core}; 3 warm-ups, median of 10 runs",
        "sample": (f"{S.rows} of {A.rows} rows (seeded random), {prods} products, full one-shot "
                   f"(count, C alloc, fill+sort, numeric): median {t:.3f} s"),
    }


def cpu_baseline_batch(pairs, target_s=1.0, seed=11):
    """rv-sparse's scalar one-shot over whole problems of the batch, a seeded
    subset of about target_s of CPU time per run: 1 thread pinned to a core,
    3 warm-ups, median of 10 runs of the subset."""
    from oracle import pyoracle

    kind = "reference" if pyoracle.ref_available() else "port"

    def one(A, B):
        return pyoracle.ref_spgemm(A, B, timed=True)[-1] if kind == "reference" else pyoracle.time_oneshot(A, B)

    subset, t, prods = [], 0.0, 0
    for i in np.random.default_rng(seed).permutation(len(pairs)):
        A, B = pairs[i]
        t += one(A, B)
        prods += int(np.diff(B.row_ptr)[A.col_idx].sum())
        subset.append(i)
        if t >= target_s:
            break
    med, core = timed_cpu(lambda: sum(one(*pairs[i]) for i in subset))
    return {
        "value": round(2 * prods / med / 1e9, 4),
        "unit": "GFLOP/s",
        "cores": 1,
        "kind": kind,
        "cpu_model": cpu_model(),
        "build": CPU_BUILD if kind == "reference" else "oracle/oracle.c, gcc -O3 -march=x86-64-v3",
        "protocol": f"pinned to core {core}; 3 warm-ups, median of 10 runs",
        "sample": (f"{len(subset)} of {len(pairs)} problems (seeded order), {prods} products, full "
                   f"one-shot each (count, C alloc, fill+sort, numeric): median {med:.3f} s per run"),
    }


ALGOS = {"hip": rvsp.ALGO_HIP, "learned": rvsp.ALGO_HIP_LEARNED}


TRAFFIC_KIND = ("L2-miss bytes of one numeric pass (rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                "scaled by a device copy of known size; MALL hits included), read from the committed "
                "profiles/traffic_<workload>.md")


def measured_traffic(workload):
    """HBM bytes of one numeric pass, from the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary of this workload (the line `numeric_hbm_bytes_per_pass: N`
    of profiles/traffic_<workload>.md), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("numeric_hbm_bytes_per_pass:"):
                return int(line.split(":")[1].strip())
    return None


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB: a dispatch of known traffic
    (mib MiB read, mib MiB written) that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters in a profiled run."""
    n = int(mib) << 20
    src, dst = rvsp.DeviceBuffer(n), rvsp.DeviceBuffer(n)
    rvsp.check("memset", rvsp.lib().rvsp_hip_memset(src.ptr, 1, n))
    rvsp.check("sync", rvsp.lib().rvsp_hip_synchronize())
    rvsp.check("copy", rvsp.lib().rvsp_hip_memcpy_d2d(dst.ptr, src.ptr, n))
    rvsp.check("sync", rvsp.lib().rvsp_hip_synchronize())
    src.free()
    dst.free()


# rv-sparse's per-run CSV schema (bench/bench.c:531-534) and field formats
# (bench.c:879-913), so GPU rows land in one table with the reference's rows
CSV_HEADER = ("label,kernel,arm,build,march,cflags,cc_version,dtype,"
              "rows,cols,nnz_a,nnz_b,nnz_c,flops,op_mean,op_max,op_var,"
              "run,time_s,gops,correct,cycles,instructions")


def csv_row(meta, run, dt, nnz_c, correct):
    gops = meta["flops"] / dt / 1e9 if dt > 0 else 0.0
    return (f"{meta['label']},{meta['kernel']},{meta['arm']},{meta['build']},{meta['march']},"
            f"{meta['cflags']},{meta['cc_version']},f32,"
            f"{meta['rows']},{meta['cols']},{meta['nnz_a']},{meta['nnz_b']},{nnz_c},"
            f"{meta['flops']:.0f},{meta['op_mean']:.4f},{meta['op_max']:.0f},{meta['op_var']:.4f},"
            f"{run},{dt:.9f},{gops:.4f},{correct},,")  # no perf counters: empty, as bench.c


def bench_csv(args):
    """Rows in rv-sparse's bench CSV schema, one per timed run, for the GPU
    raw one-shot on host arrays (hip_f32, PCIe included: what rv-sparse's
    bench times once the kernel is bound, bench.c:853-866), the device
    one-shot with inputs resident in HBM (hip_f32_dev), and rv-sparse's own
    scalar kernel on the host cores (scalar_f32, the CPU leg). `correct` is
    1 when C matches the scalar kernel's result (bench.c:316-352)."""
    import subprocess
    from oracle import pyoracle

    A, B, same, desc = load_workload(args.workload)
    L = rvsp.lib()
    d = rvsp.Descr()
    st, _, _ = d.estimate(A.struct(), B.struct())
    rvsp.check("estimate", st)
    st, ops = d.op_counts(A.rows)
    rvsp.check("op_counts", st)
    d.close()
    ops = ops.astype(np.float64)
    mean = float(ops.mean())
    label = args.csv_label if args.csv_label is not None else default_csv_label(args.workload)
    meta = {"label": label, "rows": A.rows, "cols": A.cols, "nnz_a": A.nnz, "nnz_b": B.nnz,
            "flops": 2.0 * float(ops.sum()), "op_mean": mean, "op_max": float(ops.max()),
            "op_var": float(((ops - mean) ** 2).mean())}
    try:
        ccv = subprocess.run(["hipcc", "--version"], capture_output=True, text=True,
                             timeout=60).stdout.split("\n")[0].replace(",", ";").strip() or "-"
    except Exception:
        ccv = "-"
    kind = "reference" if pyoracle.ref_available() else "port"
    ref = (pyoracle.ref_spgemm(A, B) if kind == "reference" else pyoracle.spgemm(A, B))[:3]

    def same_as_ref(rp, ci, va):
        return int(np.array_equal(rp, ref[0]) and np.array_equal(ci, ref[1]) and
                   np.array_equal(va.view(np.uint32), np.asarray(ref[2], np.float32).view(np.uint32)))

    out = [CSV_HEADER]
    times = {}
    gpu_meta = dict(meta, arm="hip", build="gfx950", march="gfx950",
                    cflags="-O3 -ffp-contract=off", cc_version=ccv)
    # GPU, host arrays in and out
    for _ in range(args.warmup):
        rvsp.spgemm_raw(A, B)
    for r in range(args.steps):
        Cg, dt = rvsp.spgemm_raw(A, B, timed=True)  # the library call alone, as bench.c times
        out.append(csv_row(dict(gpu_meta, kernel="hip_f32"), r, dt, Cg.nnz,
                           same_as_ref(Cg.row_ptr, Cg.col_idx, Cg.values)))
        times.setdefault("hip_f32", []).append(dt)
    # GPU, device-resident operands
    dA = rvsp.DeviceCsr(A)
    dB = dA if same else rvsp.DeviceCsr(B)
    a, b = dA.struct(), dB.struct()
    for r in range(-args.warmup, args.steps):
        c = rvsp.CsrStruct()
        rvsp.check("sync", L.rvsp_hip_synchronize())
        t0 = time.perf_counter()
        rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c),
                                                               rvsp.ALGO_HIP_LEARNED))
        rvsp.check("sync", L.rvsp_hip_synchronize())
        dt = time.perf_counter() - t0
        if r >= 0:
            h = rvsp.device_csr_to_host(c)
            out.append(csv_row(dict(gpu_meta, kernel="hip_f32_dev"), r, dt, c.nnz,
                               same_as_ref(h.row_ptr, h.col_idx, h.values)))
            times.setdefault("hip_f32_dev", []).append(dt)
        L.rvsp_csr_destroy(C.byref(c))
    # CPU: rv-sparse's scalar kernel (or the oracle's restatement), 1 thread
    cpu_meta = dict(meta, kernel="scalar_f32", arm="baseline", build="x86-64-v3",
                    march="x86-64-v3", cflags="-std=c11 -O3", cc_version="gcc")
    for r in range(args.csv_cpu_runs):
        if kind == "reference":
            rp, ci, va, dt = pyoracle.ref_spgemm(A, B, timed=True)
            out.append(csv_row(cpu_meta, r, dt, int(rp[-1]), same_as_ref(rp, ci, va)))
        else:
            dt = pyoracle.time_oneshot(A, B)
            out.append(csv_row(dict(cpu_meta, kernel="orc_scalar_f32"), r, dt, int(ref[0][-1]), 1))
        times.setdefault("scalar_f32", []).append(dt)
    with open(args.csv, "w") as f:
        f.write("\n".join(out) + "\n")
    # rv-sparse's analyze.py statistics over the rows just written: speedup
    # over the scalar kernel (arm "baseline") with its bootstrap interval
    from rvsp import report
    summary = report.summarize(report.load_rows(args.csv))
    report.attach_speedups(summary)
    log("\n".join(report.table(summary)))

    def num(x):
        return None if x is None or x != x else round(float(x), 4)

    print(json.dumps({"csv": args.csv, "workload": desc, "label": label, "rows": len(out) - 1,
                      "median_s": {k: round(float(np.median(v)), 6) for k, v in times.items()},
                      "median_gflops": {k: round(meta["flops"] / float(np.median(v)) / 1e9, 3)
                                        for k, v in times.items()},
                      "speedup_vs_scalar": {key[3]: {"x": num(e["speedup"]),
                                                     "ci95": [num(e["speedup_lo"]), num(e["speedup_hi"])],
                                                     "status": e["status"]}
                                            for key, e in summary.items()}}), flush=True)


def measure_oneshot(args, workload):
    """The device one-shot of one workload: the bench record (value, phases,
    roofline, parity of seeded rows, CPU baseline unless args.no_cpu, and the
    plain heuristic beside the learned path when args.algo is learned)."""
    t0 = time.time()
    A, B, same, desc = load_workload(workload)
    log(f"workload {workload}: {desc}; nnzA={A.nnz} nnzB={B.nnz} ({time.time() - t0:.1f} s)")
    dA = rvsp.DeviceCsr(A)
    dB = dA if same else rvsp.DeviceCsr(B)
    algo = ALGOS[args.algo]
    L = rvsp.lib()
    a, b = dA.struct(), dB.struct()

    def step(al=algo):
        c = rvsp.CsrStruct()
        rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c), al))
        return c

    c = step()
    nnz_c = c.nnz
    parity = None
    if not args.no_check:
        parity = spot_check(A, B, c)
        log(f"spot check: {parity['rows']} rows bit-exact vs oracle")
    L.rvsp_csr_destroy(C.byref(c))
    d = rvsp.Descr()
    d.set_memory(rvsp.MEMORY_DEVICE)
    st, _, _ = d.estimate(a, b)
    rvsp.check("estimate", st)
    products = d.products()
    d.close()

    def timed(al, tag):
        """Warm-up, then exactly args.steps one-shot products between
        device synchronisations: (ms per step, phase timings)."""
        for _ in range(args.warmup):
            c = step(al)
            L.rvsp_csr_destroy(C.byref(c))
        rvsp.check("sync", L.rvsp_hip_synchronize())
        rvsp.reset_timing()
        rvsp.set_timing(True)
        t0 = time.perf_counter()
        for i in range(args.steps):
            ts = time.perf_counter()
            c = step(al)
            L.rvsp_csr_destroy(C.byref(c))
            log(f"{tag} step {i + 1}/{args.steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms")
        rvsp.check("sync", L.rvsp_hip_synchronize())
        t1 = time.perf_counter()
        rvsp.set_timing(False)
        return (t1 - t0) * 1e3 / args.steps, rvsp.get_timing()

    other = None
    if args.algo == "learned":
        # the plain heuristic-binned path, timed the same way, for comparison
        ms_plain, ph_plain = timed(ALGOS["hip"], "plain")
        other = {"algo": "hip", "value": round(2.0 * products / (ms_plain * 1e-3) / 1e9, 3),
                 "ms_per_step": round(ms_plain, 4),
                 "phases_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in ph_plain.items()}}
    ms_step, phases = timed(algo, args.algo)
    gflops = 2.0 * products / (ms_step * 1e-3) / 1e9

    num_ms, num_calls = phases["numeric"]
    num_avg_ms = num_ms / max(num_calls, 1)
    # compulsory bytes of the numeric pass: A (+B) structure and values read,
    # C structure read, C values written (same total as the one-shot's)
    alg_bytes = W.products_bytes(A, B, nnz_c, same)
    achieved = alg_bytes / (num_avg_ms * 1e-3) / 1e9
    step_gbps = alg_bytes / (ms_step * 1e-3) / 1e9

    traffic = measured_traffic(workload)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, B)

    out = {
        "metric": METRIC,
        "value": round(gflops, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded project generator)",
        "config": {
            "workload": desc,
            "m": A.rows, "k": A.cols, "n": B.cols, "nnz_a": A.nnz, "nnz_b": B.nnz,
            "nnz_c": nnz_c, "products": products, "algo": args.algo,
            "step": "device one-shot: analysis + C alloc + fill + numeric",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "numeric pass (row-binned LDS accumulator kernels; HIP events on the caller's stream, which joins the bins' side streams)",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_kind": TRAFFIC_KIND,
            "algorithmic_bytes": alg_bytes,
            "numeric_ms": round(num_avg_ms, 4),
        },
        "parity": parity,
        "beside_roofline": beside_roofline(products, traffic, num_avg_ms),
        "phases_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in phases.items()},
        "step_roofline_frac": round(step_gbps / HBM_PEAK_GBPS, 4),
        "cpu_baseline": cpu,
    }
    if other is not None:
        out["plain"] = other
        out["learned_vs_plain"] = round(other["ms_per_step"] / ms_step, 4)
    if args.selector_eval:
        from rvsp import selector as SEL
        log("selector evaluation (heuristic / learned / exhaustive-best numeric bins) ...")
        out["selector_eval"] = SEL.evaluate_chooser(workload, A, B, log=log)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat",
                    help="rmat, mixed, uniform, stencil, batched, mtx:PATH (a local .mtx file, C=A*A), or "
                         "gen:ROWS:COLS:DENSITY:SEED (rv-sparse's generated pair, C=A*B)")
    ap.add_argument("--algo", default="learned", choices=["hip", "learned"],
                    help="learned: MLP row selectors (the plain heuristic is timed beside it)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run one device copy of this many MiB (a known-size dispatch "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--selector-eval", action="store_true",
                    help="also time heuristic vs learned vs exhaustive-best numeric bins")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="rmat: skip the stencil-reuse and batched comparisons")
    ap.add_argument("--csv", default=None,
                    help="instead of the JSON bench line: write rv-sparse bench CSV rows (one per "
                         "timed run) for hip_f32, hip_f32_dev and scalar_f32 to this file")
    ap.add_argument("--csv-label", default=None,
                    help="the CSV label column (default: rv-sparse's label for the workload)")
    ap.add_argument("--csv-cpu-runs", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU path is a design note, not built.", file=sys.stderr)
        sys.exit(2)
    if not rvsp.gpu_available():
        raise SystemExit("no HIP device visible")
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    if args.csv:
        return bench_csv(args)
    if args.workload == "stencil":
        return bench_reuse(args)
    if args.workload == "batched":
        return bench_batched(args)
    out = measure_oneshot(args, args.workload)
    if args.workload == "rmat" and not args.no_extra:
        # the other north-star comparisons, measured in the same run:
        # hipGraph replay vs per-call passes, one batch vs separate calls
        sub = argparse.Namespace(**vars(args))
        sub.no_cpu, sub.steps, sub.warmup = True, 20, 3
        log("other configs: stencil 64^3 reuse ...")
        r = measure_reuse(sub)
        sub.steps, sub.warmup = 5, 1
        log("other configs: batched 512 ...")
        b = measure_batched(sub)
        out["other_configs"] = {
            "stencil_reuse": {k: r["config"][k] for k in
                              ("graph_ms_per_pass", "per_call_ms_per_pass", "graph_speedup")}
            | {"gflops_graph": r["value"], "numeric_roofline_frac": r["roofline"]["frac"],
               "parity": r.get("parity")},
            "batched_512": {k: b["config"][k] for k in
                            ("batched_ms", "separate_calls_ms", "batched_speedup", "chunks")}
            | {"gflops_batched": b["value"], "numeric_roofline_frac": b["roofline"]["frac"],
               "parity": b.get("parity")},
        }
        # the mixed-structure product (learned vs plain, and the numeric pass
        # under heuristic, learned and exhaustive-best bins) and the uniform
        # 4096 product against rv-sparse's scalar kernel on one host core
        sub.steps, sub.warmup, sub.algo = 10, 2, "learned"
        sub.no_cpu, sub.selector_eval = True, True
        log("other configs: mixed 2^20 x 2^16 ...")
        mx = measure_oneshot(sub, "mixed")
        sub.no_cpu, sub.selector_eval = args.no_cpu, False
        log("other configs: uniform 4096 ...")
        un = measure_oneshot(sub, "uniform")

        def brief(r):
            return {"gflops": r["value"], "ms_per_step": r["ms_per_step"],
                    "phases_ms": r["phases_ms"], "plain_gflops": r["plain"]["value"],
                    "learned_vs_plain": r["learned_vs_plain"],
                    "numeric_roofline_frac": r["roofline"]["frac"], "parity": r["parity"]}

        out["other_configs"]["mixed"] = brief(mx) | {"selector_eval": mx.get("selector_eval")}
        cpu = un.get("cpu_baseline")
        out["other_configs"]["uniform_4096"] = brief(un) | {
            "cpu_baseline": cpu,
            "gpu_over_cpu": round(un["value"] / cpu["value"], 1) if cpu else None}
    print(json.dumps(out), flush=True)


def base_record(args, value, ms_step, config):
    return {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded project generator)",
        "config": config,
    }


def beside_roofline(products, traffic, numeric_ms):
    """The two figures reported next to the roofline, not as it: the
    per-product gather bound (8 bytes of B per product, which L2 / MALL
    reuse can push past the peak) and the measured HBM traffic's rate."""
    t = numeric_ms * 1e-3
    return {"gather_bytes": 8 * int(products),
            "gather_GBps": round(8.0 * products / t / 1e9, 2) if t > 0 else None,
            "measured_GBps": round(traffic / t / 1e9, 2) if traffic and t > 0 else None}


def bench_reuse(args):
    print(json.dumps(measure_reuse(args)), flush=True)


def measure_reuse(args, n_sets=16):
    """27-point stencil 64^3: analysis once, then numeric passes, each after
    rewriting A's values in place with a freshly drawn set (sets drawn up
    front with the project generator and cycled)."""
    L = rvsp.lib()
    A, _ = W.stencil()
    desc = "27-point stencil 64^3, C=A*A, N(0,1); analysis once, numeric passes with new values"
    dA = rvsp.DeviceCsr(A)
    sets = [rvsp.DeviceBuffer.from_array(rvsp.fill_values(A.nnz, 1000 + i, normal=True))
            for i in range(n_sets)]
    out = {}
    out_parity = None
    for graph in (False, True):
        d = rvsp.Descr()
        d.set_memory(rvsp.MEMORY_DEVICE)
        d.set_graph(graph)
        a = dA.struct()
        st, ws, nnz = d.estimate(a, a)
        rvsp.check("estimate", st)
        products = d.products()
        stream = d.stream()
        crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
        cci = rvsp.DeviceBuffer(4 * nnz)
        cva = rvsp.DeviceBuffer(4 * nnz)
        c = rvsp.CsrStruct(A.rows, A.cols, nnz, crp.ptr, cci.ptr, cva.ptr, 3, 0, 0)
        wsb = np.zeros(max(ws, 1), np.uint8)

        def one(p):
            rvsp.check("copy", L.rvsp_hip_copy_async(dA.va.ptr, sets[p % n_sets].ptr, 4 * A.nnz, stream))
            rvsp.check("compute", d.compute(a, a, c, wsb.ctypes.data))

        one(0)  # fills C's structure
        for p in range(args.warmup + 1):
            one(p)
        rvsp.check("wait", d.wait())
        rvsp.reset_timing()
        rvsp.set_timing(True)
        t0 = time.perf_counter()
        for p in range(args.steps):
            one(p)
        rvsp.check("wait", d.wait())
        t1 = time.perf_counter()
        rvsp.set_timing(False)
        ph = rvsp.get_timing()
        out[graph] = ((t1 - t0) * 1e3 / args.steps, ph["numeric"][0] / max(ph["numeric"][1], 1))
        if not args.no_check:
            # last pass used set (steps-1) % n_sets: check a sample of rows
            Ah = rvsp.Csr(A.rows, A.cols, A.row_ptr, A.col_idx,
                          rvsp.fill_values(A.nnz, 1000 + (args.steps - 1) % n_sets, normal=True))
            par = spot_check(Ah, Ah, c)
            log(f"graph={graph}: spot check {par['rows']} rows bit-exact")
            out_parity = par
        d.close()
    ms_graph, num_graph = out[True]
    ms_call, num_call = out[False]
    value = 2.0 * products / (ms_graph * 1e-3) / 1e9
    alg_bytes = W.products_bytes(A, A, nnz, True)
    achieved = alg_bytes / (num_graph * 1e-3) / 1e9
    rec = base_record(args, value, ms_graph, {
        "workload": desc, "m": A.rows, "nnz_a": A.nnz, "nnz_c": nnz, "products": products,
        "step": "one numeric pass (values rewritten in place on the device, then compute)",
        "graph_ms_per_pass": round(ms_graph, 4), "per_call_ms_per_pass": round(ms_call, 4),
        "graph_speedup": round(ms_call / ms_graph, 3)})
    rec["roofline"] = {"bound": "hbm", "kernel": "numeric pass (hipGraph replay)",
                       "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBPS, 4),
                       "traffic": measured_traffic(args.workload), "traffic_kind": TRAFFIC_KIND,
                       "algorithmic_bytes": alg_bytes, "numeric_ms": round(num_graph, 4)}
    rec["parity"] = out_parity
    rec["beside_roofline"] = beside_roofline(products, rec["roofline"]["traffic"], num_graph)
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline(A, A)
    return rec


def bench_batched(args):
    print(json.dumps(measure_batched(args)), flush=True)


def measure_batched(args):
    """512 independent uneven products: one stacked launch set versus 512
    separate device one-shots."""
    L = rvsp.lib()
    t0 = time.time()
    pairs = W.batched()
    log(f"batched: generated {len(pairs)} problems ({time.time() - t0:.1f} s)")
    products = sum(int(np.diff(B.row_ptr)[A.col_idx].sum()) for A, B in pairs)
    dA = [rvsp.DeviceCsr(A) for A, _ in pairs]
    dB = [rvsp.DeviceCsr(B) for _, B in pairs]
    algo = ALGOS[args.algo]

    def separate():
        for i in range(len(pairs)):
            a, b = dA[i].struct(), dB[i].struct()
            c = rvsp.CsrStruct()
            rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c), algo))
            L.rvsp_csr_destroy(C.byref(c))

    def batched():
        _, ch, nnzs = rvsp.spgemm_batched(dA, dB, algo=algo, fetch=False, with_nnz=True)
        return ch, nnzs

    parity = None
    if not args.no_check:
        from oracle import pyoracle
        got, chunks = rvsp.spgemm_batched(dA[:8], dB[:8], algo=algo)[:2]
        va_all, vd_all, rows = [], [], 0
        for i in range(8):
            one = rvsp.spgemm_device(dA[i], dB[i], algo=algo)
            assert np.array_equal(got[i].col_idx, one.col_idx)
            assert np.array_equal(got[i].values.view(np.uint32), one.values.view(np.uint32))
            rp, ci, va, vd = pyoracle.spgemm(pairs[i][0], pairs[i][1], f64=True)
            if not (np.array_equal(got[i].col_idx, ci) and
                    np.array_equal(got[i].values.view(np.uint32), va.view(np.uint32))):
                raise SystemExit(f"batched problem {i} differs from the oracle")
            va_all.append(va)
            vd_all.append(vd)
            rows += pairs[i][0].rows
        parity = parity_record(rows, np.concatenate(va_all), np.concatenate(vd_all),
                               "first 8 problems: batched = separate calls = oracle, bit-exact")
        log("batched vs separate vs oracle: first 8 problems bit-identical")
    for _ in range(max(args.warmup, 1)):
        _, nnzs = batched()
        separate()
    rvsp.check("sync", L.rvsp_hip_synchronize())
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 5)):
        separate()
    rvsp.check("sync", L.rvsp_hip_synchronize())
    ms_sep = (time.perf_counter() - t0) * 1e3 / max(1, args.steps // 5)
    rvsp.reset_timing()
    rvsp.set_timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        chunks, _ = batched()
        log(f"step {i + 1}/{args.steps}: {(time.perf_counter() - ts) * 1e3:.1f} ms")
    rvsp.check("sync", L.rvsp_hip_synchronize())
    ms_bat = (time.perf_counter() - t0) * 1e3 / args.steps
    rvsp.set_timing(False)
    ph = rvsp.get_timing()
    num_ms = ph["numeric"][0] / max(args.steps, 1)  # all chunks of one batch
    alg_bytes = sum(W.products_bytes(A, B, nz, False) for (A, B), nz in zip(pairs, nnzs))
    achieved = alg_bytes / (num_ms * 1e-3) / 1e9
    value = 2.0 * products / (ms_bat * 1e-3) / 1e9
    rec = base_record(args, value, ms_bat, {
        "workload": "512 x C_i=A_i*B_i, n_i log-uniform [256,16384], nnz/row U{4..64}, U[-1,1) FP32",
        "problems": len(pairs), "products": products, "chunks": chunks,
        "step": "whole batch: stacked analysis + fill + numeric per chunk",
        "separate_calls_ms": round(ms_sep, 3), "batched_ms": round(ms_bat, 3),
        "batched_speedup": round(ms_sep / ms_bat, 3),
        "separate_gflops": round(2.0 * products / (ms_sep * 1e-3) / 1e9, 3)})
    rec["config"]["nnz_c"] = int(sum(nnzs))
    rec["roofline"] = {"bound": "hbm", "kernel": "numeric pass of every chunk (stacked problems)",
                       "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBPS, 4),
                       "traffic": measured_traffic(args.workload), "traffic_kind": TRAFFIC_KIND,
                       "algorithmic_bytes": alg_bytes, "numeric_ms": round(num_ms, 4)}
    rec["parity"] = parity
    rec["beside_roofline"] = beside_roofline(products, rec["roofline"]["traffic"], num_ms)
    rec["phases_ms_per_batch"] = {k: round(v[0] / max(args.steps, 1), 4) for k, v in ph.items()}
    rec["cpu_baseline"] = None if args.no_cpu else cpu_baseline_batch(pairs)
    return rec


if __name__ == "__main__":
    main()
