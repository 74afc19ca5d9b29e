"""Summaries of rv-sparse bench CSV rows (the 23-column per-run schema of
bench/bench.c:531-534, which `bench.py --csv` writes): per (label, build,
cflags, kernel, dtype) group, the run count, time and GOP/s statistics,
the status, and the speedup over the baseline rows of the same label and
dtype with a bootstrap 95% interval. The statistics are the ones rv-sparse's
bench/analyze.py reports, defined the same way, so a GPU row and the CPU
scalar row of one matrix land in one table with comparable numbers:

  time / gops: median, mean, population std, min, max (analyze.py:225-255);
  IQR: median of the upper half minus median of the lower half, 0 below
      four runs (:143-155);
  spread: (max - min) / median (:259-263);
  status: FAIL if any run was wrong, noref if none was checked (:218-223);
  speedup: median(baseline) / median(group); the baseline is the group
      with arm "baseline" (or a named kernel) of the same label and dtype,
      default cflags preferred (:339-378);
  interval: 2000 bootstrap resamples (random.Random(20240517)), each
      drawing the baseline's runs then the group's runs with replacement,
      ratio of medians, 2.5th / 97.5th nearest-rank percentiles;
      significant when the interval excludes 1 (:170-200, :420-440).

Usage: python -m rvsp.report RAW.csv [--csv-out SUMMARY.csv] [--baseline KERNEL]
"""
from __future__ import annotations

import csv
import math
import random
import statistics
import sys
from collections import defaultdict

BOOT_ITERS = 2000
BOOT_SEED = 20240517
DEFAULT_CFLAGS = "-"
NAN = float("nan")


def _num(text, cast=float):
    if text is None or text == "":
        return None
    try:
        return cast(text)
    except (TypeError, ValueError):
        return None


def load_rows(path):
    """Per-run rows with numeric fields parsed; rows without a kernel or
    with an unparsable time / rate / check / nnz_c / flops are skipped."""
    out = []
    with open(path, newline="") as f:
        for raw in csv.DictReader(f):
            if not raw.get("kernel"):
                continue
            rate = raw.get("gops", raw.get("gflops"))
            vals = [_num(raw.get("time_s")), _num(rate), _num(raw.get("correct"), int),
                    _num(raw.get("nnz_c"), int), _num(raw.get("flops"))]
            if any(v is None for v in vals):
                continue
            r = dict(raw)
            r["time_s"], r["gops"], r["correct"], r["nnz_c"], r["flops"] = vals
            r["build"] = (raw.get("build") or "-").strip()
            arm = (raw.get("arm") or "").strip()
            if not arm:
                k = r["kernel"]
                arm = "intrinsic" if k.startswith("rvv_") else "scalar_unroll" if k.startswith(
                    "unroll4_") else "-"
            r["arm"] = arm
            r["march"] = (raw.get("march") or "-").strip()
            r["cflags"] = (raw.get("cflags") or "").strip() or DEFAULT_CFLAGS
            r["cc_version"] = (raw.get("cc_version") or "-").strip()
            for k in ("cycles", "instructions", "op_mean", "op_max", "op_var"):
                r[k] = _num(raw.get(k))
            out.append(r)
    return out


def iqr(values):
    if len(values) < 4:
        return 0.0
    s = sorted(values)
    half = len(s) // 2
    return statistics.median(s[(len(s) + 1) // 2:]) - statistics.median(s[:half])


def nearest_rank(sorted_values, pct):
    if not sorted_values:
        return NAN
    i = int(round(pct / 100.0 * (len(sorted_values) - 1)))
    return sorted_values[min(len(sorted_values) - 1, max(0, i))]


def bootstrap_ci(base_times, times, iters=BOOT_ITERS, seed=BOOT_SEED):
    """95% interval of median(base) / median(times) by resampling both."""
    nb, nk = len(base_times), len(times)
    if nb < 4 or nk < 4:
        return NAN, NAN
    rng = random.Random(seed)
    ratios = []
    for _ in range(iters):
        b = [base_times[rng.randrange(nb)] for _ in range(nb)]
        k = [times[rng.randrange(nk)] for _ in range(nk)]
        mk = statistics.median(k)
        if mk > 0:
            ratios.append(statistics.median(b) / mk)
    if not ratios:
        return NAN, NAN
    ratios.sort()
    return nearest_rank(ratios, 2.5), nearest_rank(ratios, 97.5)


def _stats(prefix, values):
    return {f"{prefix}_median": statistics.median(values), f"{prefix}_mean": statistics.mean(values),
            f"{prefix}_std": statistics.pstdev(values) if len(values) > 1 else 0.0,
            f"{prefix}_min": min(values), f"{prefix}_max": max(values)}


def summarize(rows):
    """{(label, build, cflags, kernel, dtype): entry}."""
    groups = defaultdict(list)
    for r in rows:
        groups[(r["label"], r["build"], r["cflags"], r["kernel"], r["dtype"])].append(r)
    out = {}
    for key, g in groups.items():
        times = [r["time_s"] for r in g]
        checks = [r["correct"] for r in g]
        e = {"arm": g[0]["arm"], "march": g[0]["march"], "cc_version": g[0]["cc_version"],
             "runs": len(g), "times": times, "nnz_c": g[0]["nnz_c"],
             "op_mean": g[0]["op_mean"], "op_max": g[0]["op_max"], "op_var": g[0]["op_var"],
             "status": "FAIL" if 0 in checks else "noref" if all(c == -1 for c in checks) else "ok"}
        e.update(_stats("time", times))
        e["time_iqr"] = iqr(times)
        e.update(_stats("gops", [r["gops"] for r in g]))
        m = e["time_median"]
        e["time_rel_spread"] = (e["time_max"] - e["time_min"]) / m if m > 0 else 0.0
        cyc = [r["cycles"] for r in g if r["cycles"] is not None]
        ins = [r["instructions"] for r in g if r["instructions"] is not None]
        e["cycles_median"] = statistics.median(cyc) if cyc else None
        e["instructions_median"] = statistics.median(ins) if ins else None
        e["ipc"] = (e["instructions_median"] / e["cycles_median"]
                    if e["cycles_median"] and e["instructions_median"] else None)
        e["cycles_per_nnz_c"] = (e["cycles_median"] / e["nnz_c"]
                                 if e["cycles_median"] and e["nnz_c"] > 0 else None)
        out[key] = e
    return out


def default_build(summary):
    """The build a named baseline kernel is taken from unless one is given:
    "gc" when present, else the only build (analyze.py:106-128)."""
    builds = {k[1] for k in summary}
    return next(iter(builds)) if len(builds) == 1 and "gc" not in builds else "gc"


def kernel_build(summary, kernel):
    """The build a named baseline kernel is taken from: the default build
    when the kernel ran there, else its only build (analyze.py:756-782)."""
    builds = sorted({k[1] for k in summary if k[3] == kernel})
    base = default_build(summary)
    if base in builds:
        return base
    if len(builds) == 1:
        return builds[0]
    raise ValueError(f"baseline kernel {kernel!r} ran in builds {builds}, none of them {base!r}: "
                     "the denominator is ambiguous")


def attach_speedups(summary, baseline_kernel=None, baseline_build=None):
    """Adds speedup, speedup_lo / _hi, significant and speedup_same_build to
    every entry; returns the (label, dtype) pairs that had no baseline."""
    bbuild = baseline_build or (kernel_build(summary, baseline_kernel) if baseline_kernel else None)

    def is_base(key, e):
        if baseline_kernel is not None:
            return key[3] == baseline_kernel and key[1] == bbuild
        return e["arm"] == "baseline"

    base, same_build = {}, {}
    for key, e in sorted(summary.items(), key=lambda kv: kv[0]):
        if not is_base(key, e):
            continue
        label, build, cflags, _, dtype = key
        # the first baseline group in key order, replaced by a default-cflags one
        for table, k in ((same_build, (label, build, dtype)), (base, (label, dtype))):
            if k not in table or cflags == DEFAULT_CFLAGS:
                table[k] = e
    missing = set()
    for key, e in summary.items():
        label, build, cflags, _, dtype = key
        m = e["time_median"]
        sb = same_build.get((label, build, dtype))
        e["speedup_same_build"] = sb["time_median"] / m if sb and sb["time_median"] and m > 0 else NAN
        b = base.get((label, dtype))
        if b is None:
            missing.add((label, dtype))
            e.update(speedup=NAN, speedup_lo=NAN, speedup_hi=NAN, significant=None)
            continue
        e["speedup"] = b["time_median"] / m if m > 0 else NAN
        if is_base(key, e) and cflags == DEFAULT_CFLAGS:
            e.update(speedup_lo=1.0, speedup_hi=1.0, significant=None)
            continue
        lo, hi = bootstrap_ci(b["times"], e["times"])
        e["speedup_lo"], e["speedup_hi"] = lo, hi
        e["significant"] = None if math.isnan(lo) or math.isnan(hi) else (lo > 1.0 or hi < 1.0)
    return missing


def _order(summary):
    return sorted(summary, key=lambda k: (k[0], k[4], k[1], k[3], k[2]))


# analyze.py:587-623, the summary CSV's columns
SUMMARY_COLUMNS = [
    "matrix", "build", "arm", "kernel", "cflags", "dtype", "runs", "time_median_s", "time_mean_s",
    "time_std_s", "time_min_s", "time_max_s", "time_iqr_s", "time_rel_spread", "gops_median",
    "gops_mean", "gops_std", "gops_min", "gops_max", "speedup_vs_gc_scalar", "speedup_ci_lo",
    "speedup_ci_hi", "speedup_significant", "speedup_vs_scalar_same_build", "nnz_c", "op_mean",
    "op_max", "op_var", "cycles_median", "instructions_median", "ipc", "cycles_per_nnz_c", "march",
    "cc_version", "status"]


def write_summary_csv(summary, path):
    def fnum(v, spec=".4f"):
        return format(v, spec) if v is not None and v == v else ""

    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(SUMMARY_COLUMNS)
        for key in _order(summary):
            label, build, cflags, kernel, dtype = key
            e = summary[key]
            sig = e.get("significant")
            w.writerow([label, build, e["arm"], kernel, cflags, dtype, e["runs"]] +
                       [format(e[k], ".9f") for k in ("time_median", "time_mean", "time_std",
                                                      "time_min", "time_max", "time_iqr")] +
                       [format(e["time_rel_spread"], ".4f")] +
                       [format(e[k], ".4f") for k in ("gops_median", "gops_mean", "gops_std",
                                                      "gops_min", "gops_max")] +
                       [fnum(e["speedup"]), fnum(e.get("speedup_lo")), fnum(e.get("speedup_hi")),
                        "" if sig is None else ("yes" if sig else "no"),
                        fnum(e.get("speedup_same_build")), e["nnz_c"],
                        fnum(e["op_mean"]), fnum(e["op_max"], ".0f"), fnum(e["op_var"]),
                        fnum(e["cycles_median"], ".0f"), fnum(e["instructions_median"], ".0f"),
                        fnum(e["ipc"]), fnum(e["cycles_per_nnz_c"]), e["march"], e["cc_version"],
                        e["status"]])


def table(summary):
    """The summary as printable lines, one group per line."""
    head = (f"{'matrix':<20} {'kernel':<16} {'build':<10} {'dt':<4} {'n':>3} {'median_ms':>10} "
            f"{'spread':>7} {'gops':>9} {'speedup':>9} {'95% CI':>17} {'sig':>4} {'st':>5}")
    lines = [head, "-" * len(head)]
    for key in _order(summary):
        label, build, _, kernel, dtype = key
        e = summary[key]
        sp = e.get("speedup", NAN)
        lo, hi = e.get("speedup_lo", NAN), e.get("speedup_hi", NAN)
        ci = f"[{lo:.2f},{hi:.2f}]" if lo == lo and hi == hi else "-"
        sig = e.get("significant")
        lines.append(f"{label:<20} {kernel:<16} {build:<10} {dtype:<4} {e['runs']:>3} "
                     f"{e['time_median'] * 1e3:>10.3f} {e['time_rel_spread'] * 100:>6.1f}% "
                     f"{e['gops_median']:>9.3f} {(f'{sp:.2f}x' if sp == sp else '-'):>9} {ci:>17} "
                     f"{('-' if sig is None else 'sig' if sig else 'ns'):>4} {e['status']:>5}")
    return lines


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        print("usage: python -m rvsp.report RAW.csv [--csv-out SUMMARY.csv] [--baseline KERNEL]")
        return 2
    path = argv[0]
    out = argv[argv.index("--csv-out") + 1] if "--csv-out" in argv else None
    bk = argv[argv.index("--baseline") + 1] if "--baseline" in argv else None
    rows = load_rows(path)
    if not rows:
        print("no valid rows in", path)
        return 1
    s = summarize(rows)
    attach_speedups(s, baseline_kernel=bk)
    print("\n".join(table(s)))
    if out:
        write_summary_csv(s, out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
