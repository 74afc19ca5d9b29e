"""rvsp.report, the summary of rv-sparse bench CSV rows: the bootstrap
interval against vectors recorded from rv-sparse's bench/analyze.py
(tests/golden/bootstrap_ci.json), the other statistics against closed
forms, the baseline choice rules, and the summary CSV's columns against the
reference's own list (tests/golden/summary_csv_header.txt). No GPU."""
import csv
import json
import math
import os
import random
import statistics

import pytest

from rvsp import report

HERE = os.path.dirname(os.path.abspath(__file__))
BOOT = json.load(open(os.path.join(HERE, "golden", "bootstrap_ci.json")))
HEADER = open(os.path.join(HERE, "golden", "bench_csv_header.txt")).read().strip()


def same_float(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


def test_bootstrap_interval_matches_recorded_vectors():
    assert (BOOT["iters"], BOOT["seed"]) == (report.BOOT_ITERS, report.BOOT_SEED)
    for c in BOOT["cases"]:
        lo, hi = report.bootstrap_ci(c["base"], c["kernel"])
        assert same_float(lo, c["lo"]) and same_float(hi, c["hi"]), c


def test_vectors_catch_a_changed_draw_order():
    """Drawing the group's runs before the baseline's (a plausible slip)
    moves the interval of the 30-against-12-runs case: the vectors pin the
    order. (With few runs each, both orders can land on the same two of the
    few possible ratios.)"""
    c = BOOT["cases"][2]
    assert (len(c["base"]), len(c["kernel"])) == (30, 12)
    rng = random.Random(report.BOOT_SEED)
    base, kern = c["base"], c["kernel"]
    ratios = []
    for _ in range(report.BOOT_ITERS):
        k = [kern[rng.randrange(len(kern))] for _ in kern]
        b = [base[rng.randrange(len(base))] for _ in base]
        ratios.append(statistics.median(b) / statistics.median(k))
    ratios.sort()
    swapped = (report.nearest_rank(ratios, 2.5), report.nearest_rank(ratios, 97.5))
    assert swapped != (c["lo"], c["hi"])


def test_closed_forms():
    assert report.iqr([4, 1, 3, 2, 5]) == 4.5 - 1.5  # halves [1,2] and [4,5]
    assert report.iqr([1, 2, 3, 4]) == 3.5 - 1.5
    assert report.iqr([1, 2, 3]) == 0.0
    v = list(range(101))
    assert report.nearest_rank(v, 2.5) == 2 and report.nearest_rank(v, 97.5) == 98
    assert report.nearest_rank([7.0], 97.5) == 7.0
    assert math.isnan(report.nearest_rank([], 50))
    # a constant group against a constant baseline: one-point interval at the ratio
    assert report.bootstrap_ci([6.0] * 5, [2.0] * 4) == (3.0, 3.0)
    assert all(math.isnan(x) for x in report.bootstrap_ci([1.0] * 3, [1.0] * 5))


def write_rows(path, rows):
    with open(path, "w") as f:
        f.write(HEADER + "\n")
        for r in rows:
            f.write(",".join(str(r.get(k, "")) for k in HEADER.split(",")) + "\n")


def row(label, kernel, arm, build, t, correct=1, cflags="-", dtype="f32", flops=2e6):
    return {"label": label, "kernel": kernel, "arm": arm, "build": build, "march": build,
            "cflags": cflags, "cc_version": "cc", "dtype": dtype, "rows": 10, "cols": 10,
            "nnz_a": 5, "nnz_b": 5, "nnz_c": 7, "flops": flops, "op_mean": 1.5, "op_max": 3,
            "op_var": 0.25, "run": 0, "time_s": t, "gops": flops / t / 1e9 if t else 0.0,
            "correct": correct, "cycles": "", "instructions": ""}


def test_summary_and_speedups(tmp_path):
    rows = [row("m", "scalar_f32", "baseline", "cpu", t) for t in (4, 1, 3, 2, 5)]
    rows += [row("m", "hip_f32", "hip", "gfx950", t) for t in (1.0, 1.0, 1.0, 1.0)]
    rows += [row("m", "bad", "hip", "gfx950", t, correct=c) for t, c in ((2, 1), (2, 0))]
    rows += [row("m", "unchecked", "hip", "gfx950", 1.5, correct=-1)]
    rows += [row("other", "hip_f32", "hip", "gfx950", 1.0)]  # no baseline for this label
    p = tmp_path / "raw.csv"
    write_rows(p, rows)
    s = report.summarize(report.load_rows(p))
    missing = report.attach_speedups(s)
    assert missing == {("other", "f32")}
    base = s[("m", "cpu", "-", "scalar_f32", "f32")]
    assert base["runs"] == 5 and base["time_median"] == 3 and base["time_mean"] == 3
    assert base["time_std"] == statistics.pstdev([1, 2, 3, 4, 5]) and base["time_iqr"] == 3.0
    assert base["time_rel_spread"] == 4 / 3
    assert base["speedup"] == 1.0 and (base["speedup_lo"], base["speedup_hi"]) == (1.0, 1.0)
    assert base["significant"] is None and base["status"] == "ok"
    hip = s[("m", "gfx950", "-", "hip_f32", "f32")]
    assert hip["speedup"] == 3.0 and hip["gops_median"] == 2e6 / 1.0 / 1e9
    assert (hip["speedup_lo"], hip["speedup_hi"]) == report.bootstrap_ci([4, 1, 3, 2, 5], [1.0] * 4)
    lo, hi = hip["speedup_lo"], hip["speedup_hi"]
    assert 1.0 <= lo <= 3.0 <= hi <= 5.0  # resampled baseline medians lie in [1, 5]
    assert hip["significant"] is (lo > 1.0 or hi < 1.0)
    assert math.isnan(hip["speedup_same_build"])  # no baseline in the gfx950 build
    assert s[("m", "gfx950", "-", "bad", "f32")]["status"] == "FAIL"
    assert s[("m", "gfx950", "-", "unchecked", "f32")]["status"] == "noref"
    assert math.isnan(s[("other", "gfx950", "-", "hip_f32", "f32")]["speedup"])


def test_default_cflags_baseline_is_preferred(tmp_path):
    rows = [row("m", "scalar_f32", "baseline", "gc", 8.0, cflags="-DX=1") for _ in range(4)]
    rows += [row("m", "scalar_f32", "baseline", "gc", 4.0) for _ in range(4)]
    rows += [row("m", "hip_f32", "hip", "gc", 2.0) for _ in range(4)]
    p = tmp_path / "raw.csv"
    write_rows(p, rows)
    s = report.summarize(report.load_rows(p))
    report.attach_speedups(s)
    assert s[("m", "gc", "-", "hip_f32", "f32")]["speedup"] == 2.0
    assert s[("m", "gc", "-", "hip_f32", "f32")]["speedup_same_build"] == 2.0
    tuned = s[("m", "gc", "-DX=1", "scalar_f32", "f32")]
    assert tuned["speedup"] == 0.5 and tuned["speedup_lo"] == tuned["speedup_hi"] == 0.5


def test_named_baseline_kernel(tmp_path):
    rows = [row("m", "hip_f32", "hip", "gfx950", 2.0) for _ in range(4)]
    rows += [row("m", "hip_f32_dev", "hip", "gfx950", 1.0) for _ in range(4)]
    rows += [row("m", "scalar_f32", "baseline", "cpu", 8.0) for _ in range(4)]
    p = tmp_path / "raw.csv"
    write_rows(p, rows)
    s = report.summarize(report.load_rows(p))
    report.attach_speedups(s, baseline_kernel="hip_f32")  # its only build, gfx950
    assert s[("m", "gfx950", "-", "hip_f32_dev", "f32")]["speedup"] == 2.0
    assert s[("m", "cpu", "-", "scalar_f32", "f32")]["speedup"] == 0.25
    rows += [row("m", "hip_f32", "hip", "other", 2.0) for _ in range(4)]
    write_rows(p, rows)
    s = report.summarize(report.load_rows(p))
    with pytest.raises(ValueError):
        report.attach_speedups(s, baseline_kernel="hip_f32")  # two builds, neither "gc"


def test_summary_csv_columns_match_reference(tmp_path):
    ref = open(os.path.join(HERE, "golden", "summary_csv_header.txt")).read().strip().split(",")
    assert report.SUMMARY_COLUMNS == ref
    rows = [row("m", "scalar_f32", "baseline", "cpu", t) for t in (4, 1, 3, 2, 5)]
    rows += [row("m", "hip_f32", "hip", "gfx950", t) for t in (1.0, 1.1, 0.9, 1.0)]
    p, out = tmp_path / "raw.csv", tmp_path / "summary.csv"
    write_rows(p, rows)
    assert report.main([str(p), "--csv-out", str(out)]) == 0
    got = list(csv.reader(open(out)))
    assert got[0] == ref and len(got) == 3 and all(len(r) == len(ref) for r in got)
    hip = dict(zip(ref, got[2]))
    assert hip["kernel"] == "hip_f32" and hip["speedup_vs_gc_scalar"] == "3.0000"
    assert hip["time_median_s"] == "1.000000000" and hip["status"] == "ok"


def test_reads_bench_py_rows(tmp_path):
    import bench
    meta = {"label": "gen_64x64_d0.05", "kernel": "hip_f32", "arm": "hip", "build": "gfx950",
            "march": "gfx950", "cflags": "-O3", "cc_version": "v", "rows": 64, "cols": 64,
            "nnz_a": 5, "nnz_b": 6, "flops": 2e6, "op_mean": 1.0, "op_max": 2.0, "op_var": 0.5}
    p = tmp_path / "raw.csv"
    p.write_text(bench.CSV_HEADER + "\n" + "\n".join(bench.csv_row(meta, r, 0.001 * (r + 1), 9, 1)
                                                     for r in range(4)) + "\n")
    rows = report.load_rows(p)
    assert len(rows) == 4 and rows[0]["label"] == "gen_64x64_d0.05" and rows[0]["cflags"] == "-O3"
    s = report.summarize(rows)
    (e,) = s.values()
    assert e["time_median"] == 0.0025 and e["nnz_c"] == 9
