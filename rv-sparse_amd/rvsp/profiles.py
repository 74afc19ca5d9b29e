"""Summaries of rocprofv3 CSV output for profiles/ (kernel-trace statistics
and per-pass HBM traffic). Used by hand after a profiled bench run:

    python -m rvsp.profiles trace  WORKLOAD KT_DIR COMMAND  > profiles/<workload>_kernel_trace.md
    python -m rvsp.profiles traffic WORKLOAD FE_DIR WR_DIR COPY_MIB COMMAND > profiles/traffic_<workload>.md

Traffic: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the last numeric
pass of the run, each scaled by what the counter reads for the run's first
device copy of COPY_MIB MiB (bench.py --copy-probe-mib), whose true traffic
is COPY_MIB MiB read and written. The result is L2-miss traffic (MALL hits
included), the line `numeric_hbm_bytes_per_pass:` that bench.py reads.
"""
from __future__ import annotations

import csv
import glob
import os
import sys

NUMERIC = ("k_num_", "k_interleave")


def short(name: str) -> str:
    n = name.replace("void ", "").split("(")[0]
    return n.replace("rvsp::NumArgsT<float, float, HIP_vector_type<int, 2u> >", "NumArgs")


def trace(workload, kt_dir, command):
    path = glob.glob(os.path.join(kt_dir, "*kernel_stats.csv"))[0]
    rows = list(csv.DictReader(open(path)))
    out = [f"# Kernel trace: {workload}", "", f"Command: `{command}`", "",
           f"Summary: `cd rv-sparse_amd && python -m rvsp.profiles trace {workload} <dir> '<command>'`", "",
           f"From `{os.path.basename(path)}` (all dispatches of the run; the bench line's "
           "phases come from HIP events in the library). Row bins of one pass run on "
           "concurrent streams, so durations of kernels that overlap `k_num_big` include "
           "time waiting for CUs and do not add up to the pass.", "",
           "| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in rows[:25]:
        out.append(f"| `{short(r['Name'])}` | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    print("\n".join(out))


def _dispatches(d):
    path = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    out = {}
    for r in csv.DictReader(open(path)):
        out.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])[1] += float(r["Counter_Value"])
    return [(k, v[0], v[1]) for k, v in sorted(out.items())]


def _last_numeric_pass(disp):
    """Dispatches of the last contiguous run of numeric kernels."""
    idx = [i for i, (_, n, _) in enumerate(disp) if any(t in n for t in NUMERIC)]
    if not idx:
        return []
    end = idx[-1]
    start = end
    while start - 1 >= 0 and any(t in disp[start - 1][1] for t in NUMERIC):
        start -= 1
    return disp[start:end + 1]


def traffic(workload, fe_dir, wr_dir, copy_mib, command):
    fe, wr = _dispatches(fe_dir), _dispatches(wr_dir)
    true_kib = int(copy_mib) * 1024
    copy_fe = next(v for _, n, v in fe if "copyBuffer" in n)
    copy_wr = next(v for _, n, v in wr if "copyBuffer" in n)
    ff, fw = true_kib / copy_fe, true_kib / copy_wr
    pf, pw = _last_numeric_pass(fe), _last_numeric_pass(wr)
    per = {}
    for _, n, v in pf:
        per.setdefault(short(n), [0.0, 0.0])[0] += v * ff * 1024
    for _, n, v in pw:
        per.setdefault(short(n), [0.0, 0.0])[1] += v * fw * 1024
    out = [f"# HBM traffic of one numeric pass: {workload}", "",
           f"Commands: `rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o fe -- {command}`, "
           "then the same with `--pmc WRITE_SIZE` (one counter per run). Summary: "
           f"`cd rv-sparse_amd && python -m rvsp.profiles traffic {workload} <fe dir> <wr dir> "
           f"{copy_mib} '<command>'`.", "",
           f"Scale: the run's first device copy moves {true_kib} KiB each way (`--copy-probe-mib "
           f"{copy_mib}`); it reads FETCH_SIZE {copy_fe:.0f} KiB and WRITE_SIZE {copy_wr:.0f} KiB, "
           f"so the factors are {ff:.4f} (fetch) and {fw:.4f} (write). These are L2-miss bytes "
           "(FETCH_SIZE / WRITE_SIZE count L2 traffic to the fabric, MALL hits included).", "",
           "Numeric pass = the last contiguous run of numeric dispatches in the run (one pass).", "",
           "| kernel | fetched bytes (scaled) | written bytes (scaled) |", "|---|---|---|"]
    for n, (f, w) in sorted(per.items(), key=lambda x: -(x[1][0] + x[1][1])):
        out.append(f"| `{n}` | {f:.4g} | {w:.4g} |")
    tf = sum(v[0] for v in per.values())
    tw = sum(v[1] for v in per.values())
    out += ["", f"numeric_hbm_bytes_per_pass: {int(tf + tw)}", f"fetched: {int(tf)}",
            f"written: {int(tw)}"]
    print("\n".join(out))


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        trace(*sys.argv[2:5])
    else:
        traffic(*sys.argv[2:7])
