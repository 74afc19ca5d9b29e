"""Writes tests/golden/bootstrap_ci.json: input/output vectors of rv-sparse's
bootstrap speedup interval (bench/analyze.py:170-200, `_bootstrap_speedup_ci`).
The interval depends on the exact sequence of draws from
random.Random(20240517), which no closed form fixes, so its outputs are
recorded from the reference function itself. Only that function is called;
the rest of analyze.py's statistics are checked against closed forms in
tests/test_report_cpu.py. Run where the reference tree is present."""
import importlib.util
import json
import os
import random

REF = os.environ.get("RVSP_REFERENCE_DIR", "/root/reference")
spec = importlib.util.spec_from_file_location("rvsp_ref_analyze", os.path.join(REF, "bench", "analyze.py"))
analyze = importlib.util.module_from_spec(spec)
spec.loader.exec_module(analyze)

gen = random.Random(7)  # the inputs only


def times(n, scale, spread, zeros=0):
    t = [round(scale * (1 + spread * gen.uniform(-1, 1)), 9) for _ in range(n)]
    for i in range(zeros):
        t[i] = 0.0
    return t


cases = [
    (times(5, 0.010, 0.05), times(5, 0.002, 0.05)),     # clear speedup, five runs each
    (times(4, 0.010, 0.30), times(7, 0.009, 0.30)),     # overlapping, uneven counts
    (times(30, 1.0, 0.02), times(12, 0.25, 0.4)),       # many baseline runs
    ([0.5, 0.5, 0.5, 0.5], [0.25, 0.25, 0.25, 0.25]),   # constant: the interval is one point
    (times(6, 0.01, 0.1), times(6, 0.01, 0.1, zeros=3)),  # some resampled medians are 0
    (times(6, 0.01, 0.1), [0.0] * 6),                   # every median 0: no interval
    (times(3, 0.01, 0.1), times(9, 0.01, 0.1)),         # too few baseline runs
    (times(10, 0.02, 0.1), times(3, 0.01, 0.1)),        # too few kernel runs
    ([0.3, 0.1, 0.2, 0.4, 0.1, 0.3], [0.2, 0.2, 0.1, 0.3]),  # ties
]
out = []
for base, kern in cases:
    lo, hi = analyze._bootstrap_speedup_ci(base, kern)
    out.append({"base": base, "kernel": kern, "lo": lo, "hi": hi})
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bootstrap_ci.json")
with open(path, "w") as f:
    json.dump({"iters": analyze.BOOT_ITERS, "seed": analyze.BOOT_SEED, "cases": out}, f, indent=1)
print("wrote", path, len(out), "cases")
