"""bench.py --csv: rows in rv-sparse's bench CSV schema (bench/bench.c:531-534,
header compared with the reference's own text in tests/golden) for the GPU
kernels and the CPU scalar kernel, on a small Matrix Market file."""
import argparse
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = open(os.path.join(HERE, "golden", "bench_csv_header.txt")).read().strip()


def test_header_matches_reference_schema():
    import bench
    assert bench.CSV_HEADER == GOLDEN
    assert len(GOLDEN.split(",")) == 23


def test_row_format_matches_bench_c():
    import bench
    meta = {"label": "L", "kernel": "hip_f32", "arm": "hip", "build": "gfx950", "march": "gfx950",
            "cflags": "-O3", "cc_version": "v", "rows": 3, "cols": 4, "nnz_a": 5, "nnz_b": 6,
            "flops": 20.0, "op_mean": 10.0 / 3, "op_max": 6.0, "op_var": 2.0 / 9}
    row = bench.csv_row(meta, 2, 0.5, 7, 1)
    f = row.split(",")
    assert len(f) == 23
    # bench.c: %.0f flops, %.4f op_mean, %.0f op_max, %.4f op_var, %d run,
    # %.9f time, %.4f gops, %d correct, then empty cycles / instructions
    assert f[13:21] == ["20", "3.3333", "6", "0.2222", "2", "0.500000000", "0.0000", "1"]
    assert f[21:] == ["", ""]


@pytest.mark.gpu
def test_csv_rows_for_gpu_and_cpu_kernels(tmp_path):
    import bench
    rng = np.random.default_rng(3)
    n, L = 1500, 12000
    r, c = rng.integers(1, n + 1, L), rng.integers(1, n + 1, L)
    r, c = np.maximum(r, c), np.minimum(r, c)
    v = np.round(rng.uniform(-1, 1, L), 4)
    mtx = tmp_path / "m.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate real symmetric\n"
                   f"{n} {n} {L}\n" + "\n".join(f"{a} {b} {x:.4f}" for a, b, x in zip(r, c, v)) + "\n")
    out = tmp_path / "rows.csv"
    args = argparse.Namespace(workload=f"mtx:{mtx}", csv=str(out), csv_label="t", csv_cpu_runs=1,
                              steps=2, warmup=1)
    bench.bench_csv(args)
    lines = out.read_text().strip().split("\n")
    assert lines[0] == GOLDEN
    rows = [ln.split(",") for ln in lines[1:]]
    assert all(len(x) == 23 for x in rows)
    kernels = [x[1] for x in rows]
    assert kernels.count("hip_f32") == 2 and kernels.count("hip_f32_dev") == 2
    assert sum(k.endswith("scalar_f32") for k in kernels) == 1
    assert all(x[20] == "1" for x in rows)  # every result equals the scalar kernel's
    assert len({x[12] for x in rows}) == 1 and len({x[13] for x in rows}) == 1  # nnz_c, flops
