"""Products of rv-sparse's generated bench inputs on the GPU: the pairs of its
`bench --gen R C DENSITY SEED` sweep (run_bench.sh:185-190; A = R x C,
B = C x R from the library's genmat), C = A*B through rvsp_spgemm_csr,
bit-exact against the oracle (oracle.c, and rv-sparse's own scalar kernel in
oracle/_ref when built); the FP64 and INT8 bench inputs through the legacy
dtypes against rv-sparse's legacy kernels; and `bench.py --csv` on a
generated pair, rows labelled as run_bench.sh labels them, with the
analyze.py speedup summary."""
import argparse
import ctypes as C
import json

import numpy as np
import pytest

import rvsp
from oracle import pyoracle

pytestmark = pytest.mark.gpu

GEN_CASES = [(512, 512, 0.02, 42), (1024, 1024, 0.01, 42), (2048, 2048, 0.005, 42),
             (4096, 4096, 0.002, 42), (300, 7000, 0.01, 5)]


def assert_bitexact(got, want, what):
    rp, ci, va = want[:3]
    np.testing.assert_array_equal(got.row_ptr, rp, err_msg=f"{what} row_ptr")
    np.testing.assert_array_equal(got.col_idx, ci, err_msg=f"{what} col_idx")
    np.testing.assert_array_equal(got.values.view(np.uint32), np.asarray(va, np.float32).view(np.uint32),
                                  err_msg=f"{what} values")


@pytest.mark.parametrize("R,Cc,D,S", GEN_CASES)
def test_generated_pairs_match_oracle(R, Cc, D, S):
    print(f"gen {R}x{Cc} d{D} seed {S}", flush=True)
    A, B = rvsp.gen_bench_pair(R, Cc, D, S)
    assert (A.rows, A.cols, B.rows, B.cols) == (R, Cc, Cc, R)
    Cg = rvsp.spgemm_csr(A, B)[1]
    assert_bitexact(Cg, pyoracle.spgemm(A, B), "oracle")
    if pyoracle.ref_available():
        assert_bitexact(Cg, pyoracle.ref_spgemm(A, B), "rv-sparse scalar")


# larger generated pairs: log-normal degrees put whole rows of C at full
# width (16384, 32768 columns), so they cross the dense and big-row paths
# (the 16384^2 pair's near-dense rows fit the one-pass bitmap's span but
# take the big-row path by their product count)
LARGE_CASES = [(16384, 16384, 0.001, 7), (32768, 32768, 0.0005, 9), (8192, 40000, 0.0008, 3)]


@pytest.mark.parametrize("R,Cc,D,S", LARGE_CASES)
def test_large_generated_pairs_match_oracle(R, Cc, D, S):
    print(f"gen large {R}x{Cc} d{D} seed {S}", flush=True)
    A, B = rvsp.gen_bench_pair(R, Cc, D, S)
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    want = pyoracle.spgemm(A, B)
    if (R, Cc) == (16384, 16384):
        # above both thresholds for running row bins on concurrent streams
        # (2^21 products, 2^24 outputs): count, fill and numeric all fan out
        products = int(np.diff(B.row_ptr)[A.col_idx].astype(np.int64).sum())
        assert products >= 1 << 21 and int(want[0][-1]) >= 1 << 24
    for algo in (rvsp.ALGO_HIP, rvsp.ALGO_HIP_LEARNED):
        assert_bitexact(rvsp.spgemm_device(dA, dB, algo=algo), want, f"oracle algo {algo}")


def legacy_product(lib, a, b, din, dout):
    c = rvsp.CsrStruct()
    opt = rvsp.Options(rvsp.BACKEND_SCALAR, din, dout)
    st = lib.rvsp_spgemm_csr(C.byref(a), C.byref(b), C.byref(c), C.byref(opt))
    assert st == 0, st
    n = c.nnz
    ct, dt = (C.c_double, np.float64) if dout == rvsp.DTYPE_FP64 else (C.c_int32, np.int32)
    rp = np.ctypeslib.as_array(C.cast(c.row_ptr, C.POINTER(C.c_int32)), (c.rows + 1,)).copy()
    ci = np.ctypeslib.as_array(C.cast(c.col_idx, C.POINTER(C.c_int32)), (n,)).copy() if n else np.zeros(0, np.int32)
    va = np.ctypeslib.as_array(C.cast(c.values, C.POINTER(ct)), (n,)).copy() if n else np.zeros(0, dt)
    lib.rvsp_csr_destroy(C.byref(c))
    return rp, ci, va


@pytest.mark.parametrize("din,dout", [(rvsp.DTYPE_FP64, rvsp.DTYPE_FP64), (rvsp.DTYPE_INT8, rvsp.DTYPE_INT32)])
def test_generated_legacy_dtypes_match_reference(din, dout):
    if not pyoracle.ref_available():
        pytest.skip("rv-sparse reference library not built (oracle/build_ref.sh)")
    L, R = rvsp.lib(), pyoracle.ref()
    R.rvsp_spgemm_csr.restype = C.c_int
    R.rvsp_spgemm_csr.argtypes = [C.c_void_p] * 4
    R.rvsp_csr_destroy.restype = None
    R.rvsp_csr_destroy.argtypes = [C.c_void_p]
    a, b = rvsp.CsrStruct(), rvsp.CsrStruct()
    print(f"gen legacy dtype {din}", flush=True)
    assert L.rvsp_genmat_bench_pair(1024, 1024, 0.01, 42, din, C.byref(a), C.byref(b)) == 0
    got = legacy_product(L, a, b, din, dout)
    want = legacy_product(R, a, b, din, dout)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g.view(np.uint64) if g.dtype == np.float64 else g,
                                      w.view(np.uint64) if w.dtype == np.float64 else w)
    L.rvsp_csr_destroy(C.byref(a))
    L.rvsp_csr_destroy(C.byref(b))


def test_bench_csv_on_generated_pair(tmp_path, capsys):
    import bench
    out = tmp_path / "rows.csv"
    args = argparse.Namespace(workload="gen:512:512:0.02:42", csv=str(out), csv_label=None,
                              csv_cpu_runs=4, steps=4, warmup=1)
    bench.bench_csv(args)
    line = [x for x in capsys.readouterr().out.strip().split("\n") if x.startswith("{")][-1]
    res = json.loads(line)
    assert res["label"] == "gen_512x512_d0.02"  # run_bench.sh:1432
    rows = [r.split(",") for r in out.read_text().strip().split("\n")[1:]]
    assert all(r[0] == "gen_512x512_d0.02" and r[20] == "1" for r in rows)
    sp = res["speedup_vs_scalar"]
    (cpu,) = [k for k in sp if k.endswith("scalar_f32")]  # rv-sparse's kernel, or the oracle's port
    assert sp[cpu]["x"] == 1.0 and sp[cpu]["ci95"] != [None, None]
    for k in ("hip_f32", "hip_f32_dev"):
        lo, hi = sp[k]["ci95"]
        assert sp[k]["status"] == "ok" and lo is not None and lo <= sp[k]["x"] <= hi
