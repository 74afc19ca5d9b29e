"""rv-sparse's own callers of rv_sparse.h, compiled unchanged.

oracle/build_ref.sh compiles rv-sparse's two API tests
(tests/test_spgemm_reference_f32.c, tests/test_spgemm_descr_f32.c) and its
three examples (examples/spgemm_csr_f32.c, spgemm_csr_descr_f32.c,
spgemm_csr_reuse_f32.c) from where they lie, twice each: against this
build's include/ linked to librvsparse_hip.so (`.hip`), and against
rv-sparse's own include/ linked to its scalar library (`.ref`, CPU).

The GPU test runs both and requires exit status 0, no FAIL line, the tests'
PASS summaries, and byte-identical output from the two builds: the same
caller source gets the same answers from the HIP library. The binaries are
built in the container that holds the reference sources; where they are
absent the tests skip.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLERS = os.path.join(ROOT, "oracle", "_ref", "callers")
NAMES = ["test_spgemm_reference_f32", "test_spgemm_descr_f32", "spgemm_csr_f32",
         "spgemm_csr_descr_f32", "spgemm_csr_reuse_f32"]


def _bin(name, kind):
    p = os.path.join(CALLERS, f"{name}.{kind}")
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (needs rv-sparse's sources at build time)")
    return p


@pytest.mark.parametrize("name", NAMES)
def test_caller_links_against_hip_library(name):
    """Every rv_sparse.h symbol the unchanged caller uses resolves in
    librvsparse_hip.so (no CPU library in between)."""
    exe = _bin(name, "hip")
    out = subprocess.run(["ldd", "-r", exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "librvsparse_hip.so" in out.stdout
    assert "not found" not in out.stdout and "undefined symbol" not in (out.stdout + out.stderr)
    assert "librvsp_ref" not in out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_caller_runs_unchanged_on_gpu(name):
    hip = subprocess.run([_bin(name, "hip")], capture_output=True, text=True, timeout=120)
    ref = subprocess.run([_bin(name, "ref")], capture_output=True, text=True, timeout=120)
    print(hip.stdout, flush=True)
    assert ref.returncode == 0, ref.stdout + ref.stderr
    assert hip.returncode == 0, hip.stdout + hip.stderr
    assert "FAIL" not in hip.stdout
    if name.startswith("test_"):
        assert f"{name}: PASS" in hip.stdout
    if name == "test_spgemm_descr_f32":
        assert hip.stdout.count("[ok  ]") >= 27
    assert hip.stdout == ref.stdout
