"""Real-matrix path: a Matrix Market file read by rvsp_mtx_load, then
C = A*A on the GPU (host one-shot and device one-shot), bit-exact against
the CPU oracle, as rv-sparse's bench runs its SuiteSparse A*A cases
(bench/bench.c:722-736). The files are synthetic: no matrix collection is
available offline."""
import numpy as np
import pytest

import rvsp
from oracle import pyoracle
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu


def write_mtx(path, n, L, seed, symmetric, pattern=False):
    rng = np.random.default_rng(seed)
    r = rng.integers(1, n + 1, L)
    c = rng.integers(1, n + 1, L)
    if symmetric:
        r, c = np.maximum(r, c), np.minimum(r, c)
    field = "pattern" if pattern else "real"
    kind = "symmetric" if symmetric else "general"
    lines = [f"%%MatrixMarket matrix coordinate {field} {kind}", "% synthetic", f"{n} {n} {L}"]
    if pattern:
        lines += [f"{a} {b}" for a, b in zip(r, c)]
    else:
        v = np.round(rng.uniform(-1, 1, L), 4)
        lines += [f"{a} {b} {x:.4f}" for a, b, x in zip(r, c, v)]
    path.write_text("\n".join(lines) + "\n")
    return str(path)


@pytest.mark.parametrize("symmetric,pattern", [(True, False), (False, False), (True, True)])
def test_mtx_a_times_a(tmp_path, symmetric, pattern):
    path = write_mtx(tmp_path / "a.mtx", 3000, 40000, 5, symmetric, pattern)
    A = rvsp.mtx_load(path)
    if A.cols != A.rows:  # A*A needs a square pattern (bench.c's inferred cols)
        pytest.skip("inferred column count differs from the row count")
    print(f"mtx A*A n={A.rows} nnz={A.nnz}", flush=True)
    rp, ci, va = pyoracle.spgemm(A, A)[:3]
    _, C_ = rvsp.spgemm_csr(A, A)
    assert_bitexact(C_, rp, ci, va, "mtx host one-shot")
    dA = rvsp.DeviceCsr(A)
    assert_bitexact(rvsp.spgemm_device(dA, dA), rp, ci, va, "mtx device one-shot")
