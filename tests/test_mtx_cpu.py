"""Matrix Market reading (rvsp_mtx_load), checked bit for bit against the
restatement in oracle/mtx.py and against rv-sparse's own reader
(tools/src/mtx_to_csr_formatter.c, built into oracle/_ref when the reference
is present), on synthetic files that cover the banner variants, comments,
unsorted and repeated entries, symmetric mirroring, short files and the
error cases."""
import numpy as np
import pytest

import rvsp
from oracle import mtx as omtx


def write(tmp_path, name, banner, dims, entries, comments=("% generated",)):
    p = tmp_path / name
    lines = [banner] + list(comments) + [dims] + entries
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def check_against_oracles(path):
    got = rvsp.mtx_load(path)
    M, N, rp, ci, va = omtx.load(path)
    assert (got.rows, got.cols) == (M, N)
    np.testing.assert_array_equal(got.row_ptr, rp)
    np.testing.assert_array_equal(got.col_idx, ci)
    np.testing.assert_array_equal(got.values.view(np.uint32), va.view(np.uint32))
    if omtx.ref_available():
        ref = omtx.ref_load(path)
        assert ref is not None
        rM, rN, rrp, rci, rva = ref
        assert got.rows == rM and (got.nnz == 0 or got.cols == rN)
        np.testing.assert_array_equal(got.row_ptr, rrp)
        np.testing.assert_array_equal(got.col_idx, rci)
        np.testing.assert_array_equal(got.values.view(np.uint32), rva.view(np.uint32))
    # canonical CSR whatever the file held
    assert got.row_ptr[0] == 0 and np.all(np.diff(got.row_ptr) >= 0)
    for i in range(got.rows):
        row = got.col_idx[got.row_ptr[i]:got.row_ptr[i + 1]]
        assert np.all(np.diff(row) > 0)
    return got


def test_general_real_unsorted_with_duplicates(tmp_path):
    ents = ["3 1 0.5", "1 3 -2", "1 1 0.1", "2 2 4", "1 1 0.2", "1 1 0.3", "3 3 1e-3", "1 2 7.25"]
    path = write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate real general", "3 3 8", ents,
                 comments=("% a comment", "%another"))
    A = check_against_oracles(path)
    # closed form: repeated (1,1) summed in file order, in float
    s = np.float32(np.float32(np.float32(0.1) + np.float32(0.2)) + np.float32(0.3))
    assert A.values[0].view(np.uint32) == s.view(np.uint32)
    np.testing.assert_array_equal(A.row_ptr, [0, 3, 4, 6])
    np.testing.assert_array_equal(A.col_idx, [0, 1, 2, 1, 0, 2])


def test_symmetric_mirrors_off_diagonal(tmp_path):
    ents = ["1 1 2", "2 1 -1", "3 1 0.75", "3 2 -1", "3 3 2", "2 2 5"]
    path = write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate real symmetric", "3 3 6", ents)
    A = check_against_oracles(path)
    assert A.nnz == 3 + 2 * 3  # three diagonal entries, three mirrored pairs
    D = np.zeros((3, 3), np.float32)
    for i in range(3):
        for p in range(A.row_ptr[i], A.row_ptr[i + 1]):
            D[i, A.col_idx[p]] = A.values[p]
    np.testing.assert_array_equal(D, D.T)


def test_pattern_files_get_ones(tmp_path):
    path = write(tmp_path, "p.mtx", "%%MatrixMarket matrix coordinate pattern general", "4 5 4",
                 ["1 5", "4 2", "2 3", "1 1"])
    A = check_against_oracles(path)
    assert np.all(A.values == 1.0) and A.cols == 5
    path = write(tmp_path, "ps.mtx", "%%MatrixMarket matrix coordinate pattern symmetric", "3 3 3",
                 ["2 1", "3 3", "3 1"])
    A = check_against_oracles(path)
    assert A.nnz == 5 and np.all(A.values == 1.0)


def test_integer_field_short_file_and_inferred_columns(tmp_path):
    # 5 declared entries, 3 present: reading stops at the end; the last two
    # columns are empty, so cols = 3 (largest column + 1), not the declared 5
    path = write(tmp_path, "i.mtx", "%%MatrixMarket matrix coordinate integer general", "4 5 5",
                 ["1 1 3", "4 3 -7", "2 2 1"])
    A = check_against_oracles(path)
    assert A.nnz == 3 and A.cols == 3 and A.rows == 4


@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_files_match_reader_and_reference(tmp_path, sym, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 200))
    L = int(rng.integers(50, 2000))
    r = rng.integers(1, n + 1, L)
    c = rng.integers(1, n + 1, L)
    if sym:
        r, c = np.maximum(r, c), np.minimum(r, c)  # lower triangle
    # three-decimal values: no decimal sits near a float32 rounding boundary
    v = np.round(rng.uniform(-10, 10, L), 3)
    ents = [f"{a} {b} {x:.3f}" for a, b, x in zip(r, c, v)]
    banner = "%%MatrixMarket matrix coordinate real " + ("symmetric" if sym else "general")
    path = write(tmp_path, "r.mtx", banner, f"{n} {n} {L}", ents)
    check_against_oracles(path)


def test_errors(tmp_path):
    L = rvsp.lib()
    st, A = rvsp.mtx_load(str(tmp_path / "missing.mtx"), raise_on_error=False)
    assert st == rvsp.ERROR_INVALID_ARGUMENT and A is None
    bad_row = write(tmp_path, "b.mtx", "%%MatrixMarket matrix coordinate real general", "2 2 2",
                    ["1 1 1", "3 1 1"])
    assert rvsp.mtx_load(bad_row, raise_on_error=False)[0] == rvsp.ERROR_INVALID_ARGUMENT
    zero_col = write(tmp_path, "z.mtx", "%%MatrixMarket matrix coordinate real general", "2 2 1",
                     ["1 0 1"])
    assert rvsp.mtx_load(zero_col, raise_on_error=False)[0] == rvsp.ERROR_INVALID_ARGUMENT
    no_dims = tmp_path / "n.mtx"
    no_dims.write_text("%%MatrixMarket matrix coordinate real general\n% only comments\n")
    assert rvsp.mtx_load(str(no_dims), raise_on_error=False)[0] == rvsp.ERROR_INVALID_ARGUMENT
    cplx = write(tmp_path, "c.mtx", "%%MatrixMarket matrix coordinate complex general", "2 2 1",
                 ["1 1 1 2"])
    assert rvsp.mtx_load(cplx, raise_on_error=False)[0] == rvsp.ERROR_INVALID_ARGUMENT
    import ctypes as C
    assert L.rvsp_mtx_load(None, C.byref(rvsp.CsrStruct())) == rvsp.ERROR_NULL_POINTER


def test_oracle_restatement_catches_unstable_order(tmp_path):
    """A planted mistake in the restatement (repeated columns summed in the
    opposite order) is visible on a file whose sum depends on the order."""
    ents = ["1 1 1e8", "1 1 -1e8", "1 1 1"]
    path = write(tmp_path, "o.mtx", "%%MatrixMarket matrix coordinate real general", "1 1 3", ents)
    _, _, _, _, va = omtx.load(path)
    file_order = np.float32(np.float32(np.float32(1e8) + np.float32(-1e8)) + np.float32(1))
    reversed_order = np.float32(np.float32(np.float32(1) + np.float32(-1e8)) + np.float32(1e8))
    assert file_order != reversed_order
    assert va[0] == file_order
    assert rvsp.mtx_load(path).values[0] == file_order
