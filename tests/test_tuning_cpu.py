"""The tuning table: class function, file format and value checks (CPU;
nothing here launches a kernel)."""
import os

import pytest

import rvsp

GOOD = """# a comment
0 2048 16 512 16384 2048 2097152 16777216
5 4096 32 256 32768 1024 0 4611686018427387904
"""


def load(tmp_path, text):
    p = tmp_path / "table.txt"
    p.write_text(text)
    return rvsp.lib().rvsp_tuning_load_file(str(p).encode())


@pytest.fixture(autouse=True)
def restore(tmp_path):
    yield
    # back to the shipped table (or the defaults when none ships)
    here = os.path.dirname(rvsp.LIB_PATH)
    shipped = os.path.join(here, "..", "tuning", "table.txt")
    if os.path.exists(shipped):
        assert rvsp.lib().rvsp_tuning_load_file(shipped.encode()) == 0
    else:
        assert load(tmp_path, "") == 0
    rvsp.tuning_force(None)


def test_class_buckets():
    cls = rvsp.lib().rvsp_tuning_class
    # estimated products nnz(A) * nnz(B) / k: <2^21, <2^25, <2^29, more; x (n > 16384)
    assert cls(1000, 1000, 1000, 1000) == 0
    assert cls(1000, 1000, 1000, 16385) == 1
    assert cls(1 << 20, 1 << 24, 1 << 20, 100) == 2      # 2^24 products
    assert cls(1 << 20, 1 << 28, 1 << 20, 1 << 20) == 5  # 2^28
    assert cls(1 << 24, 1 << 24, 1 << 16, 1 << 14) == 6  # 2^32
    assert cls(10, 10, 0, 10) == 0


def test_shipped_or_default_table_is_valid():
    d = rvsp.tuning_defaults()
    assert d["big_window_outputs"] == 2048 and d["big_window_chunks"] == 16
    assert d["hash_max_u"] == 2048 and d["seg_products"] == 16384
    for c in range(rvsp.TUNING_CLASSES):
        e = rvsp.tuning_get(c)
        assert 64 <= e["hash_max_u"] <= 2048
        assert e["sym_big_threads"] in (256, 512, 1024)


def test_load_file_sets_listed_classes_and_defaults_the_rest(tmp_path):
    assert load(tmp_path, GOOD) == 0
    e5 = rvsp.tuning_get(5)
    assert (e5["big_window_outputs"], e5["big_window_chunks"], e5["sym_big_threads"],
            e5["seg_products"], e5["hash_max_u"], e5["fan_min_products"]) == (4096, 32, 256, 32768, 1024, 0)
    assert e5["fan_min_fill"] == 1 << 62
    assert rvsp.tuning_get(3) == rvsp.tuning_defaults()


@pytest.mark.parametrize("bad", [
    "5 3000 16 512 16384 2048 0 0\n",        # not a launchable window shape
    "5 2048 16 300 16384 2048 0 0\n",        # block size
    "5 2048 16 512 100 2048 0 0\n",          # segment too small
    "5 2048 16 512 16384 4096 0 0\n",        # hash limit above the largest table
    "5 2048 16 512 16384 2048 -1 0\n",       # negative threshold
    "8 2048 16 512 16384 2048 0 0\n",        # no such class
    "5 2048 16 512 16384 2048 0\n",          # a field missing
    "5 2048 16 512 16384 2048 0 0\n5 2048 16 512 16384 2048 0 0\n",  # class twice
])
def test_load_file_rejects_bad_tables_and_keeps_the_old_one(tmp_path, bad):
    before = [rvsp.tuning_get(c) for c in range(rvsp.TUNING_CLASSES)]
    assert load(tmp_path, bad) == rvsp.ERROR_INVALID_ARGUMENT
    assert [rvsp.tuning_get(c) for c in range(rvsp.TUNING_CLASSES)] == before


def test_set_and_force_check_values():
    L = rvsp.lib()
    t = rvsp.Tuning.of(dict(rvsp.tuning_defaults(), hash_max_u=5000))
    assert L.rvsp_tuning_set(1, t) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_tuning_force(t) == rvsp.ERROR_INVALID_ARGUMENT
    ok = rvsp.Tuning.of(dict(rvsp.tuning_defaults(), hash_max_u=512))
    assert L.rvsp_tuning_set(-1, ok) == rvsp.ERROR_INVALID_ARGUMENT
    assert L.rvsp_tuning_set(1, ok) == 0
    assert rvsp.tuning_get(1)["hash_max_u"] == 512
