"""Autotuner of the library's launch parameters (rvsp_tuning_t).

    python -m rvsp.autotune [--out PATH] [--quick]

Times seeded training products of every workload class on the GPU and
writes the table the analysis loads (default: tuning/table.txt beside the
library). For each class, a coordinate sweep from the built-in defaults:
each parameter in turn over its candidate values (the others held), the
class's total device one-shot time (median of 5 runs after 2 warm-ups per
product) decides, and a value replaces the current one only when it is at
least 2 % faster. Parameters never change results; the tuned entry of every
class is checked once more for bit-identical C against the defaults before
the table is written (tests/test_tuning_gpu.py checks every entry against
the oracle).

Training products use seeds distinct from the benchmark instances: R-MAT
(scales 14-18), uniform rows, rv-sparse's generator, the mixed power-law x
banded product (scaled) and 27-point stencils.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import os
import sys
import time

import numpy as np

import rvsp
from rvsp import workloads as W

CANDIDATES = {
    "hash_max_u": [256, 512, 1024, 2048],
    "big_window": [(2048, 16), (4096, 32), (2048, 32), (1024, 16), (1024, 8)],
    "sym_big_threads": [256, 512, 1024],
    "seg_products": [8192, 16384, 32768, 65536, 98304],
    "fan_min_products": [1 << 19, 1 << 21, 1 << 23, 1 << 62],
    "fan_min_fill": [1 << 22, 1 << 24, 1 << 26, 1 << 62],
}
# parameters that only matter where B has columns enough for big rows
WIDE_ONLY = ("big_window", "sym_big_threads", "seg_products")


def training_products(quick=False):
    """(name, A, B) of every training product (seeds differ from the bench's)."""
    out = [("uniform-1024", *W.uniform(n=1024, per_row=16, seed=71)),
           ("uniform-4096", *W.uniform(n=4096, per_row=16, seed=72)),
           ("gen-4096-d0.002", *rvsp.gen_bench_pair(4096, 4096, 0.002, 43)),
           ("rmat-14", *W.rmat(scale=14, seed=73)),
           ("stencil-32", *W.stencil(g=32, seed=78)),
           ("mixed-2^17", *W.mixed(m=1 << 17, k=1 << 13, n=1 << 17, seed_a=76, seed_b=77))]
    if not quick:
        out += [("gen-16384-d0.001", *rvsp.gen_bench_pair(16384, 16384, 0.001, 43)),
                ("rmat-16", *W.rmat(scale=16, seed=74)),
                ("rmat-18", *W.rmat(scale=18, seed=75)),
                ("stencil-64", *W.stencil(g=64, seed=79)),
                ("mixed-2^20", *W.mixed(seed_a=80, seed_b=81))]
    return out


class Case:
    """One training product on the device, timed as a device one-shot."""

    def __init__(self, name, A, B):
        self.name = name
        self.cls = rvsp.tuning_class(A, B)
        self.dA = rvsp.DeviceCsr(A)
        self.dB = self.dA if B is A else rvsp.DeviceCsr(B)

    def _once(self, keep=False):
        L = rvsp.lib()
        a, b = self.dA.struct(), self.dB.struct()
        c = rvsp.CsrStruct()
        rvsp.check("csr_device", L.rvsp_spgemm_csr_device_algo(C.byref(a), C.byref(b), C.byref(c),
                                                               rvsp.ALGO_HIP))
        if keep:
            return c
        L.rvsp_csr_destroy(C.byref(c))
        return None

    def time_ms(self, warmup=2, runs=5):
        L = rvsp.lib()
        for _ in range(warmup):
            self._once()
        ts = []
        for _ in range(runs):
            rvsp.check("sync", L.rvsp_hip_synchronize())
            t0 = time.perf_counter()
            self._once()
            rvsp.check("sync", L.rvsp_hip_synchronize())
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    def digest(self):
        """Hash of C (row pointer, columns, value bits)."""
        c = self._once(keep=True)
        h = rvsp.device_csr_to_host(c)
        rvsp.lib().rvsp_csr_destroy(C.byref(c))
        d = hashlib.blake2b(digest_size=16)
        for arr in (h.row_ptr, h.col_idx, h.values.view(np.uint32)):
            d.update(np.ascontiguousarray(arr).tobytes())
        return d.hexdigest()


def with_value(entry, key, value):
    e = dict(entry)
    if key == "big_window":
        e["big_window_outputs"], e["big_window_chunks"] = value
    else:
        e[key] = value
    return e


def current(entry, key):
    if key == "big_window":
        return (entry["big_window_outputs"], entry["big_window_chunks"])
    return entry[key]


def class_time(cases, entry):
    rvsp.tuning_force(entry)
    try:
        return sum(c.time_ms() for c in cases)
    finally:
        rvsp.tuning_force(None)


def tune_class(cls, cases, log, min_gain=0.02):
    entry = rvsp.tuning_defaults()
    base = best = class_time(cases, entry)
    log(f"class {cls}: {[c.name for c in cases]} defaults {base:.3f} ms")
    wide = cls % 2 == 1
    for key, values in CANDIDATES.items():
        if key in WIDE_ONLY and not wide:
            continue
        for v in values:
            if v == current(entry, key):
                continue
            trial = with_value(entry, key, v)
            t = class_time(cases, trial)
            log(f"  {key}={v}: {t:.3f} ms (best {best:.3f})")
            if t < best * (1.0 - min_gain):
                entry, best = trial, t
    return entry, base, best


def write_table(path, entries, notes):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("# rvsp launch-parameter table, written by: python -m rvsp.autotune\n")
        f.write("# class outputs chunks sym_big_threads seg_products hash_max_u "
                "fan_min_products fan_min_fill\n")
        for n in notes:
            f.write(f"# {n}\n")
        for cls in sorted(entries):
            e = entries[cls]
            f.write(f"{cls} {e['big_window_outputs']} {e['big_window_chunks']} "
                    f"{e['sym_big_threads']} {e['seg_products']} {e['hash_max_u']} "
                    f"{e['fan_min_products']} {e['fan_min_fill']}\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--out", default=os.path.join(here, "tuning", "table.txt"))
    ap.add_argument("--quick", action="store_true", help="small training products only")
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if not rvsp.gpu_available():
        raise SystemExit("autotune: no HIP device visible")
    for var in ("RVSP_BIG_WINDOW", "RVSP_SYM_BIG_THREADS"):
        if os.environ.pop(var, None) is not None:
            log(f"autotune: ignoring {var} (it would override every trial)")
    by_class = {}
    for name, A, B in training_products(args.quick):
        case = Case(name, A, B)
        by_class.setdefault(case.cls, []).append(case)
    entries, notes = {}, []
    for cls in sorted(by_class):
        cases = by_class[cls]
        entry, base, best = tune_class(cls, cases, log)
        # the tuned entry must not change any result
        rvsp.tuning_force(rvsp.tuning_defaults())
        want = [c.digest() for c in cases]
        rvsp.tuning_force(entry)
        got = [c.digest() for c in cases]
        rvsp.tuning_force(None)
        if got != want:
            raise SystemExit(f"autotune: class {cls} entry changed a result: {entry}")
        entries[cls] = entry
        notes.append(f"class {cls} ({', '.join(c.name for c in cases)}): "
                     f"{base:.3f} -> {best:.3f} ms ({base / best:.3f}x)")
        log(notes[-1])
    write_table(args.out, entries, notes)
    print(f"autotune: wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
