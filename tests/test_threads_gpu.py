"""The library called from several host threads at once.

Distinct descriptors are independent (each has its own stream and plan),
and one-shots and batches run on a stream of the calling thread, so
threads neither share mutable state nor queue behind each other. Four
threads run host-memory descriptors, device-memory descriptors with reuse
(per call and hipGraph-replayed), device one-shots and batches concurrently, several rounds each; every
result must equal the CPU oracle's bit for bit. (ctypes releases the GIL
during each library call, so the calls do overlap.)
"""
import ctypes as C
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import rvsp
from rvsp import workloads as W
from oracle import pyoracle
from helpers import assert_bitexact

pytestmark = pytest.mark.gpu

PRODUCTS = {
    "uniform": lambda: W.uniform(n=1024, per_row=16, seed=61),
    "rmat": lambda: W.rmat(scale=13, seed=62),
    "mixed": lambda: W.mixed(m=1 << 13, k=1 << 9, n=1 << 13, maxlen=256, seed_a=63, seed_b=64),
    "stencil": lambda: W.stencil(g=10, seed=65),
}


@pytest.fixture(scope="module")
def cases():
    rvsp.lib()  # loaded once, before the threads start
    out = {}
    for name, make in PRODUCTS.items():
        A, B = make()
        out[name] = (A, B, pyoracle.spgemm(A, B))
    return out


def host_descriptor(A, B):
    d = rvsp.Descr()
    a, b = A.struct(), B.struct()
    st, ws, nnz = d.estimate(a, b)
    assert st == 0
    rp = np.zeros(A.rows + 1, np.int32)
    ci = np.zeros(max(nnz, 1), np.int32)
    va = np.zeros(max(nnz, 1), np.float32)
    c = rvsp.CsrStruct(A.rows, B.cols, nnz, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 3, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    out = []
    for _ in range(2):  # the second compute reuses the structure
        assert d.compute(a, b, c, wsb.ctypes.data) == 0
        out.append(rvsp.Csr(A.rows, B.cols, rp.copy(), ci[:nnz].copy(), va[:nnz].copy()))
    d.close()
    return out


def device_descriptor(A, B, graph=False):
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    d = rvsp.Descr()
    assert d.set_memory(rvsp.MEMORY_DEVICE) == 0
    assert d.set_graph(graph) == 0
    a, b = dA.struct(), dB.struct()
    st, ws, nnz = d.estimate(a, b)
    assert st == 0
    crp = rvsp.DeviceBuffer(4 * (A.rows + 1))
    cci = rvsp.DeviceBuffer(4 * max(nnz, 1))
    cva = rvsp.DeviceBuffer(4 * max(nnz, 1))
    c = rvsp.CsrStruct(A.rows, B.cols, nnz, crp.ptr, cci.ptr, cva.ptr, 3, 0, 0)
    wsb = np.zeros(max(ws, 1), np.uint8)
    out = []
    for _ in range(4):  # fill, then reuse passes (captured, then replayed, with graph)
        assert d.compute(a, b, c, wsb.ctypes.data) == 0
        assert d.wait() == 0
        out.append(rvsp.Csr(A.rows, B.cols, crp.to_array(np.int32, A.rows + 1),
                            cci.to_array(np.int32, nnz), cva.to_array(np.float32, nnz)))
    d.close()
    return out


def oneshot(A, B):
    return [rvsp.spgemm_device(rvsp.DeviceCsr(A), rvsp.DeviceCsr(B))]


def batched(A, B):
    dA, dB = rvsp.DeviceCsr(A), rvsp.DeviceCsr(B)
    got, _ = rvsp.spgemm_batched([dA, dA], [dB, dB])
    return got


KINDS = {"host": host_descriptor, "device": device_descriptor,
         "graph": lambda A, B: device_descriptor(A, B, graph=True), "oneshot": oneshot,
         "batched": batched}


def test_concurrent_threads_distinct_descriptors(cases):
    jobs = [(kind, name) for rnd in range(3) for kind in KINDS for name in PRODUCTS]
    rng = np.random.default_rng(7)
    rng.shuffle(jobs)

    def run(job):
        kind, name = job
        A, B, _ = cases[name]
        try:
            return job, KINDS[kind](A, B)
        except AssertionError as e:
            raise AssertionError(f"job {kind} {name}: {e}") from None

    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(run, jobs))
    for (kind, name), outs in results:
        rp, ci, va = cases[name][2]
        for i, got in enumerate(outs):
            assert_bitexact(got, rp, ci, va, f"{kind} {name} result {i}")
    print(f"{len(jobs)} jobs on 4 threads, all bit-exact", flush=True)


def test_each_thread_has_its_own_oneshot_stream():
    L = rvsp.lib()
    L.rvsp_hip_oneshot_stream.restype = C.c_void_p
    L.rvsp_hip_oneshot_stream.argtypes = []
    seen = {}

    def grab(i):
        s1 = L.rvsp_hip_oneshot_stream()
        s2 = L.rvsp_hip_oneshot_stream()
        assert s1 and s1 == s2  # stable within a thread
        seen[i] = s1

    threads = [threading.Thread(target=grab, args=(i,)) for i in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert len(set(seen.values())) == 3
