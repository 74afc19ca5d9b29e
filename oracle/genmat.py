"""Binding of rv-sparse's own synthetic generator (oracle/_ref/
librvsp_genmat_ref.so, built from tools/src/genmat.c by build_ref.sh), the
checker of genmat_generate_csr.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.

The reference draws its values from the process-wide rand() stream
(srand(seed) at genmat.c:873); calls here are serialised by the caller.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF_GENMAT_PATH = os.path.join(HERE, "_ref", "librvsp_genmat_ref.so")
_lib = None


def ref_available() -> bool:
    return os.path.exists(REF_GENMAT_PATH)


def _ref():
    global _lib
    if _lib is None:
        import rvsp
        lib = C.CDLL(REF_GENMAT_PATH)
        lib.genmat_generate_csr.restype = rvsp.GenmatCsr
        lib.genmat_generate_csr.argtypes = [rvsp.GenmatParams]
        lib.genmat_default_params.restype = rvsp.GenmatParams
        lib.genmat_default_params.argtypes = [C.c_int, C.c_int]
        lib.genmat_free_csr.restype = None
        lib.genmat_free_csr.argtypes = [C.POINTER(rvsp.GenmatCsr)]
        _lib = lib
    return _lib


def ref_default_params(rows, cols):
    return _ref().genmat_default_params(rows, cols)


def ref_generate(params):
    """(rows, cols, row_ptr, col_idx, values float64) from rv-sparse's own
    generator, or None on its error matrix."""
    import rvsp
    lib = _ref()
    m = lib.genmat_generate_csr(params)
    if m.nrows < 0:
        return None
    out = rvsp.genmat_arrays(m)
    lib.genmat_free_csr(C.byref(m))
    return out
