"""ctypes wrapper of the oracle's C dense matvec (test infrastructure only)."""
import ctypes
import os

import numpy as np

from .build_oracle import OUT, build

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(OUT):
            build(verbose=False)
        L = ctypes.CDLL(OUT)
        vp = ctypes.c_void_p
        L.oracle_dense_rows.argtypes = [vp, vp, vp, vp, vp, ctypes.c_long, ctypes.c_long,
                                        ctypes.c_long, ctypes.c_int]
        L.oracle_dense_rows.restype = ctypes.c_int
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def dense_rows(C, A, sigma, self_integrals=None, row0=0, row1=None, nthreads=0):
    """Rows [row0, row1) of the FP64 dense matvec (core/bem_kernel.py:157-230)."""
    C = np.ascontiguousarray(C, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    si = None if self_integrals is None else np.ascontiguousarray(self_integrals, dtype=np.float64)
    n = C.shape[0]
    row1 = n if row1 is None else row1
    out = np.empty(row1 - row0)
    vp = ctypes.c_void_p
    lib().oracle_dense_rows(C.ctypes.data_as(vp), A.ctypes.data_as(vp), s.ctypes.data_as(vp),
                            None if si is None else si.ctypes.data_as(vp),
                            out.ctypes.data_as(vp), n, row0, row1, int(nthreads))
    return out


def max_threads():
    return int(lib().oracle_max_threads())
