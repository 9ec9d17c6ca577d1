"""Shared helpers for the parity tests: run the GPU path and the oracle on the same input."""
import numpy as np
import scipy.sparse as sp

from oracle import numeric as ON
from oracle import scaling as OS


def flat_from_gpu_symbolic(S):
    return (np.asarray(S.sstart, np.int64), np.asarray(S.upd_ptr, np.int64), np.asarray(S.upd_idx, np.int64),
            np.asarray(S.child_ptr, np.int64), np.asarray(S.child_idx, np.int64))


class _SymView:
    """Adapter giving the oracle the qf of a library SymbolicMF."""

    def __init__(self, S):
        self.qf = np.asarray(S.qf, np.int64)


def oracle_splu(A, S, scale=OS.SCALE_SUM, backend=0):
    """The oracle's splu on the same symbolic arrays the GPU plan was built from."""
    return ON.splu_mf(A, scale=scale, S=_SymView(S), flat=flat_from_gpu_symbolic(S), backend=backend)


def max_rel(a, b):
    a = a.toarray() if sp.issparse(a) else np.asarray(a)
    b = b.toarray() if sp.issparse(b) else np.asarray(b)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def same_pattern(X, Y):
    return (np.array_equal(X.indptr, Y.indptr) and np.array_equal(X.indices, Y.indices))


class DeviceArrays:
    """Device copies of numpy arrays through the HIP runtime (for the *_device entry points)."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.ct.c_void_p()
        assert self.h.hipMalloc(self.ct.byref(p), max(a.nbytes, 16)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        self.ptrs.append(p)
        return p

    def download(self, p, like):
        out = np.empty_like(np.ascontiguousarray(like))
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []
