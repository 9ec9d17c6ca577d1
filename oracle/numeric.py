"""Numeric multifrontal LU, solves and refinement -- oracle front end (test infrastructure only).

Wraps oracle/mf_oracle.c (built by oracle/Makefile into oracle/liboracle.so) and
adds the host glue of the reference path:

  splu_mf(A, scale, ...)      src/interface.jl:93-105 (method=:multifrontal branch)
  multifrontal_lu(As, S)      src/multifrontal.jl:146-358 given a SymbolicMF
  solve(F, b, refine)         src/interface.jl:123-177
  dense_restricted_lu(V, ss)  brute force: dense Gaussian elimination with the pivot
                              search confined to each supernode's rows, the textbook
                              routine the supernodal algorithm must agree with.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import scaling as _scaling
from . import symbolic as _symbolic

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_BLAS_SET = False
_BLAS_LIB = None

i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(os.path.join(_HERE, "mf_oracle.c")):
            build()
        L = ctypes.CDLL(path)
        L.mfo_factor.restype = ctypes.c_int
        L.mfo_factor.argtypes = [ctypes.c_int64, i64p, i64p, f64p, ctypes.c_int64, i64p, i64p, i64p,
                                 i64p, i64p, ctypes.c_int, ctypes.c_int, i64p, i64p, f64p, i64p, i64p,
                                 f64p, i64p, ctypes.POINTER(ctypes.c_int64)]
        L.mfo_colptrs.restype = None
        L.mfo_colptrs.argtypes = [ctypes.c_int64, ctypes.c_int64, i64p, i64p, i64p, i64p, i64p]
        L.mfo_solve_factor.restype = None
        L.mfo_solve_factor.argtypes = [ctypes.c_int64, i64p, i64p, f64p, i64p, i64p, f64p, i64p, i64p,
                                       ctypes.c_void_p, f64p, f64p, f64p]
        L.mfo_residual.restype = None
        L.mfo_residual.argtypes = [ctypes.c_int64, i64p, i64p, f64p, f64p, f64p, f64p, f64p]
        L.mfo_set_blas.restype = None
        L.mfo_set_blas.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _LIB = L
    return _LIB


def openblas_path():
    import numpy
    d = os.path.join(os.path.dirname(numpy.__file__), "..", "numpy.libs")
    c = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
    return os.path.abspath(c[0]) if c else None


def use_openblas(threads: int | None = None) -> bool:
    """Point backend 1 at numpy's bundled OpenBLAS (ILP64); optionally set its thread count."""
    global _BLAS_SET, _BLAS_LIB
    path = openblas_path()
    if path is None:
        return False
    if _BLAS_LIB is None:
        _BLAS_LIB = ctypes.CDLL(path)
    B = _BLAS_LIB
    addr = lambda name: ctypes.cast(getattr(B, name), ctypes.c_void_p).value
    lib().mfo_set_blas(addr("scipy_dgetrf_64_"), addr("scipy_dtrsm_64_"), addr("scipy_dgemm_64_"))
    if threads is not None:
        f = B.scipy_openblas_set_num_threads64_      # C API: void (int)
        f.argtypes = [ctypes.c_int]
        f.restype = None
        f(int(threads))
    _BLAS_SET = True
    return True


def flatten_symbolic(S):
    """(sstart, upd_ptr, upd_idx, child_ptr, child_idx) 0-based int64 arrays of a SymbolicMF."""
    ns = S.nsuper
    sstart = np.asarray(S.sstart, dtype=np.int64)
    upd = [S.colstruct[S.sstart[s + 1] - 1] for s in range(ns)]
    upd_ptr = np.zeros(ns + 1, dtype=np.int64)
    upd_ptr[1:] = np.cumsum([len(u) for u in upd])
    upd_idx = np.fromiter((i for u in upd for i in u), dtype=np.int64, count=int(upd_ptr[-1]))
    child_ptr = np.zeros(ns + 1, dtype=np.int64)
    child_ptr[1:] = np.cumsum([len(c) for c in S.childsupers])
    child_idx = np.fromiter((i for c in S.childsupers for i in c), dtype=np.int64, count=int(child_ptr[-1]))
    return sstart, upd_ptr, upd_idx, child_ptr, child_idx


@dataclass
class MFFactor:
    L: sp.csc_matrix
    U: sp.csc_matrix
    p: np.ndarray       # row permutation: As[p, q] = L U (0-based)
    q: np.ndarray
    prow: np.ndarray    # factor row -> V row
    info_col: int       # first exactly-singular pivot column (0-based) or -1
    Rs: np.ndarray | None = None
    A: sp.csc_matrix | None = None


class SingularException(Exception):
    def __init__(self, col):
        super().__init__(f"singular pivot at column {col}")
        self.col = col


def multifrontal_lu(As, S=None, flat=None, backend: int = 0, check: bool = True) -> MFFactor:
    """src/multifrontal.jl:146-358 on the oracle's C loop; `S` defaults to symbolic_mf(As)."""
    As = As.tocsc()
    n = As.shape[0]
    if As.shape[1] != n:
        raise ValueError("multifrontal_lu requires a square matrix")
    if S is None:
        S = _symbolic.symbolic_mf(As)
    qf = np.asarray(S.qf, dtype=np.int64)
    sstart, upd_ptr, upd_idx, child_ptr, child_idx = flat if flat is not None else flatten_symbolic(S)
    V = As[qf, :][:, qf].tocsc()
    V.sort_indices()
    Vp = V.indptr.astype(np.int64)
    Vi = V.indices.astype(np.int64)
    Vx = np.ascontiguousarray(V.data, dtype=np.float64)
    ns = len(sstart) - 1
    Lp = np.zeros(n + 1, dtype=np.int64)
    Up = np.zeros(n + 1, dtype=np.int64)
    lib().mfo_colptrs(n, ns, sstart, upd_ptr, upd_idx, Lp, Up)
    nL, nU = int(Lp[-1]), int(Up[-1])
    Li = np.zeros(max(nL, 1), dtype=np.int64); Lx = np.zeros(max(nL, 1))
    Ui = np.zeros(max(nU, 1), dtype=np.int64); Ux = np.zeros(max(nU, 1))
    prow = np.zeros(max(n, 1), dtype=np.int64)
    info = ctypes.c_int64(-1)
    if backend == 1 and not _BLAS_SET:
        if not use_openblas():
            raise RuntimeError("OpenBLAS backend unavailable")
    st = lib().mfo_factor(n, Vp, Vi, Vx, ns, sstart, upd_ptr, upd_idx, child_ptr, child_idx,
                          int(backend), int(check), Lp, Li, Lx, Up, Ui, Ux, prow, ctypes.byref(info))
    if st != 0:
        raise SingularException(int(info.value))
    prow = prow[:n]
    L = sp.csc_matrix((Lx[:nL], Li[:nL], Lp), shape=(n, n))
    U = sp.csc_matrix((Ux[:nU], Ui[:nU], Up), shape=(n, n))
    return MFFactor(L, U, qf[prow], qf, prow, int(info.value))


def splu_mf(A, scale: int = _scaling.SCALE_SUM, check: bool = True, backend: int = 0, S=None, flat=None):
    """src/interface.jl:93-105: scale rows, then the multifrontal factor; keeps Rs and A."""
    A = A.tocsc()
    n = A.shape[1]
    if A.shape[0] != n:
        raise ValueError("splu requires a square matrix")
    Rs = _scaling.row_scaling(A, scale)
    As = A if scale == _scaling.SCALE_NONE else _scaling.apply_row_scaling(A, Rs)
    F = multifrontal_lu(As, S=S, flat=flat, backend=backend, check=check)
    F.Rs = Rs
    F.A = A
    return F


def solve_factor(F: MFFactor, b, Rs=None):
    """src/interface.jl:123-138 (Rs given) / src/solve.jl:66-77 (Rs None)."""
    n = F.L.shape[0]
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    w = np.zeros(n)
    L, U = F.L, F.U
    rs = None if Rs is None else np.ascontiguousarray(Rs, dtype=np.float64)
    lib().mfo_solve_factor(n, L.indptr.astype(np.int64), L.indices.astype(np.int64), L.data,
                           U.indptr.astype(np.int64), U.indices.astype(np.int64), U.data,
                           F.p.astype(np.int64), F.q.astype(np.int64),
                           None if rs is None else rs.ctypes.data, b, x, w)
    return x


def residual(A, x, b):
    """r = b - A x with the column-oriented SpMV of SparseArrays (src/interface.jl:173)."""
    A = A.tocsc()
    n = A.shape[0]
    r = np.zeros(n)
    w = np.zeros(n)
    lib().mfo_residual(n, A.indptr.astype(np.int64), A.indices.astype(np.int64),
                       np.ascontiguousarray(A.data, dtype=np.float64),
                       np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64), r, w)
    return r


def solve(F: MFFactor, b, refine: int = 0):
    """src/interface.jl:170-177: solve, then `refine` fixed refinement steps."""
    x = solve_factor(F, b, F.Rs)
    for _ in range(refine):
        r = residual(F.A, x, b)
        x = x + solve_factor(F, r, F.Rs)
    return x


def backward_error(A, x, b):
    """Normwise (Rigal-Gaches) backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    r = b - A @ x
    Ainf = abs(A).sum(axis=1).max()
    return np.abs(r).max() / (Ainf * np.abs(x).max() + np.abs(b).max())


def dense_restricted_lu(V, sstart, tie="first", stats=None):
    """Dense elimination of V with pivots chosen, column by column, as the first row of
    maximum magnitude among the not-yet-pivoted rows of the column's own supernode.

    Returns (prow, L, U) with V[prow, :] = L U (dense).  Brute-force check of mfo_factor.
    tie="last" takes the last row of maximum magnitude instead (a deliberately wrong rule the
    tests use to show that ties decide interchanges); `stats` (a dict) receives the number of
    steps whose maximum magnitude is held by more than one candidate row ("ties") and the
    number of steps that interchange rows ("swaps").
    """
    W = np.array(V.toarray() if sp.issparse(V) else V, dtype=np.float64)
    n = W.shape[0]
    perm = np.arange(n)
    ties = swaps = 0
    for s in range(len(sstart) - 1):
        c1, c2 = sstart[s], sstart[s + 1] - 1
        for j in range(c1, c2 + 1):
            seg = np.abs(W[j:c2 + 1, j])
            at_max = np.flatnonzero(seg == seg.max())
            ties += int(at_max.size > 1)
            p = j + int(at_max[0] if tie == "first" else at_max[-1])
            swaps += int(p != j)
            if p != j:
                W[[j, p], :] = W[[p, j], :]
                perm[[j, p]] = perm[[p, j]]
            if W[j, j] != 0.0:
                W[j + 1:, j] = W[j + 1:, j] * (1.0 / W[j, j])
            W[j + 1:, j + 1:] -= np.outer(W[j + 1:, j], W[j, j + 1:])
    L = np.tril(W, -1) + np.eye(n)
    U = np.triu(W)
    if stats is not None:
        stats["ties"] = ties
        stats["swaps"] = swaps
    return perm, L, U
