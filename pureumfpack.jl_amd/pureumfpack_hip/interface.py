"""Host-side mirror of PureUMFPACK.jl's multifrontal interface, running on libmfgpu.

Same names, argument meanings and error behaviour as the reference:
  row_scaling(A, kind)                    src/scaling.jl:92-115
  symbolic_mf(A, q=None)                  src/symbolic.jl:211-226 (C++ analysis in the library)
  multifrontal_lu(A, q=None, check=True)  src/multifrontal.jl:146-358  -> MFFactorization
  splu(A, method="multifrontal", ...)     src/interface.jl:93-120      -> PureLU
  solve(F, b, refine=0)                   src/interface.jl:170-177, src/solve.jl:66-77
  ldiv(F, b)  (Julia's F \\ b)            src/interface.jl:179, src/solve.jl:79
Indices exposed to Python are 0-based; the C ABI underneath is 1-based like Julia.

Every numeric step runs on the GPU through the C ABI; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import MF_BADARG, MF_DIM, MF_OK, MF_SINGULAR, c_f64p, c_i64p

SCALE_NONE, SCALE_SUM, SCALE_MAX = 0, 1, 2
_PREC = {"fp64": _lib.MF_FP64, "fp32": _lib.MF_FP32, "bf16": _lib.MF_BF16}
_PIVOT = {"auto": _lib.MF_PIVOT_AUTO, "partial": _lib.MF_PIVOT_PARTIAL}
_SOLVE = {"auto": _lib.MF_SOLVE_AUTO, "fronts": _lib.MF_SOLVE_FRONTS, "columns": _lib.MF_SOLVE_COLUMNS}


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class SingularException(ArithmeticError):
    """Julia's LinearAlgebra.SingularException; `info` is the 1-based factor column."""

    def __init__(self, info):
        super().__init__(f"SingularException({info})")
        self.info = info


class DeviceError(RuntimeError):
    pass


def _check(rc, what=""):
    if rc == MF_OK:
        return
    msg = _lib.load().mf_last_error().decode(errors="replace")
    if rc == MF_DIM:
        raise DimensionMismatch(msg or what)
    if rc == MF_BADARG:
        raise ArgumentError(msg or what)
    raise DeviceError(f"{what}: {msg} (code {rc})")


def _csc(A):
    A = sp.csc_matrix(A)
    if not A.has_sorted_indices:
        A = A.sorted_indices()
    return A


def _pattern_1based(A):
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int64) + 1
    Ai = np.ascontiguousarray(A.indices, dtype=np.int64) + 1
    return Ap, Ai


def row_scaling(A, kind=SCALE_SUM):
    """Row scale factors Rs (src/scaling.jl:92-115): 1/sum|a_i.| (SUM), 1/max|a_i.| (MAX) or ones;
    rows without a nonzero magnitude get 1.  Sums run in CSC storage order like the reference."""
    A = _csc(A)
    n = A.shape[0]
    Rs = np.ones(n)
    if kind == SCALE_NONE:
        return Rs
    acc = np.zeros(n)
    mag = np.abs(A.data)
    if kind == SCALE_SUM:
        np.add.at(acc, A.indices, mag)
    elif kind == SCALE_MAX:
        np.maximum.at(acc, A.indices, mag)
    else:
        raise ArgumentError(f"unknown scale kind {kind}")
    nz = acc > 0
    Rs[nz] = 1.0 / acc[nz]
    return Rs


class SymbolicMF:
    """Result of symbolic_mf (src/symbolic.jl:203-209), 0-based.  `colstruct` is kept only for
    each supernode's last column (`upd`), which is all the numeric phase reads."""

    def __init__(self, n, qf, parent, sstart, upd_ptr, upd_idx, child_ptr, child_idx, ncolstruct):
        self.n = n
        self.qf = qf
        self.parent = parent
        self.sstart = sstart
        self.upd_ptr = upd_ptr
        self.upd_idx = upd_idx
        self.child_ptr = child_ptr
        self.child_idx = child_idx
        self.ncolstruct = ncolstruct

    @property
    def nsuper(self):
        return len(self.sstart) - 1

    def upd(self, s):
        return self.upd_idx[self.upd_ptr[s]:self.upd_ptr[s + 1]]

    def children(self, s):
        return self.child_idx[self.child_ptr[s]:self.child_ptr[s + 1]]

    def predicted_fill(self):
        """nnz(L) + nnz(U) (src/symbolic.jl:230-237)."""
        return 2 * (self.n + self.ncolstruct)


RELAX_DEFAULT = (0.05, 0)   # (max explicit-zero fraction, max pivot block merged regardless)


def symbolic_mf(A, q=None, ordering="amd", relax=None):
    """Symbolic analysis (src/symbolic.jl:211-226) on the host, in the library's C++ code.

    `q` (0-based) fixes the order before postordering like `symbolic_mf(A; q)`; otherwise
    `ordering` is "amd" (amd_order_sym on A+A^T, the reference default) or "colamd" (AMD on
    A^T A, src/amd.jl:183-188).  `relax` (opt-in, no reference counterpart) applies relaxed
    supernode amalgamation: True for RELAX_DEFAULT or a (max_zero_fraction, max_small_np)
    pair; the factors then store explicit zeros and no longer have the reference's pattern."""
    A = _csc(A)
    n = A.shape[1]
    if A.shape[0] != n:
        raise DimensionMismatch("symbolic_mf requires a square matrix")
    lib = _lib.load()
    Ap, Ai = _pattern_1based(A)
    qq = None
    if q is not None:
        qq = np.ascontiguousarray(q, dtype=np.int64) + 1
        mode = _lib.MF_ORDER_GIVEN
    elif ordering == "amd":
        mode = _lib.MF_ORDER_AMD
    elif ordering == "colamd":
        mode = _lib.MF_ORDER_COLAMD
    else:
        raise ArgumentError(f"unknown ordering {ordering}")
    h = ctypes.c_void_p()
    _check(lib.mf_analyze(n, _lib.ptr(Ap), _lib.ptr(Ai), mode, _lib.ptr(qq), ctypes.byref(h)), "mf_analyze")
    try:
        if relax:
            zf, snp = RELAX_DEFAULT if relax is True else relax
            merged = ctypes.c_int64()
            _check(lib.mf_symbolic_relax(h, float(zf), int(snp), ctypes.byref(merged)), "mf_symbolic_relax")
        ns, nu, nc, ncs = (ctypes.c_int64() for _ in range(4))
        _check(lib.mf_symbolic_sizes(h, ctypes.byref(ns), ctypes.byref(nu), ctypes.byref(nc), ctypes.byref(ncs)))
        ns, nu, nc = ns.value, nu.value, nc.value
        qf = np.zeros(n, np.int64); parent = np.zeros(n, np.int64); sstart = np.zeros(ns + 1, np.int64)
        upd_ptr = np.zeros(ns + 1, np.int64); upd_idx = np.zeros(max(nu, 1), np.int64)
        child_ptr = np.zeros(ns + 1, np.int64); child_idx = np.zeros(max(nc, 1), np.int64)
        _check(lib.mf_symbolic_export(h, *(_lib.ptr(x) for x in (qf, parent, sstart, upd_ptr, upd_idx,
                                                                 child_ptr, child_idx))))
    finally:
        lib.mf_symbolic_destroy(h)
    return SymbolicMF(n, qf - 1, parent - 1, sstart - 1, upd_ptr - 1, upd_idx[:nu] - 1, child_ptr - 1,
                      child_idx[:nc] - 1, ncs.value)


class _Plan:
    """Owns an mf_plan (device schedule + front storage); destroyed with the Python object."""

    def __init__(self, A, S: SymbolicMF, precision="fp64", small_nf=0, use_graph=False, nbatch=1,
                 pivoting="auto", pivot_tol=0.0, solve_path="auto"):
        lib = _lib.load()
        if precision not in _PREC:
            raise ArgumentError(f"unknown precision {precision}")
        if pivoting not in _PIVOT:
            raise ArgumentError(f"unknown pivoting {pivoting}")
        if solve_path not in _SOLVE:
            raise ArgumentError(f"unknown solve_path {solve_path}")
        self.lib = lib
        self.n = A.shape[0]
        self.nnz = A.nnz
        self.S = S
        self.precision = precision
        opt = _lib.mf_options(_PREC[precision], int(small_nf), int(bool(use_graph)), int(nbatch),
                              _PIVOT[pivoting], float(pivot_tol), _SOLVE[solve_path])
        Ap, Ai = _pattern_1based(A)
        arrs = [np.ascontiguousarray(x, dtype=np.int64) + 1 for x in
                (S.qf, S.sstart, S.upd_ptr, S.upd_idx, S.child_ptr, S.child_idx)]
        h = ctypes.c_void_p()
        _check(lib.mf_plan_create(self.n, _lib.ptr(arrs[0]), S.nsuper, _lib.ptr(arrs[1]), _lib.ptr(arrs[2]),
                                  _lib.ptr(arrs[3]), _lib.ptr(arrs[4]), _lib.ptr(arrs[5]), _lib.ptr(Ap),
                                  _lib.ptr(Ai), ctypes.byref(opt), ctypes.byref(h)), "mf_plan_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.mf_plan_destroy(h)
            self.h = None

    def stats(self):
        st = _lib.mf_plan_stats()
        _check(self.lib.mf_plan_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def factor(self, Ax, Rs, check=True):
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        rs = None if Rs is None else np.ascontiguousarray(Rs, dtype=np.float64)
        info = ctypes.c_int64(0)
        rc = self.lib.mf_factor(self.h, _lib.ptr(Ax, c_f64p), _lib.ptr(rs, c_f64p), int(bool(check)),
                                ctypes.byref(info))
        if rc == MF_SINGULAR:
            raise SingularException(int(info.value))
        _check(rc, "mf_factor")
        return int(info.value)

    def solve(self, B, refine=0, tol=0.0):
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 1
        Bm = np.asfortranarray(B.reshape(self.n, -1))
        if Bm.shape[0] != self.n:
            raise DimensionMismatch("right-hand side has the wrong length")
        nrhs = Bm.shape[1]
        X = np.zeros((self.n, nrhs), order="F")
        rep = _lib.mf_solve_report()
        _check(self.lib.mf_solve(self.h, nrhs, _lib.ptr(Bm, c_f64p), self.n, _lib.ptr(X, c_f64p), self.n,
                                 int(refine), float(tol), ctypes.byref(rep)), "mf_solve")
        self.last_report = {f: getattr(rep, f) for f, _ in rep._fields_}
        return X[:, 0].copy() if vec else X

    def factor_batch(self, Ax, Rs, check=True):
        """mf_refactor_batch: Ax is nmat x nnz (values in the plan pattern's CSC order), Rs is
        nmat x n or None.  Returns the per-matrix singular columns (0 = nonsingular)."""
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        nmat = Ax.shape[0]
        if Ax.shape[1] != self.nnz:
            raise DimensionMismatch("batch values do not match the pattern")
        rs = None if Rs is None else np.ascontiguousarray(Rs, dtype=np.float64)
        info = np.zeros(nmat, np.int64)
        rc = self.lib.mf_refactor_batch(self.h, nmat, _lib.ptr(Ax, c_f64p), _lib.ptr(rs, c_f64p), _lib.ptr(info))
        if rc == MF_SINGULAR and check:
            raise SingularException(int(info[np.nonzero(info)[0][0]]))
        if rc != MF_SINGULAR:
            _check(rc, "mf_refactor_batch")
        self.nmat = nmat
        return info

    def solve_batch(self, B, refine=0, tol=0.0):
        """mf_solve_batch: B is nmat x n x nrhs (or nmat x n); returns X of the same shape."""
        B = np.asarray(B, dtype=np.float64)
        vec = B.ndim == 2
        B3 = B[:, :, None] if vec else B
        nmat, n, nrhs = B3.shape
        if n != self.n:
            raise DimensionMismatch("right-hand sides have the wrong length")
        Bt = np.ascontiguousarray(np.transpose(B3, (0, 2, 1)))   # per matrix: n x nrhs, column-major
        Xt = np.zeros_like(Bt)
        rep = _lib.mf_solve_report()
        _check(self.lib.mf_solve_batch(self.h, nmat, nrhs, _lib.ptr(Bt, c_f64p), _lib.ptr(Xt, c_f64p), int(refine),
                                       float(tol), ctypes.byref(rep)), "mf_solve_batch")
        self.last_report = {f: getattr(rep, f) for f, _ in rep._fields_}
        X = np.transpose(Xt, (0, 2, 1))
        return X[:, :, 0].copy() if vec else np.ascontiguousarray(X)

    def export(self, mat=None):
        lib = self.lib
        nL, nU = ctypes.c_int64(), ctypes.c_int64()
        _check(lib.mf_nnz(self.h, ctypes.byref(nL), ctypes.byref(nU)))
        n = self.n
        Lp = np.zeros(n + 1, np.int64); Li = np.zeros(nL.value, np.int64); Lx = np.zeros(nL.value)
        Up = np.zeros(n + 1, np.int64); Ui = np.zeros(nU.value, np.int64); Ux = np.zeros(nU.value)
        p = np.zeros(n, np.int64); q = np.zeros(n, np.int64)
        args = (_lib.ptr(Lp), _lib.ptr(Li), _lib.ptr(Lx, c_f64p), _lib.ptr(Up), _lib.ptr(Ui), _lib.ptr(Ux, c_f64p),
                _lib.ptr(p), _lib.ptr(q))
        if mat is None:
            _check(lib.mf_export_factors(self.h, *args), "export")
        else:
            _check(lib.mf_export_factors_batch(self.h, int(mat), *args), "export")
        L = sp.csc_matrix((Lx, Li - 1, Lp - 1), shape=(n, n))
        U = sp.csc_matrix((Ux, Ui - 1, Up - 1), shape=(n, n))
        return L, U, p - 1, q - 1


def plan_estimate(A, S: SymbolicMF, precision="fp64", small_nf=0, nbatch=1, pivoting="auto"):
    """Sizes and schedule statistics of the device plan for A's pattern (factor store, workspace
    arena, flops, launches), computed on the host by mf_plan_estimate without touching the GPU."""
    lib = _lib.load()
    if precision not in _PREC:
        raise ArgumentError(f"unknown precision {precision}")
    if pivoting not in _PIVOT:
        raise ArgumentError(f"unknown pivoting {pivoting}")
    A = _csc(A)
    opt = _lib.mf_options(_PREC[precision], int(small_nf), 0, int(nbatch), _PIVOT[pivoting], 0.0)
    Ap, Ai = _pattern_1based(A)
    arrs = [np.ascontiguousarray(x, dtype=np.int64) + 1 for x in
            (S.qf, S.sstart, S.upd_ptr, S.upd_idx, S.child_ptr, S.child_idx)]
    st = _lib.mf_plan_stats()
    _check(lib.mf_plan_estimate(A.shape[0], _lib.ptr(arrs[0]), S.nsuper, _lib.ptr(arrs[1]), _lib.ptr(arrs[2]),
                                _lib.ptr(arrs[3]), _lib.ptr(arrs[4]), _lib.ptr(arrs[5]), _lib.ptr(Ap), _lib.ptr(Ai),
                                ctypes.byref(opt), ctypes.byref(st)), "mf_plan_estimate")
    return {f: getattr(st, f) for f, _ in st._fields_}


class MFFactorization:
    """What multifrontal_lu returns (GPLUFactorization, src/gplu.jl:45-54): A[p, q] = L U.
    L, U, p and q are exported from the device on first access."""

    def __init__(self, plan: _Plan, Rs=None, A=None):
        self._plan = plan
        self.Rs = Rs
        self.A = A
        self._exp = None

    def _export(self):
        if self._exp is None:
            self._exp = self._plan.export()
        return self._exp

    L = property(lambda self: self._export()[0])
    U = property(lambda self: self._export()[1])
    p = property(lambda self: self._export()[2])
    q = property(lambda self: self._export()[3])

    @property
    def shape(self):
        return (self._plan.n, self._plan.n)

    def size(self, i=None):
        return self.shape if i is None else self.shape[i - 1]

    def solve(self, b, refine=0, tol=0.0):
        return self._plan.solve(b, refine=refine, tol=tol)

    @property
    def last_report(self):
        return getattr(self._plan, "last_report", None)


class PureLU(MFFactorization):
    """High-level factorization (src/interface.jl:38-48): (Rs .* A)[p, q] = L U; keeps Rs and A."""


def multifrontal_lu(A, q=None, tol=None, check=True, *, precision="fp64", symbolic=None, small_nf=0,
                    use_graph=False, pivoting="auto", pivot_tol=0.0):
    """Numeric multifrontal LU on the GPU (src/multifrontal.jl:146-358).  `tol` is accepted and
    ignored, as in the reference (:121).  Returns an MFFactorization with A[p, q] = L U."""
    A = _csc(A)
    n = A.shape[1]
    if A.shape[0] != n:
        raise DimensionMismatch("multifrontal_lu requires a square matrix")
    S = symbolic if symbolic is not None else symbolic_mf(A, q=q)
    plan = _Plan(A, S, precision=precision, small_nf=small_nf, use_graph=use_graph, pivoting=pivoting,
                 pivot_tol=pivot_tol)
    plan.factor(A.data, None, check=check)
    return MFFactorization(plan, None, A)


def splu(A, method="gplu", ordering="amd", tol=0.1, scale=SCALE_SUM, check=True, *, precision="fp64",
         symbolic=None, small_nf=0, use_graph=False, pivoting="auto", pivot_tol=0.0, solve_path="auto"):
    """High-level factorization (src/interface.jl:93-120).  Only method="multifrontal" is served
    by this package; method="gplu" (the reference default, a different algorithm) raises
    NotImplementedError, any other method raises ArgumentError.  `ordering` and `tol` are
    ignored for the multifrontal method, as in the reference (:103-105)."""
    A = _csc(A)
    n = A.shape[1]
    if A.shape[0] != n:
        raise DimensionMismatch("splu requires a square matrix")
    if method == "multifrontal":
        Rs = row_scaling(A, scale)
        S = symbolic if symbolic is not None else symbolic_mf(A)
        plan = _Plan(A, S, precision=precision, small_nf=small_nf, use_graph=use_graph, pivoting=pivoting,
                     pivot_tol=pivot_tol, solve_path=solve_path)
        plan.factor(A.data, None if scale == SCALE_NONE else Rs, check=check)
        return PureLU(plan, Rs, A)
    if method == "gplu":
        raise NotImplementedError("method='gplu' is not part of the GPU numeric phase")
    raise ArgumentError(f"unknown method {method}")


class BatchedLU:
    """Many matrices sharing one sparsity pattern: one symbolic analysis and one device plan;
    every launch of the factor and solve schedules covers the whole batch (grid dimension y
    = matrix), optionally replayed from a captured hipGraph.  The per-matrix semantics are
    those of splu(A_m; method=:multifrontal) and solve(F_m, b; refine) (src/interface.jl:93-177).

    values: nmat x nnz in the CSC order of `pattern` (the pattern's own data are ignored)."""

    def __init__(self, pattern, nbatch, precision="fp32", scale=SCALE_SUM, symbolic=None, use_graph=True,
                 pivoting="auto", pivot_tol=0.0, solve_path="auto"):
        A = _csc(pattern)
        if A.shape[0] != A.shape[1]:
            raise DimensionMismatch("BatchedLU requires square matrices")
        self.pattern = A
        self.scale = scale
        self.S = symbolic if symbolic is not None else symbolic_mf(A)
        self._plan = _Plan(A, self.S, precision=precision, use_graph=use_graph, nbatch=nbatch,
                           pivoting=pivoting, pivot_tol=pivot_tol, solve_path=solve_path)
        self.nbatch = int(nbatch)
        self.Rs = None

    def matrix(self, values_m):
        """Matrix m as a scipy CSC matrix with the shared pattern."""
        A = self.pattern
        return sp.csc_matrix((np.asarray(values_m, dtype=np.float64), A.indices, A.indptr), shape=A.shape)

    def refactor(self, values, check=True):
        """Factor every matrix of the batch (row scaling per matrix as splu computes it)."""
        values = np.ascontiguousarray(values, dtype=np.float64)
        if values.ndim != 2 or values.shape[1] != self.pattern.nnz:
            raise DimensionMismatch("values must be nmat x nnz(pattern)")
        if values.shape[0] > self.nbatch:
            raise ArgumentError("more matrices than the plan's nbatch")
        if self.scale == SCALE_NONE:
            self.Rs = None
        else:
            self.Rs = np.stack([row_scaling(self.matrix(v), self.scale) for v in values])
        self.values = values
        return self._plan.factor_batch(values, self.Rs, check=check)

    def solve(self, B, refine=0, tol=0.0):
        """X[m] = A_m^-1 B[m] for B of shape nmat x n (x nrhs)."""
        return self._plan.solve_batch(B, refine=refine, tol=tol)

    def factors(self, m):
        """(L, U, p, q) of matrix m: (Rs_m .* A_m)[p, q] = L U."""
        return self._plan.export(mat=m)

    @property
    def last_report(self):
        return getattr(self._plan, "last_report", None)


def solve(F: MFFactorization, b, refine: int = 0, tol: float = 0.0):
    """solve(F, b; refine) (src/interface.jl:170-177): `refine` fixed refinement steps; tol > 0
    stops early once the normwise backward error is <= tol.  b may be n or n x nrhs."""
    return F.solve(b, refine=refine, tol=tol)


def ldiv(F: MFFactorization, b):
    """Julia's F \\ b (src/interface.jl:179): solve with refine = 0."""
    return F.solve(b, refine=0)
