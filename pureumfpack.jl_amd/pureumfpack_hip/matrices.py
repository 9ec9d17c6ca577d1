"""Test-matrix generators for the multifrontal path.

`poisson2d`, `poisson3d`, `randmat` and `arrowband` follow the generators in
bench/matrices.jl:5-86 (same grid numbering, same stencil values, same CSC
pattern).  `randmat` and `arrowband` draw from numpy's PCG64 instead of Julia's
MersenneTwister, so their pattern *shape* matches the reference but their
values and exact positions do not.

The remaining generators define the large workloads the GPU path is measured
on: the diagonally perturbed 3D Poisson matrix, the batch of unsymmetric
matrices sharing one 3D Poisson pattern, and the 3D first-order-upwind
convection-diffusion operator.

All matrices are returned as `scipy.sparse.csc_matrix` (0-based, sorted row
indices, float64), i.e. the same stored pattern Julia's `sparse(I,J,V,n,n)`
produces.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _csc(rows, cols, vals, n):
    A = sp.csc_matrix((np.asarray(vals, dtype=np.float64),
                       (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def _stencil(k: int, dims: int):
    """Row/col index arrays of the 2*dims+1 point Laplacian stencil on a k^dims grid.

    Grid numbering follows bench/matrices.jl:11 and :31 (first coordinate fastest).
    Returns (rows, cols, kind) where kind is 0 for the diagonal and 1 + 2*d + s
    for the neighbour in direction d (s=0: minus side, s=1: plus side).
    """
    n = k ** dims
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx.copy()
    for _ in range(dims):
        coords.append(rem % k)
        rem //= k
    rows = [idx]
    cols = [idx]
    kinds = [np.zeros(n, dtype=np.int64)]
    stride = 1
    for d in range(dims):
        c = coords[d]
        lo = c > 0
        rows.append(idx[lo]); cols.append(idx[lo] - stride)
        kinds.append(np.full(int(lo.sum()), 1 + 2 * d, dtype=np.int64))
        hi = c < k - 1
        rows.append(idx[hi]); cols.append(idx[hi] + stride)
        kinds.append(np.full(int(hi.sum()), 2 + 2 * d, dtype=np.int64))
        stride *= k
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(kinds)


def poisson2d(k: int) -> sp.csc_matrix:
    """5-point Laplacian on a k x k grid (bench/matrices.jl:5-22): diag 4, off-diag -1."""
    r, c, kind = _stencil(k, 2)
    v = np.where(kind == 0, 4.0, -1.0)
    return _csc(r, c, v, k * k)


def poisson3d(k: int) -> sp.csc_matrix:
    """7-point Laplacian on a k^3 grid (bench/matrices.jl:25-44): diag 6, off-diag -1."""
    r, c, kind = _stencil(k, 3)
    v = np.where(kind == 0, 6.0, -1.0)
    return _csc(r, c, v, k ** 3)


def randmat(n: int, deg: int, seed: int = 1) -> sp.csc_matrix:
    """Random unsymmetric matrix of density deg/n plus (deg+1) I (shape of bench/matrices.jl:47-51).

    Values of the random part are uniform [0,1) like Julia's `sprand`; the RNG
    is numpy's, so entries differ from the reference's MersenneTwister draw.
    """
    rng = np.random.default_rng(seed)
    R = sp.random(n, n, density=deg / n, format="csc", random_state=rng, dtype=np.float64)
    A = (R + sp.identity(n, format="csc") * (deg + 1.0)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def arrowband(n: int, bw: int, seed: int = 3) -> sp.csc_matrix:
    """Arrowhead + band unsymmetric matrix (shape of bench/matrices.jl:54-86)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for j in range(n):
        rows.append(j); cols.append(j); vals.append(float(n))
        for d in range(1, bw + 1):
            if j + d < n:
                rows += [j, j + d]; cols += [j + d, j]; vals += list(rng.standard_normal(2))
    for j in range(n):
        if rng.random() < 0.3:
            rows += [0, j]; cols += [j, 0]; vals += list(rng.standard_normal(2))
    return _csc(rows, cols, vals, n)


def poisson3d_perturbed(k: int, seed: int = 3, amp: float = 0.1) -> sp.csc_matrix:
    """poisson3d(k) + Diagonal(uniform[0, amp]) -- the mixed-precision workload."""
    A = poisson3d(k)
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.0, amp, A.shape[0])
    return (A + sp.diags(d, format="csc")).tocsc()


def poisson3d_batch_values(k: int, nmat: int, seed: int = 4):
    """Values of `nmat` unsymmetric matrices sharing the poisson3d(k) pattern.

    Diagonal 6 + U[0,1]; each off-diagonal entry -U[0.5,1.5], drawn independently,
    so the values are unsymmetric.  Returns (pattern, Ax) with `pattern` the
    csc_matrix of poisson3d(k) and `Ax` an (nmat, nnz) float64 array of values in
    the pattern's CSC order.
    """
    P = poisson3d(k)
    rng = np.random.default_rng(seed)
    nnz = P.nnz
    rows = P.indices
    cols = np.repeat(np.arange(P.shape[1]), np.diff(P.indptr))
    diag = rows == cols
    Ax = np.empty((nmat, nnz), dtype=np.float64)
    for m in range(nmat):
        v = -rng.uniform(0.5, 1.5, nnz)
        v[diag] = 6.0 + rng.uniform(0.0, 1.0, int(diag.sum()))
        Ax[m] = v
    return P, Ax


def convdiff3d(k: int, vel=(1.0, 0.5, 0.25), peclet: float = 5.0) -> sp.csc_matrix:
    """3D first-order-upwind convection-diffusion on a k^3 grid (h = 1, Dirichlet rows dropped).

    eps = max|v| / Pe.  Diagonal = 6 eps + sum_d v_d; the upwind neighbour in
    direction d is -eps - v_d, the downwind neighbour -eps.  With v_d > 0 the
    upwind neighbour is the minus side.
    """
    eps = max(abs(v) for v in vel) / peclet
    r, c, kind = _stencil(k, 3)
    v = np.empty(r.shape[0], dtype=np.float64)
    v[kind == 0] = 6.0 * eps + sum(vel)
    for d in range(3):
        up = -eps - vel[d] if vel[d] > 0 else -eps
        down = -eps if vel[d] > 0 else -eps + vel[d]
        v[kind == 1 + 2 * d] = up
        v[kind == 2 + 2 * d] = down
    return _csc(r, c, v, k ** 3)


def spd6(seed: int = 2024) -> sp.csc_matrix:
    """Dense SPD 6x6 M*M' + 6I stored sparse (test/multifrontal_test.jl:31-36 shape)."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((6, 6))
    return sp.csc_matrix(M @ M.T + 6.0 * np.eye(6))


# ---- matrices that make the in-block partial pivoting interchange rows ----

def sprand_identity(n: int, density: float | None = None, seed: int = 1) -> sp.csc_matrix:
    """sprand(n, n, 12/n) + I, the unsymmetric family of test/pureumfpack_test.jl:106: random
    uniform[0,1) entries at density 12/n plus the identity.  The diagonal is no larger than
    the off-diagonal entries, so the pivoting moves most rows."""
    rng = np.random.default_rng(seed)
    d = 12.0 / n if density is None else density
    R = sp.random(n, n, density=d, format="csc", random_state=rng, dtype=np.float64)
    A = (R + sp.identity(n, format="csc")).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


def dense_random(n: int, seed: int = 3) -> sp.csc_matrix:
    """Dense standard-normal n x n matrix stored sparse: one supernode, one front."""
    rng = np.random.default_rng(seed)
    return sp.csc_matrix(rng.standard_normal((n, n)))


def signs(n: int, density: float, seed: int) -> sp.csc_matrix:
    """Random +-1 matrix (diagonal always stored).  Under SCALE_NONE the candidate magnitudes of
    a pivot column are small integers, so exact ties occur at many steps and the first-maximum
    rule (BLAS idamax) decides the interchanges."""
    rng = np.random.default_rng(seed)
    D = np.where(rng.random((n, n)) < density, rng.choice([-1.0, 1.0], (n, n)), 0.0)
    np.fill_diagonal(D, rng.choice([-1.0, 1.0], n))
    rows, cols = np.nonzero(D)
    return _csc(rows, cols, D[rows, cols], n)


def poisson3d_weak(k: int, seed: int = 7, lo: float = 2.0, hi: float = 3.0) -> sp.csc_matrix:
    """The batched workload's unsymmetric values on the poisson3d(k) pattern (off-diagonals
    -U[0.5,1.5]) with the diagonal weakened to U[lo, hi]: not diagonally dominant, so the
    large fronts interchange rows."""
    P, Ax = poisson3d_batch_values(k, 1, seed=seed)
    A = P.copy()
    A.data = Ax[0].copy()
    rng = np.random.default_rng(seed + 100)
    A.setdiag(rng.uniform(lo, hi, A.shape[0]))
    A = A.tocsc()
    A.sort_indices()
    return A
