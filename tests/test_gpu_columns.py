"""Column-parallel solve sweeps (solve_path="columns": one lane per right-hand-side column of
the batch, vectors interleaved by column) against the per-front sweeps and the oracle.

The two sweep families read the same factors and apply the same operations (extend-add in
list order, pivot part permuted by the block pivoting, unit lower L, then U) with different
summation orders, so FP64 solutions agree to rounding; the oracle's solve (src/solve.jl:10-39
through oracle/numeric.py) pins both.  Cases cover row interchanges (per-matrix pivot orders
in one batch), pivot blocks and fronts wider than one 64-row register chunk, column counts
that do not fill a wave, and graph replay.
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import numeric as ON

from _util import oracle_splu

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _bwd(A, X, B):
    R = B - A @ X
    return (np.abs(R).max(axis=0) / (abs(A).sum(axis=1).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max()


@pytest.mark.parametrize("name,make", [("sprand-400", lambda: M.sprand_identity(400, seed=3)),
                                       ("poisson3d-12", lambda: M.poisson3d(12)),
                                       ("dense-200", lambda: M.dense_random(200, seed=1))])
@pytest.mark.parametrize("nrhs", [64, 37])
def test_columns_match_fronts_and_oracle_fp64(name, make, nrhs):
    A = make()
    S = I.symbolic_mf(A)
    npv = np.diff(np.asarray(S.sstart))
    nf = npv + np.diff(np.asarray(S.upd_ptr))
    print(f"[case] columns fp64 {name} nrhs={nrhs}: max np {npv.max()} max nf {nf.max()}", flush=True)
    B = np.random.default_rng(4).standard_normal((A.shape[0], nrhs))
    Fc = I.splu(A, method="multifrontal", symbolic=S, solve_path="columns")
    Ff = I.splu(A, method="multifrontal", symbolic=S, solve_path="fronts")
    Xc, Xf = I.solve(Fc, B), I.solve(Ff, B)
    assert _rel(Xc, Xf) <= 1e-12
    assert _bwd(A, Xc, B) <= 1e-14
    R = oracle_splu(A, S)
    for j in (0, nrhs - 1):
        assert _rel(Xc[:, j], ON.solve(R, B[:, j])) <= 1e-12
    # refinement through the columns path (FP64 residual, accumulated update)
    X1 = I.solve(Fc, B, refine=2)
    assert _bwd(A, X1, B) <= 1e-15 * 4


def _pivoting_batch(n, nmat, seed):
    """nmat matrices on the sprand(n, n, 12/n) + I pattern, each with its own values of that
    family (uniform[0, 1) plus 1 on the diagonal): each has its own row interchanges."""
    P = M.sprand_identity(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    rows = P.indices
    cols = np.repeat(np.arange(n), np.diff(P.indptr))
    Ax = rng.uniform(0.0, 1.0, (nmat, P.nnz))
    Ax[:, rows == cols] += 1.0
    return P, Ax


@pytest.mark.parametrize("nrhs,nmat", [(64, 6), (5, 40), (3, 5)])
def test_batch_columns_with_interchanges_fp64(nrhs, nmat):
    P, Ax = _pivoting_batch(300, nmat, seed=8)
    n = P.shape[0]
    print(f"[case] batch columns fp64 sprand-300 nrhs={nrhs} nmat={nmat}", flush=True)
    rhs = np.random.default_rng(2).standard_normal((nmat, n, nrhs))
    Bc = I.BatchedLU(P, nmat, precision="fp64", solve_path="columns", use_graph=False)
    Bf = I.BatchedLU(P, nmat, precision="fp64", solve_path="fronts", use_graph=False, symbolic=Bc.S)
    Bc.refactor(Ax)
    Bf.refactor(Ax)
    Xc, Xf = Bc.solve(rhs), Bf.solve(rhs)
    p0, q0 = Bc.factors(0)[2:]
    p1 = Bc.factors(1)[2]
    assert (p0 != q0).sum() > 0 and not np.array_equal(p0, p1)   # interchanges, per matrix
    for m in range(nmat):
        A = Bc.matrix(Ax[m])
        assert _bwd(A, Xc[m], rhs[m]) <= 1e-13
        assert _rel(Xc[m], Xf[m]) <= 1e-9
    R = oracle_splu(Bc.matrix(Ax[0]), Bc.S)
    assert _rel(Xc[0][:, 0], ON.solve(R, rhs[0][:, 0])) <= 1e-10


def test_batch_columns_fp32_refined_and_graph_replay():
    """FP32 fronts, FP64 refinement to 1e-14 through the columns path (chosen automatically at
    64 x 16 columns), and the captured sweeps replayed bit for bit."""
    nmat, nrhs = 16, 64
    P, Ax = M.poisson3d_batch_values(8, nmat, seed=6)
    n = P.shape[0]
    rhs = np.random.default_rng(3).standard_normal((nmat, n, nrhs))
    print(f"[case] batch columns fp32 poisson3d-8 nmat={nmat} nrhs={nrhs}", flush=True)
    Ba = I.BatchedLU(P, nmat, precision="fp32", use_graph=False)
    Bc = I.BatchedLU(P, nmat, precision="fp32", solve_path="columns", use_graph=True, symbolic=Ba.S)
    Ba.refactor(Ax)
    Bc.refactor(Ax)
    Xa = Ba.solve(rhs, refine=30, tol=1e-14)
    assert Ba.last_report["converged"] == 1
    X1 = Bc.solve(rhs, refine=30, tol=1e-14)
    X2 = Bc.solve(rhs, refine=30, tol=1e-14)
    assert np.array_equal(Xa, X1) and np.array_equal(X1, X2)
    for m in range(nmat):
        assert _bwd(Ba.matrix(Ax[m]), Xa[m], rhs[m]) <= 1e-14


def test_front_sweeps_stable_after_column_sweeps():
    """Regression: the backward triangle kernel of two-block pivot parts (np in (64, 128]) once
    read rows past np of its LDS vector, multiplied by zero factor entries; stale LDS from an
    earlier kernel (the FP64 tiles of the column-parallel transposes) made that 0 * NaN.  The
    per-front sweeps of one plan must give the same bits before and after another plan's
    column-parallel solves."""
    k, nmat, nrhs = 10, 6, 4     # top front: np = 100, two diagonal blocks
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=5)
    n = P.shape[0]
    rhs = np.random.default_rng(9).standard_normal((nmat, n, nrhs))
    B1 = I.BatchedLU(P, nmat, precision="fp32", use_graph=False, solve_path="fronts")
    B1.refactor(Ax)
    Xref = B1.solve(rhs)
    assert np.isfinite(Xref).all()
    A = M.poisson3d(12)
    F = I.splu(A, method="multifrontal", solve_path="columns")
    Bx = np.random.default_rng(4).standard_normal((A.shape[0], 64))
    for _ in range(4):
        I.solve(F, Bx)
        for _ in range(3):
            assert np.array_equal(B1.solve(rhs), Xref)
