"""Pins the CPU oracle against what the reference publishes and what the mathematics fixes.

Published facts (no Julia here, so the reference itself cannot run):
  * fill of AMD + symbolic on the reference's matrices (results/SUMMARY.md:32-39,52,56) and
    nsuper / max front (results/SCALING_DIAGNOSIS.md:24-30);
  * the row_scaling doctest (src/scaling.jl:86-89) and the solve doctests
    (src/interface.jl:85-90,161-167; src/multifrontal.jl:136-143).
Independent checks:
  * brute force: dense elimination with per-supernode restricted pivoting gives the oracle's
    prow, L and U; dense symbolic elimination gives predicted_fill;
  * the loop backend and the OpenBLAS backend (the reference's own getrf/trsm/gemm calls)
    agree;
  * planted mistakes (a transposed operand, a dropped extend-add, a wrong tie-break) are caught.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pureumfpack_hip import matrices as M
from oracle import amd as OA
from oracle import numeric as ON
from oracle import scaling as OS
from oracle import symbolic as OY


def test_row_scaling_doctest():
    A = sp.csc_matrix(np.array([[1.0, 2.0], [0.0, 4.0]]))
    assert np.array_equal(OS.row_scaling(A, OS.SCALE_SUM), [0.3333333333333333, 0.25])
    assert np.array_equal(OS.row_scaling(A, OS.SCALE_MAX), [0.5, 0.25])
    assert np.array_equal(OS.row_scaling(A, OS.SCALE_NONE), [1.0, 1.0])
    Z = sp.csc_matrix(np.array([[0.0, 0.0], [0.0, 4.0]]))
    assert OS.row_scaling(Z, OS.SCALE_SUM)[0] == 1.0


@pytest.mark.parametrize("name,A,fill", [
    ("poisson2d-32", lambda: M.poisson2d(32), 23800),
    ("poisson2d-50", lambda: M.poisson2d(50), 71826),
    ("poisson2d-64", lambda: M.poisson2d(64), 134400),
    ("poisson2d-100", lambda: M.poisson2d(100), 412664),
    ("poisson3d-16", lambda: M.poisson3d(16), 562028),
])
def test_amd_published_fill(name, A, fill):
    S = OY.symbolic_mf(A())
    assert OY.predicted_fill(S) == fill


@pytest.mark.parametrize("k,nsuper,maxnf", [(12, 1199, 206), (16, 2792, 427)])
def test_published_supernodes(k, nsuper, maxnf):
    S = OY.symbolic_mf(M.poisson3d(k))
    nf = max((S.sstart[s + 1] - S.sstart[s]) + len(S.colstruct[S.sstart[s + 1] - 1]) for s in range(S.nsuper))
    assert S.nsuper == nsuper and nf == maxnf


@pytest.mark.slow
def test_published_poisson3d_24():
    S = OY.symbolic_mf(M.poisson3d(24))
    assert OY.predicted_fill(S) == 3749118 and S.nsuper == 9348


def test_amd_is_permutation_and_reduces_fill():
    for A in (M.randmat(300, 6), M.arrowband(300, 4), M.poisson3d(6)):
        n = A.shape[0]
        for q in (OA.amd_order_sym(A), OA.colamd_order(A)):
            assert np.array_equal(np.sort(q), np.arange(n))
        Snat = OY.symbolic_mf(A, q=np.arange(n))
        assert OY.predicted_fill(OY.symbolic_mf(A)) < OY.predicted_fill(Snat)


@pytest.mark.parametrize("A", [M.poisson2d(6), M.poisson3d(4), M.randmat(60, 4), M.arrowband(50, 3)])
def test_symbolic_invariants_brute_force(A):
    S = OY.symbolic_mf(A)
    n = A.shape[0]
    assert np.array_equal(np.sort(S.qf), np.arange(n))
    assert all(p == -1 or p > j for j, p in enumerate(S.parent))
    B = A[S.qf, :][:, S.qf]
    assert OY.predicted_fill(S) == OY.dense_symbolic_fill(B)
    # supernode columns nest: struct(j) = {j+1..c2} U struct(c2)
    for s in range(S.nsuper):
        c1, c2 = S.sstart[s], S.sstart[s + 1] - 1
        for j in range(c1, c2):
            assert S.colstruct[j] == list(range(j + 1, c2 + 1)) + S.colstruct[c2]


MATS = [M.spd6(), M.poisson2d(8), M.poisson3d(6), M.randmat(200, 6), M.arrowband(100, 4), M.convdiff3d(5)]


@pytest.mark.parametrize("A", MATS)
def test_numeric_matches_dense_restricted_elimination(A):
    F = ON.splu_mf(A)
    As = OS.apply_row_scaling(A, F.Rs)
    S = OY.symbolic_mf(As)
    V = As[S.qf, :][:, S.qf]
    perm, Ld, Ud = ON.dense_restricted_lu(V, S.sstart)
    assert np.array_equal(perm, F.prow)
    scale = np.abs(V.toarray()).max()
    assert np.abs(F.L.toarray() - Ld).max() <= 1e-13
    assert np.abs(F.U.toarray() - Ud).max() <= 1e-13 * scale
    rec = abs(As[F.p, :][:, F.q] - F.L @ F.U).max() / abs(As).max()
    assert rec <= 1e-14


@pytest.mark.parametrize("A", MATS)
def test_blas_backend_matches_loops(A):
    F0 = ON.splu_mf(A, backend=0)
    F1 = ON.splu_mf(A, backend=1)
    assert np.array_equal(F0.p, F1.p)
    assert np.array_equal(F0.L.indices, F1.L.indices) and np.array_equal(F0.U.indices, F1.U.indices)
    assert np.abs(F0.L.data - F1.L.data).max() <= 1e-13 * np.abs(F0.L.data).max()
    assert np.abs(F0.U.data - F1.U.data).max() <= 1e-13 * np.abs(F0.U.data).max()


def test_factor_structure_invariants():
    """istril(L), unit diag(L) first in each column, istriu(U) with the diagonal last, and
    nnz(L) + nnz(U) == predicted_fill (test/multifrontal_test.jl:23-51)."""
    A = M.randmat(300, 6)
    F = ON.splu_mf(A)
    L, U = F.L, F.U
    assert sp.triu(L, 1).nnz == 0 and sp.tril(U, -1).nnz == 0
    n = A.shape[0]
    for j in range(n):
        assert L.indices[L.indptr[j]] == j and L.data[L.indptr[j]] == 1.0
        assert U.indices[U.indptr[j + 1] - 1] == j
        assert np.all(np.diff(U.indices[U.indptr[j]:U.indptr[j + 1]]) > 0)
    assert L.nnz + U.nnz == OY.predicted_fill(OY.symbolic_mf(OS.apply_row_scaling(A, F.Rs)))


def test_unpivoted_case_is_doolittle():
    """A diagonally dominant tridiagonal matrix: every supernode pivots on its diagonal, so the
    factor equals the textbook unpivoted LU of V."""
    n = 40
    T = sp.diags([-1.0, 4.0, -1.5], [-1, 0, 1], shape=(n, n), format="csc")
    F = ON.splu_mf(T, scale=OS.SCALE_NONE)
    V = T[F.q, :][:, F.q].toarray()
    W = V.copy()
    for k in range(n):
        W[k + 1:, k] /= W[k, k]
        W[k + 1:, k + 1:] -= np.outer(W[k + 1:, k], W[k, k + 1:])
    assert np.array_equal(F.p, F.q)
    assert np.abs(F.L.toarray() - (np.tril(W, -1) + np.eye(n))).max() <= 1e-15
    assert np.abs(F.U.toarray() - np.triu(W)).max() <= 1e-14


def test_solve_doctests_and_refine():
    A = sp.csc_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    want = [0.6666666666666666, -0.3333333333333333]
    F = ON.splu_mf(A)
    assert np.allclose(ON.solve(F, np.array([1.0, 0.0])), want, rtol=0, atol=1e-16)
    assert np.allclose(ON.solve(F, np.array([1.0, 0.0]), refine=1), want, rtol=0, atol=1e-16)
    R = ON.multifrontal_lu(A)
    assert np.allclose(ON.solve_factor(R, np.array([1.0, 0.0])), want, rtol=0, atol=1e-16)


def test_solve_against_dense():
    rng = np.random.default_rng(11)
    for A in (M.poisson2d(12), M.randmat(150, 5)):
        b = rng.standard_normal(A.shape[0])
        x = ON.solve(ON.splu_mf(A), b, refine=1)
        xd = np.linalg.solve(A.toarray(), b)
        assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max()
        assert ON.backward_error(A, x, b) <= 1e-15


def test_singular_check():
    D = sp.csc_matrix(np.diag([2.0, 0.0, 5.0]))
    with pytest.raises(ON.SingularException):
        ON.splu_mf(D, scale=OS.SCALE_NONE)
    F = ON.splu_mf(D, scale=OS.SCALE_NONE, check=False)
    assert F.info_col >= 0


# ---- planted mistakes: each must be detected by the checks above ----

def test_planted_transposed_operand_is_caught():
    A = M.randmat(80, 5)
    F = ON.splu_mf(A)
    As = OS.apply_row_scaling(A, F.Rs)
    wrong = F.U.T @ F.L.T                       # transposed operands
    assert abs(As[F.p, :][:, F.q] - wrong).max() > 1e-3


def test_planted_dropped_update_is_caught():
    A = M.poisson2d(8)
    F = ON.splu_mf(A)
    As = OS.apply_row_scaling(A, F.Rs)
    S = OY.symbolic_mf(As)
    V = As[S.qf, :][:, S.qf].toarray()
    # eliminate without the Schur update of the last pivot column: the dense check must disagree
    W = V.copy()
    n = W.shape[0]
    for k in range(n - 2):
        W[k + 1:, k] /= W[k, k]
        if k != n // 2:
            W[k + 1:, k + 1:] -= np.outer(W[k + 1:, k], W[k, k + 1:])
    assert np.abs(np.triu(W) - F.U.toarray()).max() > 1e-6


def test_planted_tiebreak_is_caught():
    """A tie in pivot magnitude must resolve to the first row (idamax); the last-row rule gives a
    different prow, which the brute-force comparison detects."""
    V = np.array([[1.0, 2.0, 0.0], [-1.0, 3.0, 1.0], [1.0, 0.0, 4.0]])
    perm, _, _ = ON.dense_restricted_lu(V, [0, 3])
    assert perm[0] == 0                     # first of the tied rows 0, 1, 2


# ---- pivoting: matrices whose in-block partial pivoting interchanges rows ----
# (the reference's sprand(400, 400, 12/400) + I family of test/pureumfpack_test.jl:106, dense
# random fronts, +-1 matrices with exact magnitude ties, and the batched workload's unsymmetric
# values with a weakened diagonal)

PIVOT_MATS = {
    "sprand-400": (lambda: M.sprand_identity(400, seed=1), OS.SCALE_SUM),
    "sprand-1000": (lambda: M.sprand_identity(1000, seed=2), OS.SCALE_SUM),
    "dense-40": (lambda: M.dense_random(40, seed=5), OS.SCALE_SUM),
    "dense-100": (lambda: M.dense_random(100, seed=3), OS.SCALE_SUM),
    "signs-60": (lambda: M.signs(60, 0.3, 1), OS.SCALE_NONE),
    "signs-150": (lambda: M.signs(150, 0.5, 3), OS.SCALE_NONE),
    "weak-8": (lambda: M.poisson3d_weak(8), OS.SCALE_SUM),
}


def _restricted(A, scale, **kw):
    F = ON.splu_mf(A, scale=scale)
    As = A if scale == OS.SCALE_NONE else OS.apply_row_scaling(A, F.Rs)
    S = OY.symbolic_mf(As)
    V = As[S.qf, :][:, S.qf]
    return F, As, V, S, ON.dense_restricted_lu(V, S.sstart, **kw)


@pytest.mark.parametrize("name", list(PIVOT_MATS))
def test_pivoting_matches_dense_restricted_elimination(name):
    """The oracle's prow equals the brute-force restricted elimination's on matrices where the
    pivoting moves rows (src/multifrontal.jl:247-275,363-375), and its L (relabelled through
    invperm(prow), :347-350) and U equal the dense factors."""
    A, scale = PIVOT_MATS[name]
    A = A()
    stats = {}
    F, As, V, S, (perm, Ld, Ud) = _restricted(A, scale, stats=stats)
    n = A.shape[0]
    moved = int((F.prow != np.arange(n)).sum())
    assert moved > 0 and stats["swaps"] > 0
    assert np.array_equal(perm, F.prow)
    assert np.array_equal(F.p, np.asarray(S.qf)[perm])
    # pivoting restricted to a supernode's rows bounds neither L21 nor the growth of U (weak-8:
    # max |L| ~ 110); the blocked and the dense elimination round differently, so the values
    # agree relative to the largest factor entry
    assert np.abs(F.L.toarray() - Ld).max() <= 1e-12 * np.abs(Ld).max()
    assert np.abs(F.U.toarray() - Ud).max() <= 1e-12 * np.abs(Ud).max()
    rec = abs(As[F.p, :][:, F.q] - F.L @ F.U).max() / abs(As).max()
    assert rec <= 1e-14


@pytest.mark.parametrize("name", list(PIVOT_MATS))
def test_pivoting_blas_backend_matches_loops(name):
    """The OpenBLAS backend (the reference's own getrf/trsm/gemm calls) and the loop backend give
    the same interchanges on matrices that pivot."""
    A, scale = PIVOT_MATS[name]
    A = A()
    F0 = ON.splu_mf(A, scale=scale, backend=0)
    F1 = ON.splu_mf(A, scale=scale, backend=1)
    assert (F0.prow != np.arange(A.shape[0])).any()
    assert np.array_equal(F0.p, F1.p) and np.array_equal(F0.prow, F1.prow)
    assert np.array_equal(F0.L.indices, F1.L.indices) and np.array_equal(F0.U.indices, F1.U.indices)
    assert np.abs(F0.L.data - F1.L.data).max() <= 1e-12 * np.abs(F0.L.data).max()
    assert np.abs(F0.U.data - F1.U.data).max() <= 1e-12 * np.abs(F0.U.data).max()


@pytest.mark.parametrize("name", ["signs-60", "signs-150"])
def test_pivoting_ties_take_the_first_maximum(name):
    """+-1 matrices under SCALE_NONE: the pivot columns hold small integers, so exact ties of the
    maximum magnitude occur and the elimination is exact.  The oracle must take the first
    row of maximum magnitude (BLAS idamax, src/multifrontal.jl:248 lu!(A11, RowMaximum()))."""
    A, scale = PIVOT_MATS[name]
    A = A()
    stats = {}
    F, As, V, S, (perm, Ld, Ud) = _restricted(A, scale, stats=stats)
    assert stats["ties"] > 0
    assert np.array_equal(perm, F.prow)
    assert np.array_equal(F.L.toarray(), Ld) and np.array_equal(F.U.toarray(), Ud)
    # planted mistake: the last-maximum rule changes the interchanges, so the tests above see it
    perm_last, _, _ = ON.dense_restricted_lu(V, S.sstart, tie="last")
    assert not np.array_equal(perm_last, F.prow)


def test_pivoting_planted_unswapped_r12_is_caught():
    """Planted mistake: dropping the R12 row interchanges (src/multifrontal.jl:256-263) breaks the
    reconstruction on a matrix that pivots."""
    A = M.sprand_identity(400, seed=1)
    F = ON.splu_mf(A)
    As = OS.apply_row_scaling(A, F.Rs)
    S = OY.symbolic_mf(As)
    V = As[S.qf, :][:, S.qf].toarray()
    W = V.copy()
    n = W.shape[0]
    for s in range(S.nsuper):
        c1, c2 = S.sstart[s], S.sstart[s + 1] - 1
        for j in range(c1, c2 + 1):
            p = j + int(np.argmax(np.abs(W[j:c2 + 1, j])))
            if p != j:
                W[[j, p], :c2 + 1] = W[[p, j], :c2 + 1]      # the columns beyond c2 keep their rows
            W[j + 1:, j] /= W[j, j]
            W[j + 1:, j + 1:] -= np.outer(W[j + 1:, j], W[j, j + 1:])
    assert np.abs(np.triu(W) - F.U.toarray()).max() > 1e-3
