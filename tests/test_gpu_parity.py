"""GPU parity of the FP64 multifrontal path against the CPU oracle (run on an MI355X).

Each case builds the symbolic analysis once (library C++ analysis, itself checked against
the oracle in test_analysis.py), factors on the GPU through the C ABI and with the oracle's
restatement of src/multifrontal.jl on the same arrays, and requires
  * identical p, q and prow (row pivots);
  * identical L and U CSC patterns, including the unsorted relabelled L update rows
    (src/multifrontal.jl:342-350);
  * factor values within 1e-12 relative (max |GPU - oracle| / max |oracle| per factor);
  * reconstruction (Rs .* A)[p, q] = L U to 1e-13 relative.
small_nf=1 routes every front with nf > 1 through the large-front kernels (column assembly,
register panel getrf, grouped MFMA gemm); the default routes small fronts through the fused
LDS kernel, so both paths are covered on every matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import numeric as ON
from oracle import scaling as OS

from _util import max_rel, oracle_splu, same_pattern

pytestmark = pytest.mark.gpu

CASES = {
    "spd6": lambda: M.spd6(),
    "poisson2d-8": lambda: M.poisson2d(8),
    "poisson2d-24": lambda: M.poisson2d(24),
    "poisson2d-50": lambda: M.poisson2d(50),
    "poisson3d-6": lambda: M.poisson3d(6),
    "poisson3d-10": lambda: M.poisson3d(10),
    "poisson3d-16": lambda: M.poisson3d(16),
    "randmat-500": lambda: M.randmat(500, 6),
    "arrowband-400": lambda: M.arrowband(400, 4),
    "convdiff-8": lambda: M.convdiff3d(8),
}


def _check_parity(A, F, R, tol=1e-12):
    assert np.array_equal(F.p, R.p), "row permutation p differs"
    assert np.array_equal(F.q, R.q), "column permutation q differs"
    assert same_pattern(F.L, R.L), "L pattern / row order differs"
    assert same_pattern(F.U, R.U), "U pattern differs"
    assert max_rel(F.L.data, R.L.data) <= tol
    assert max_rel(F.U.data, R.U.data) <= tol


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("small_nf,pivoting", [(0, "auto"), (1, "auto"), (1, "partial")])
def test_factor_parity(name, small_nf, pivoting):
    """pivoting="auto" tries the interchange-free block schedule for the large fronts and
    falls back to the partial-pivoting panels when a multiplier exceeds 1; "partial" always
    runs the panels.  Both must reproduce the oracle exactly."""
    A = CASES[name]()
    print(f"[case] {name} small_nf={small_nf} pivoting={pivoting} n={A.shape[0]}", flush=True)
    S = I.symbolic_mf(A)
    F = I.splu(A, method="multifrontal", symbolic=S, small_nf=small_nf, pivoting=pivoting)
    R = oracle_splu(A, S)
    _check_parity(A, F, R)
    if pivoting == "partial":
        assert F._plan.stats()["last_schedule"] == 0
    As = OS.apply_row_scaling(A, F.Rs)
    rec = abs(As[F.p, :][:, F.q] - F.L @ F.U).max() / abs(As).max()
    assert rec <= 1e-13


@pytest.mark.parametrize("scale", [OS.SCALE_NONE, OS.SCALE_SUM, OS.SCALE_MAX])
def test_scalings(scale):
    A = M.poisson3d(8)
    S = I.symbolic_mf(A)
    F = I.splu(A, method="multifrontal", scale=scale, symbolic=S)
    R = oracle_splu(A, S, scale=scale)
    _check_parity(A, F, R)
    b = np.random.default_rng(7).standard_normal(A.shape[0])
    x = I.solve(F, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-8
    xr = ON.solve(R, b)
    assert np.abs(x - xr).max() / np.abs(xr).max() <= 1e-12


def test_poisson3d_24_parity():
    """The FP64 parity workload: 3D Poisson 24^3 (n = 13824), AMD, SCALE_SUM, 1 RHS."""
    A = M.poisson3d(24)
    S = I.symbolic_mf(A)
    assert S.predicted_fill() == 3749118 and S.nsuper == 9348
    F = I.splu(A, method="multifrontal", symbolic=S)
    R = oracle_splu(A, S)
    _check_parity(A, F, R)
    # no interchange is needed on this matrix, so the interchange-free schedule must have held
    assert F._plan.stats()["last_schedule"] == 1
    Fp = I.splu(A, method="multifrontal", symbolic=S, pivoting="partial")
    _check_parity(A, Fp, R)
    b = np.random.default_rng(2).standard_normal(A.shape[0])
    x = I.solve(F, b)
    xr = ON.solve(R, b)
    assert np.abs(x - xr).max() / np.abs(xr).max() <= 1e-12
    assert F.last_report["backward_error"] <= 1e-15


def test_solve_refine_and_agreement():
    """Residual and agreement bounds of test/multifrontal_test.jl:53-61 (with scipy's SuperLU
    standing in for UMFPACK), plus fixed-count refinement semantics."""
    rng = np.random.default_rng(2024)
    for A in (M.poisson2d(30), M.poisson3d(8), M.poisson3d(12), M.randmat(800, 6)):
        b = rng.standard_normal(A.shape[0])
        F = I.splu(A, method="multifrontal")
        x = I.ldiv(F, b)
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-8
        xs = spla.splu(A.tocsc()).solve(b)
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= 1e-7
        x1 = I.solve(F, b, refine=1)
        assert F.last_report["iterations"] == 1
        assert np.linalg.norm(A @ x1 - b) / np.linalg.norm(b) <= 1e-8


def test_multi_rhs_matches_columns():
    A = M.poisson3d(10)
    F = I.splu(A, method="multifrontal")
    B = np.random.default_rng(3).standard_normal((A.shape[0], 5))
    X = I.solve(F, B, refine=2)
    for j in range(5):
        xj = I.solve(F, B[:, j], refine=2)
        assert np.abs(X[:, j] - xj).max() <= 1e-14 * np.abs(xj).max()


def test_doctests():
    A = sp.csc_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    want = np.array([0.6666666666666666, -0.3333333333333333])
    assert np.allclose(I.ldiv(I.splu(A, method="multifrontal"), np.array([1.0, 0.0])), want, rtol=0, atol=1e-16)
    assert np.allclose(I.solve(I.splu(A, method="multifrontal"), np.array([1.0, 0.0]), refine=1), want,
                       rtol=0, atol=1e-16)
    assert np.allclose(I.ldiv(I.multifrontal_lu(A), np.array([1.0, 0.0])), want, rtol=0, atol=1e-16)


def test_edge_cases():
    one = sp.csc_matrix(np.array([[4.0]]))
    assert abs(I.ldiv(I.splu(one, method="multifrontal"), np.array([8.0]))[0] - 2.0) <= 1e-15
    D = sp.csc_matrix(np.diag([2.0, 3.0, 5.0, 7.0]))
    x = I.ldiv(I.splu(D, method="multifrontal"), np.array([2.0, 6.0, 15.0, 28.0]))
    assert np.allclose(x, [1, 2, 3, 4], rtol=1e-15)
    with pytest.raises(I.ArgumentError):
        I.splu(D, method="nonsense")
    with pytest.raises(I.DimensionMismatch):
        I.splu(sp.csc_matrix(np.ones((2, 3))), method="multifrontal")


def test_singular_pivot():
    """An exactly-zero block pivot raises SingularException(column) like
    src/multifrontal.jl:250-251; with check=False the factorization completes."""
    D = sp.csc_matrix(np.diag([2.0, 0.0, 5.0]))
    S = I.symbolic_mf(D)
    with pytest.raises(I.SingularException) as ei:
        I.splu(D, method="multifrontal", scale=I.SCALE_NONE, symbolic=S)
    col = int(np.nonzero(S.qf == 1)[0][0]) + 1          # factor column holding the zero
    assert ei.value.info == col
    F = I.splu(D, method="multifrontal", scale=I.SCALE_NONE, check=False, symbolic=S)
    assert F.U.shape == (3, 3)


def test_subnormal_pivot_in_interchange_free_block():
    """A dense 100 x 100 matrix (one front, a full 64-column block) whose factor column 5 is
    zero except a subnormal diagonal: the interchange-free block meets a tiny pivot with no
    multiplier above the bound, so the block is redone with getf2's rules (division by the
    tiny pivot), and the factors must still match the oracle."""
    rng = np.random.default_rng(17)
    n = 100
    D = rng.uniform(-1.0, 1.0, (n, n)) + np.diag(np.full(n, 2.0 * n))
    A0 = sp.csc_matrix(D)
    S = I.symbolic_mf(A0)
    t = 5
    qc = int(S.qf[t])
    D[:, qc] = 0.0
    D[qc, qc] = 1e-310
    rows, cols = np.nonzero(np.ones_like(D))   # the dense pattern, zeros kept as stored entries
    A = sp.csc_matrix((D[rows, cols], (rows, cols)), shape=D.shape)
    assert A.nnz == n * n
    print(f"[case] dense n={n} subnormal pivot in factor column {t}", flush=True)
    F = I.splu(A, method="multifrontal", scale=I.SCALE_NONE, symbolic=S)
    R = oracle_splu(A, S, scale=OS.SCALE_NONE)
    _check_parity(A, F, R)


def test_deterministic():
    A = M.randmat(500, 6)
    S = I.symbolic_mf(A)
    F1 = I.splu(A, method="multifrontal", symbolic=S, small_nf=1)
    F2 = I.splu(A, method="multifrontal", symbolic=S, small_nf=1)
    assert np.array_equal(F1.L.data, F2.L.data) and np.array_equal(F1.U.data, F2.U.data)


@pytest.mark.parametrize("name,k", [("poisson3d-10", 10), ("poisson3d-24", 24)])
def test_many_rhs_chunks_match_columns(name, k):
    """More right-hand sides than one chunk of the solve kernels (chunks run side by side in
    grid z): every column equals its own single-RHS solve."""
    A = M.poisson3d(k)
    F = I.splu(A, method="multifrontal")
    B = np.random.default_rng(13).standard_normal((A.shape[0], 21))
    X = I.solve(F, B, refine=1)
    for j in (0, 7, 8, 15, 16, 20):
        xj = I.solve(F, B[:, j], refine=1)
        assert np.abs(X[:, j] - xj).max() <= 1e-14 * np.abs(xj).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replayed_solves_match_eager(precision):
    """With use_graph the solve sweeps are captured once and replayed: every replay (several
    solves in a row, with refinement steps in between, and a change of the RHS count that forces
    a new capture) gives the eager plan's result bit for bit."""
    A = M.poisson3d_perturbed(12, seed=5)
    S = I.symbolic_mf(A)
    Fe = I.splu(A, method="multifrontal", symbolic=S, precision=precision)
    Fg = I.splu(A, method="multifrontal", symbolic=S, precision=precision, use_graph=True)
    rng = np.random.default_rng(6)
    for nrhs in (3, 3, 5, 3):
        B = rng.standard_normal((A.shape[0], nrhs))
        Xe = I.solve(Fe, B, refine=3, tol=1e-15)
        Xg = I.solve(Fg, B, refine=3, tol=1e-15)
        assert np.array_equal(Xe, Xg)
        assert Fe.last_report["iterations"] == Fg.last_report["iterations"]
