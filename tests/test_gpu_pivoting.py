"""GPU parity on matrices whose in-block partial pivoting interchanges rows (run on an MI355X).

The FP64 path must reproduce the oracle's prow, p, q, the L/U patterns (including the
relabelled, unsorted L update rows of src/multifrontal.jl:342-350) and the factor values on
matrices where the pivoting of src/multifrontal.jl:247-275 actually moves rows:
  * sprand(n, n, 12/n) + I, the family of the reference's test/pureumfpack_test.jl:106;
  * dense random fronts (one supernode: n <= 64 goes through the fused small-front kernel,
    larger ones through the large-front kernels);
  * +-1 matrices under SCALE_NONE, where exact magnitude ties decide the interchanges (the
    first row of maximum magnitude wins, as BLAS idamax);
  * the batched workload's unsymmetric values with a weakened diagonal on 3D Poisson patterns
    (diagonals chosen so that no pivot block is near singular: with a pivot of 3e-5, as
    poisson3d_weak(12) has, the oracle's own loop and OpenBLAS backends already differ by 1e-11,
    and a 1e-12 bar would measure the conditioning, not the kernels).
Every case runs with small_nf 0 (small fronts in the fused LDS kernel) and 1 (every front
with nf > 1 through the large-front kernels), and with pivoting "auto" (interchange-free
blocks with a multiplier check, redone with the pivoted panels when the check fails) and
"partial" (always the pivoted panels).  The oracle itself is pinned on these matrices against
dense restricted elimination in tests/test_oracle.py.
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import scaling as OS

from _util import max_rel, oracle_splu, same_pattern

pytestmark = pytest.mark.gpu

CASES = {
    "sprand-400": (lambda: M.sprand_identity(400, seed=1), OS.SCALE_SUM),
    "sprand-2000": (lambda: M.sprand_identity(2000, seed=2), OS.SCALE_SUM),
    "dense-40": (lambda: M.dense_random(40, seed=5), OS.SCALE_SUM),
    "dense-100": (lambda: M.dense_random(100, seed=3), OS.SCALE_SUM),
    "dense-300": (lambda: M.dense_random(300, seed=4), OS.SCALE_SUM),
    "signs-60": (lambda: M.signs(60, 0.3, 1), OS.SCALE_NONE),
    "signs-150": (lambda: M.signs(150, 0.5, 3), OS.SCALE_NONE),
    "weak-6": (lambda: M.poisson3d_weak(6), OS.SCALE_SUM),
    "weak-12": (lambda: M.poisson3d_weak(12, lo=3.0, hi=4.0), OS.SCALE_SUM),
    "weak-16": (lambda: M.poisson3d_weak(16), OS.SCALE_SUM),
}


def _moved_in_large_fronts(S, prow, small_nf):
    """(rows the oracle moves inside fronts the large-front kernels factor, all rows moved,
    whether there are such fronts); large means nf > small_nf, the library's default being 64."""
    thr = small_nf if small_nf > 0 else 64
    ss, up = np.asarray(S.sstart), np.asarray(S.upd_ptr)
    npiv = np.diff(ss)
    nf = npiv + np.diff(up)
    sn_of = np.repeat(np.arange(S.nsuper), npiv)
    moved = prow != np.arange(len(prow))
    return int((moved & (nf[sn_of] > thr)).sum()), int(moved.sum()), bool((nf > thr).any())


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("small_nf,pivoting", [(0, "auto"), (0, "partial"), (1, "auto"), (1, "partial")])
def test_pivoting_parity(name, small_nf, pivoting):
    make, scale = CASES[name]
    A = make()
    print(f"[case] {name} small_nf={small_nf} pivoting={pivoting} n={A.shape[0]}", flush=True)
    S = I.symbolic_mf(A)
    R = oracle_splu(A, S, scale=scale)
    big, moved, has_large = _moved_in_large_fronts(S, R.prow, small_nf)
    assert moved > 0, "the case must interchange rows"
    F = I.splu(A, method="multifrontal", scale=scale, symbolic=S, small_nf=small_nf, pivoting=pivoting)
    assert np.array_equal(F.p, R.p), "row permutation p differs"
    assert np.array_equal(F.q, R.q)
    assert same_pattern(F.L, R.L), "L pattern / row order differs"
    assert same_pattern(F.U, R.U), "U pattern differs"
    # values relative to the largest factor entry (restricted pivoting lets L21 and U grow)
    assert max_rel(F.L.data, R.L.data) <= 1e-12
    assert max_rel(F.U.data, R.U.data) <= 1e-12
    last = F._plan.stats()["last_schedule"]
    if pivoting == "partial" or big > 0 or not has_large:
        assert last == 0, "a multiplier above 1 in a large front must send AUTO to the pivoted panels"
    else:
        assert last == 1, "no interchange in a large front: the interchange-free blocks hold"
    As = A if scale == OS.SCALE_NONE else OS.apply_row_scaling(A, F.Rs)
    rec = abs(As[F.p, :][:, F.q] - F.L @ F.U).max() / abs(As).max()
    assert rec <= 1e-13 * max(1.0, abs(F.L).max() * abs(F.U).max() / abs(As).max())


@pytest.mark.parametrize("name", ["sprand-2000", "weak-12", "signs-150"])
def test_pivoting_solve_matches_oracle(name):
    """Solves through the pivoted factors (prow applied on the device) agree with the oracle's."""
    from oracle import numeric as ON
    make, scale = CASES[name]
    A = make()
    S = I.symbolic_mf(A)
    F = I.splu(A, method="multifrontal", scale=scale, symbolic=S)
    R = oracle_splu(A, S, scale=scale)
    b = np.random.default_rng(9).standard_normal(A.shape[0])
    x = I.solve(F, b, refine=1)
    xr = ON.solve(R, b, refine=1)
    assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) <= 1e-10


def test_auto_retries_the_interchange_free_schedule():
    """The fallback is per call: after a matrix that needs interchanges, a matrix of the same
    pattern that needs none is factored by the interchange-free blocks again."""
    A = M.poisson3d_weak(12)
    S = I.symbolic_mf(A)
    plan = I._Plan(A, S)
    plan.factor(A.data, I.row_scaling(A))
    assert plan.stats()["last_schedule"] == 0
    D = M.poisson3d(12)
    plan.factor(D.data, I.row_scaling(D))
    assert plan.stats()["last_schedule"] == 1
    plan.factor(A.data, I.row_scaling(A))
    assert plan.stats()["last_schedule"] == 0


def test_batch_redoes_only_the_matrices_that_pivot():
    """A batch mixing matrices that pivot and matrices that do not: every matrix matches its own
    single-matrix oracle factorization."""
    P = M.poisson3d(8)
    S = I.symbolic_mf(P)
    weak = M.poisson3d_weak(8)
    _, Ax = M.poisson3d_batch_values(8, 3, seed=11)
    vals = np.stack([Ax[0], weak.data, Ax[1], weak.data * 1.5, Ax[2]])
    B = I.BatchedLU(P, 5, precision="fp64", symbolic=S, use_graph=False)
    B.refactor(vals)
    st = B._plan.stats()
    assert st["last_redone"] == 2 and st["last_schedule"] == 0
    for m in range(5):
        Am = B.matrix(vals[m])
        R = oracle_splu(Am, S)
        L, U, p, q = B.factors(m)
        assert np.array_equal(p, R.p) and np.array_equal(q, R.q)
        assert same_pattern(L, R.L) and same_pattern(U, R.U)
        assert max_rel(L.data, R.L.data) <= 1e-12 and max_rel(U.data, R.U.data) <= 1e-12


@pytest.mark.parametrize("pivoting", ["auto", "partial"])
def test_pivot_block_taller_than_registers(pivoting):
    """A dense 3300 x 3300 front (one supernode, np = 3300 > 3072 rows): the pivoted panels of
    its first blocks run in the kernel that keeps the panel in memory (any height), the rest
    in the register panels; both must give the oracle's interchanges and factors."""
    A = M.dense_random(3300, seed=8)
    print(f"[case] dense-3300 pivoting={pivoting}", flush=True)
    S = I.symbolic_mf(A)
    assert S.nsuper == 1
    R = oracle_splu(A, S, backend=1)
    assert (R.prow != np.arange(A.shape[0])).sum() > 3000
    F = I.splu(A, method="multifrontal", symbolic=S, pivoting=pivoting)
    assert np.array_equal(F.p, R.p) and np.array_equal(F.q, R.q)
    assert same_pattern(F.L, R.L) and same_pattern(F.U, R.U)
    # a 3300-row dense elimination rounds differently in every order: the oracle's own loop and
    # OpenBLAS backends differ by 2.3e-12 (L) and 6.1e-13 (U) on this matrix
    assert max_rel(F.L.data, R.L.data) <= 1e-11
    assert max_rel(F.U.data, R.U.data) <= 1e-11
    assert F._plan.stats()["last_schedule"] == 0
