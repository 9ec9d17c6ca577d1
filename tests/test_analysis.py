"""The library's host analysis (mf_analyze, C++) against the oracle and the published figures.
Runs on the CPU: mf_analyze makes no GPU call."""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import amd as OA
from oracle import numeric as ON
from oracle import scaling as OS
from oracle import symbolic as OY

from _util import oracle_splu

MATS = {
    "poisson2d-8": lambda: M.poisson2d(8),
    "poisson2d-50": lambda: M.poisson2d(50),
    "poisson3d-6": lambda: M.poisson3d(6),
    "poisson3d-16": lambda: M.poisson3d(16),
    "randmat-300": lambda: M.randmat(300, 6),
    "arrowband-200": lambda: M.arrowband(200, 4),
    "convdiff-8": lambda: M.convdiff3d(8),
    "diag4": lambda: M.poisson2d(1),
}


@pytest.mark.parametrize("ordering", ["amd", "colamd"])
@pytest.mark.parametrize("name", list(MATS))
def test_analysis_matches_oracle(name, ordering):
    A = MATS[name]()
    G = I.symbolic_mf(A, ordering=ordering)
    O = OY.symbolic_mf(A) if ordering == "amd" else OY.symbolic_mf(A, q=OA.colamd_order(A))
    ss, up, ui, cp, ci = ON.flatten_symbolic(O)
    assert np.array_equal(G.qf, O.qf)
    assert np.array_equal(G.parent, np.asarray(O.parent))
    assert np.array_equal(G.sstart, ss)
    assert np.array_equal(G.upd_ptr, up) and np.array_equal(G.upd_idx, ui)
    assert np.array_equal(G.child_ptr, cp) and np.array_equal(G.child_idx, ci)
    assert G.predicted_fill() == OY.predicted_fill(O)


def test_given_order():
    A = M.randmat(200, 5)
    q = np.random.default_rng(5).permutation(200)
    G = I.symbolic_mf(A, q=q)
    O = OY.symbolic_mf(A, q=q)
    assert np.array_equal(G.qf, O.qf) and G.predicted_fill() == OY.predicted_fill(O)


@pytest.mark.parametrize("k,fill,nsuper", [(16, 562028, 2792), (20, 1684564, 5445), (24, 3749118, 9348)])
def test_published_poisson3d(k, fill, nsuper):
    G = I.symbolic_mf(M.poisson3d(k))
    assert G.predicted_fill() == fill and G.nsuper == nsuper


def test_published_poisson2d():
    assert I.symbolic_mf(M.poisson2d(100)).predicted_fill() == 412664


def test_bad_inputs():
    import scipy.sparse as sp
    with pytest.raises(I.DimensionMismatch):
        I.symbolic_mf(sp.csc_matrix(np.ones((2, 3))))
    with pytest.raises(I.ArgumentError):
        I.symbolic_mf(M.poisson2d(3), q=np.array([0, 0, 1, 2, 3, 4, 5, 6, 7]))
    with pytest.raises(I.ArgumentError):
        I.symbolic_mf(M.poisson2d(3), ordering="nested")


def test_row_scaling_matches_oracle():
    from oracle import scaling as OS
    for A in (M.randmat(300, 6), M.convdiff3d(6)):
        for k in (OS.SCALE_NONE, OS.SCALE_SUM, OS.SCALE_MAX):
            assert np.array_equal(I.row_scaling(A, k), OS.row_scaling(A, k))


def _threshold_edge_matrix(n, seed):
    """Sparse random matrix whose first rows hold exactly thr - 1, thr, thr + 1 and thr + 2
    entries, thr = floor(min(n - 2, max(16, 10 sqrt(n)))) the dense-row threshold of
    colamd_order (src/amd.jl:96)."""
    import math
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    thr = int(math.floor(min(n - 2, max(16.0, 10.0 * math.sqrt(n)))))
    D = (rng.random((n, n)) < 3.0 / n).astype(float)
    np.fill_diagonal(D, 1.0)
    for r, cnt in enumerate((thr - 1, thr, thr + 1, thr + 2)):
        D[r, :] = 0.0
        D[r, rng.choice(n, cnt, replace=False)] = 1.0
    rows, cols = np.nonzero(D)
    return sp.csc_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, n)), thr


@pytest.mark.parametrize("n,seed", [(36, 1), (40, 2), (150, 3), (400, 4)])
def test_colamd_ata_pattern_brute_force(n, seed):
    """The A^T A pattern of colamd_order (src/amd.jl:91-104): a dense boolean product over the
    rows of A with at most thr entries, diagonal dropped.  Pins the oracle's construction; the
    library's (mf_analysis.cpp) then gives the same COLAMD q as the oracle on these matrices."""
    A, thr = _threshold_edge_matrix(n, seed)
    rc = np.diff(A.tocsr().indptr)
    assert rc[0] == thr - 1 and rc[1] == thr and rc[2] == thr + 1 and rc[3] == thr + 2
    P = (A.toarray() != 0).astype(np.int64)
    keep = rc <= thr
    C = (P[keep].T @ P[keep]) > 0
    np.fill_diagonal(C, False)
    cp, ci = OA.ata_pattern(A)
    got = np.zeros((n, n), bool)
    for j in range(n):
        got[ci[cp[j]:cp[j + 1]], j] = True
        assert np.all(np.diff(ci[cp[j]:cp[j + 1]]) > 0)
    assert np.array_equal(got, C)
    # the boundary matters: dropping the row with exactly thr entries (an off-by-one threshold)
    # or keeping the row with thr + 1 gives another pattern
    for wrong in (rc <= thr - 1, rc <= thr + 1):
        W = (P[wrong].T @ P[wrong]) > 0
        np.fill_diagonal(W, False)
        assert not np.array_equal(W, C)
    G = I.symbolic_mf(A, ordering="colamd")
    assert np.array_equal(G.qf, OY.symbolic_mf(A, q=OA.colamd_order(A)).qf)


@pytest.mark.parametrize("name", ["poisson3d-16", "randmat-300", "convdiff-8", "arrowband-200"])
@pytest.mark.parametrize("relax", [True, (0.1, 8), (0.3, 32)])
def test_relaxed_amalgamation_structure(name, relax):
    """Relaxed amalgamation (opt-in) keeps qf, merges supernodes into their parents and only adds
    explicit zeros: every column's relaxed row set contains its exact structure (the oracle's
    colstruct, src/symbolic.jl:122-160), supernodes stay contiguous with nested structures, each
    child's update rows lie in its parent's front, and nnz(L) + nnz(U) = predicted_fill."""
    A = MATS[name]()
    S0 = I.symbolic_mf(A)
    S1 = I.symbolic_mf(A, relax=relax)
    assert np.array_equal(S0.qf, S1.qf)
    assert S1.nsuper < S0.nsuper and S1.predicted_fill() >= S0.predicted_fill()
    O = OY.symbolic_mf(A)
    n = A.shape[0]
    ss = np.asarray(S1.sstart)
    total = 0
    for t in range(S1.nsuper):
        c1, c2 = ss[t], ss[t + 1] - 1
        upd = set(int(v) for v in S1.upd(t))
        assert all(v > c2 for v in upd)
        for j in range(c1, c2 + 1):
            rows = set(range(j + 1, c2 + 1)) | upd
            assert set(O.colstruct[j]) <= rows
            total += len(rows)
        front = set(range(c1, c2 + 1)) | upd
        for c in S1.children(t):
            assert c < t and set(int(v) for v in S1.upd(c)) <= front
    assert S1.predicted_fill() == 2 * (n + total)
    # the numeric oracle factors the relaxed structure exactly (explicit zeros included)
    R = oracle_splu(A, S1)
    As = OS.apply_row_scaling(A, R.Rs)
    assert R.L.nnz + R.U.nnz == S1.predicted_fill()
    assert abs(As[R.p, :][:, R.q] - R.L @ R.U).max() <= 1e-13 * abs(As).max()


def test_relaxed_amalgamation_argument_checks():
    A = M.poisson3d(4)
    with pytest.raises(I.ArgumentError):
        I.symbolic_mf(A, relax=(1.5, 0))
    with pytest.raises(I.ArgumentError):
        I.symbolic_mf(A, relax=(0.1, -1))
