"""The library's host analysis (mf_analyze, C++) against the oracle and the published figures.
Runs on the CPU: mf_analyze makes no GPU call."""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import amd as OA
from oracle import numeric as ON
from oracle import symbolic as OY

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
