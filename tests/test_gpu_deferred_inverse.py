"""The pivot-block inversions of the top level's fronts are left out of the factorization and
run by the first solve after it, on a second stream beside the sweeps of the levels below
(which read no inverse of the top level), joined just before the top level's sweep.

Every schedule of them gives the same bits: inside the first solve, eager or replayed from a
captured graph; after several factorizations without a solve; after a solve on the
column-parallel path (which reads no inverse and so leaves them pending); and at the end of
the factorization when every launch is timed.  The solution is pinned by the oracle's solve
(src/solve.jl:10-39 through oracle/numeric.py).
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import numeric as ON

from _util import oracle_splu

pytestmark = pytest.mark.gpu


def test_deferred_top_inversion_schedules_agree():
    A = M.poisson3d(24)
    S = I.symbolic_mf(A)
    npv = np.diff(np.asarray(S.sstart))
    print(f"[case] deferred inversion poisson3d-24: max np {npv.max()}", flush=True)
    assert npv.max() > 128   # the top front keeps inverses
    Rs = I.row_scaling(A, I.SCALE_SUM)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(A.shape[0])
    Bw = rng.standard_normal((A.shape[0], 600))   # enough columns for the column-parallel path
    ref = None
    for graph in (False, True):
        P = I._Plan(A, S, use_graph=graph)
        P.factor(A.data, Rs)
        x1 = P.solve(b)                       # the inversions run inside this solve
        assert np.array_equal(P.solve(b), x1)  # and not again
        P.factor(A.data, Rs)
        P.factor(A.data, Rs)                  # two factorizations, then one solve
        assert np.array_equal(P.solve(b), x1)
        P.factor(A.data, Rs)
        Xw = P.solve(Bw)                      # column path: the inversions stay pending
        assert np.abs(A @ Xw - Bw).max() <= 1e-10 * np.abs(Bw).max()
        assert np.array_equal(P.solve(b), x1)  # the fronts path then runs them
        if ref is None:
            ref = x1
        assert np.array_equal(x1, ref)
    # with every launch timed they stay at the end of the factorization
    P = I._Plan(A, S)
    assert P.lib.mf_plan_set_profiling(P.h, 1) == 0
    P.factor(A.data, Rs)
    assert np.array_equal(P.solve(b), ref)
    R = oracle_splu(A, S)
    xo = ON.solve(R, b)
    assert np.abs(ref - xo).max() <= 1e-12 * np.abs(xo).max()
