"""Relaxed supernode amalgamation on the GPU (opt-in; run on an MI355X).

The relaxed analysis (mf_symbolic_relax) merges supernodes into their parents and stores
explicit zeros; the factorization of that structure is still exact.  The oracle factors the
same relaxed structure on the CPU, so the FP64 path must reproduce its p, q, L/U pattern and
values exactly as on the reference's structure; the solutions agree with the unrelaxed path,
and the mixed-precision paths still refine to their backward-error targets with fewer levels
and launches.
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import scaling as OS

from _util import max_rel, oracle_splu, same_pattern

pytestmark = pytest.mark.gpu

CASES = {
    "poisson3d-16": (lambda: M.poisson3d(16), OS.SCALE_SUM),
    "sprand-2000": (lambda: M.sprand_identity(2000, seed=2), OS.SCALE_SUM),
    "weak-16": (lambda: M.poisson3d_weak(16), OS.SCALE_SUM),
    "convdiff-12": (lambda: M.convdiff3d(12), OS.SCALE_SUM),
}


@pytest.mark.parametrize("relax", [True, (0.1, 8)])
@pytest.mark.parametrize("name", list(CASES))
def test_relaxed_fp64_matches_oracle(name, relax):
    make, scale = CASES[name]
    A = make()
    print(f"[case] relaxed {name} relax={relax}", flush=True)
    S0 = I.symbolic_mf(A)
    S = I.symbolic_mf(A, relax=relax)
    assert S.nsuper < S0.nsuper
    R = oracle_splu(A, S, scale=scale)
    F = I.splu(A, method="multifrontal", scale=scale, symbolic=S)
    assert np.array_equal(F.p, R.p) and np.array_equal(F.q, R.q)
    assert same_pattern(F.L, R.L) and same_pattern(F.U, R.U)
    assert F.L.nnz + F.U.nnz == S.predicted_fill()
    assert max_rel(F.L.data, R.L.data) <= 1e-12
    assert max_rel(F.U.data, R.U.data) <= 1e-12
    b = np.random.default_rng(4).standard_normal(A.shape[0])
    x = I.solve(F, b, refine=1)
    x0 = I.solve(I.splu(A, method="multifrontal", scale=scale, symbolic=S0), b, refine=1)
    assert np.abs(x - x0).max() <= 1e-11 * np.abs(x0).max()


def test_relaxed_mixed_poisson48():
    """The 48^3 workload on the relaxed analysis: fewer levels and launches, FP32 fronts refined
    to backward error <= 1e-14."""
    A = M.poisson3d_perturbed(48, seed=3)
    S0 = I.symbolic_mf(A)
    S = I.symbolic_mf(A, relax=True)
    e0, e1 = I.plan_estimate(A, S0, precision="fp32"), I.plan_estimate(A, S, precision="fp32")
    print(f"[case] relaxed poisson3d-48: levels {e0['nlevels']} -> {e1['nlevels']}, launches {e0['launches']} -> "
          f"{e1['launches']}", flush=True)
    assert e1["nlevels"] < e0["nlevels"] and e1["launches"] < e0["launches"]
    B = np.random.default_rng(3).standard_normal((A.shape[0], 8))
    F = I.splu(A, method="multifrontal", symbolic=S, precision="fp32")
    X = I.solve(F, B, refine=30, tol=1e-14)
    assert F.last_report["converged"] == 1
    R = B - A @ X
    be = np.abs(R).max(axis=0) / (abs(A).sum(axis=1).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))
    assert be.max() <= 1e-14
