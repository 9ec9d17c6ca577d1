"""The measured workloads at their full sizes (run on an MI355X).

* 3D Poisson 48^3 + Diagonal(U[0, 0.1]) (n = 110592), AMD, SCALE_SUM, 8 RHS randn(3): the FP64
  path against the oracle (p, q, L/U pattern and row order, values), then FP32 fronts with FP64
  refinement to a normwise backward error <= 1e-14 on every RHS, its error reported against
  the FP64 solve, which is itself checked against the oracle's solve.
* 512 matrices on the 3D Poisson 16^3 pattern (unsymmetric values), 64 RHS each, FP32 fronts +
  FP64 refinement, hipGraph replay: the backward error of every matrix and RHS checked on the
  host; oracle parity of the FP64 batch path for some matrices.
* 3D convection-diffusion 64^3 (n = 262144), COLAMD, BF16 Schur complements with per-tile
  power-of-two scaling, FP32 getrf/trsm, FP64 refinement to <= 1e-12: fits one GPU with the
  factor store and the liveness-planned arena (every front resident would take 616 GB).
The reference's own tests stop at far smaller sizes (test/multifrontal_test.jl:53-61); the
bounds are theirs (residual, agreement) tightened to what FP64 refinement reaches.
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import numeric as ON

from _util import max_rel, oracle_splu, same_pattern

pytestmark = pytest.mark.gpu


def _bwd(A, X, B):
    """Normwise backward error of every column: ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    R = B - A @ X
    an = abs(A).sum(axis=1).max()
    return np.abs(R).max(axis=0) / (an * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))


@pytest.fixture(scope="module")
def poisson48():
    A = M.poisson3d_perturbed(48, seed=3)
    S = I.symbolic_mf(A)
    B = np.random.default_rng(3).standard_normal((A.shape[0], 8))
    return A, S, B


def test_poisson48_fp64_matches_oracle(poisson48):
    A, S, B = poisson48
    print(f"[case] poisson3d-48 perturbed fp64 n={A.shape[0]} fill={S.predicted_fill()}", flush=True)
    F = I.splu(A, method="multifrontal", symbolic=S)
    R = oracle_splu(A, S, backend=1)
    assert np.array_equal(F.p, R.p) and np.array_equal(F.q, R.q)
    assert same_pattern(F.L, R.L) and same_pattern(F.U, R.U)
    assert max_rel(F.L.data, R.L.data) <= 1e-12
    assert max_rel(F.U.data, R.U.data) <= 1e-12
    X = I.solve(F, B)
    for j in (0, 7):
        xr = ON.solve(R, B[:, j])
        assert np.abs(X[:, j] - xr).max() <= 1e-12 * np.abs(xr).max()
    assert _bwd(A, X, B).max() <= 1e-15


def test_poisson48_mixed_refined_vs_fp64(poisson48):
    """FP32 fronts, FP64 residual and update: refinement restores backward error <= 1e-14 on all
    8 RHS; the error against the FP64 solve is reported and bounded by the conditioning."""
    A, S, B = poisson48
    print(f"[case] poisson3d-48 perturbed fp32 + refinement, 8 RHS", flush=True)
    F64 = I.splu(A, method="multifrontal", symbolic=S)
    X64 = I.solve(F64, B)
    del F64
    F32 = I.splu(A, method="multifrontal", symbolic=S, precision="fp32")
    X = I.solve(F32, B, refine=30, tol=1e-14)
    rep = F32.last_report
    be = _bwd(A, X, B)
    err = np.abs(X - X64).max(axis=0) / np.abs(X64).max(axis=0)
    print(f"mixed: {rep['iterations']} refinement steps, device bwd {rep['backward_error']:.2e}, host bwd "
          f"{be.max():.2e}, max rel error vs the FP64 solve {err.max():.2e}", flush=True)
    assert rep["converged"] == 1 and rep["backward_error"] <= 1e-14
    assert be.max() <= 1e-14
    assert err.max() <= 1e-12


def test_batch_512_poisson16_full():
    nmat, k, nrhs = 512, 16, 64
    print(f"[case] batch {nmat} x poisson3d-{k}, {nrhs} RHS, fp32 + refinement, graph", flush=True)
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=4)
    S = I.symbolic_mf(P)
    B = I.BatchedLU(P, nmat, precision="fp32", symbolic=S, use_graph=True)
    info = B.refactor(Ax)
    assert not np.any(info)
    rhs = np.random.default_rng(3).standard_normal((nmat, P.shape[0], nrhs))
    X = B.solve(rhs, refine=30, tol=1e-14)
    assert B.last_report["converged"] == 1
    worst = 0.0
    for m in range(nmat):
        worst = max(worst, float(_bwd(B.matrix(Ax[m]), X[m], rhs[m]).max()))
    print(f"batch: {B.last_report['iterations']} refinement steps, worst host bwd {worst:.2e}", flush=True)
    assert worst <= 1e-14
    # FP64 batch path: factors of a few matrices against the oracle
    pick = [0, 255, 511]
    B64 = I.BatchedLU(P, len(pick), precision="fp64", symbolic=S, use_graph=True)
    B64.refactor(Ax[pick])
    for i, m in enumerate(pick):
        R = oracle_splu(B.matrix(Ax[m]), S)
        L, U, p, q = B64.factors(i)
        assert np.array_equal(p, R.p) and np.array_equal(q, R.q)
        assert same_pattern(L, R.L) and same_pattern(U, R.U)
        assert max_rel(L.data, R.L.data) <= 1e-12 and max_rel(U.data, R.U.data) <= 1e-12


@pytest.fixture(scope="module")
def convdiff64():
    A = M.convdiff3d(64)
    S = I.symbolic_mf(A, ordering="colamd")
    b = np.random.default_rng(5).standard_normal((A.shape[0], 1))
    return A, S, b


def test_convdiff64_colamd_bf16_refined(convdiff64):
    A, S, b = convdiff64
    st = I.plan_estimate(A, S, precision="bf16")
    print(f"[case] convdiff-64 colamd bf16: fill {S.predicted_fill()}, store {st['store_bytes'] / 1e9:.2f} GB, "
          f"arena {st['arena_bytes'] / 1e9:.2f} GB, resident fronts would be {st['resident_front_bytes'] / 1e9:.0f} GB",
          flush=True)
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
    x = I.solve(F, b, refine=60, tol=1e-12)
    rep = F.last_report
    be = _bwd(A, x, b).max()
    print(f"bf16: {rep['iterations']} refinement steps, device bwd {rep['backward_error']:.2e}, host bwd {be:.2e}",
          flush=True)
    assert rep["converged"] == 1 and be <= 1e-12


def test_convdiff64_colamd_fp64_agrees(convdiff64):
    """The FP64 path on the same workload (store 8.7 GB, arena 6.6 GB): the BF16 result's error
    against it."""
    A, S, b = convdiff64
    print("[case] convdiff-64 colamd fp64", flush=True)
    F = I.splu(A, method="multifrontal", symbolic=S)
    x64 = I.solve(F, b, refine=1)
    assert _bwd(A, x64, b).max() <= 1e-15
    del F
    Fb = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
    x = I.solve(Fb, b, refine=60, tol=1e-12)
    err = np.abs(x - x64).max() / np.abs(x64).max()
    print(f"bf16 vs fp64: max rel error {err:.2e}", flush=True)
    assert err <= 1e-9
