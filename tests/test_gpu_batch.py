"""Same-pattern batches on the GPU (mf_refactor_batch / mf_solve_batch through BatchedLU).

Every matrix of a batch must come out exactly as a single-matrix factorization would: the
FP64 batch is checked matrix by matrix against the oracle (identical p, q and L/U patterns,
values within 1e-12), the mixed-precision batch against the backward-error target, and a
captured-graph replay against the eager launches.
"""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M
from oracle import scaling as OS

from _util import max_rel, oracle_splu, same_pattern

pytestmark = pytest.mark.gpu


def _bwd_err(A, x, b):
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    r = b - A @ x
    return np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("k,nmat", [(6, 5), (8, 3)])
def test_batch_fp64_parity(k, nmat):
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=11)
    print(f"[case] batch fp64 k={k} nmat={nmat}", flush=True)
    B = I.BatchedLU(P, nmat, precision="fp64", use_graph=False)
    info = B.refactor(Ax)
    assert np.all(info == 0)
    for m in range(nmat):
        A = B.matrix(Ax[m])
        L, U, p, q = B.factors(m)
        R = oracle_splu(A, B.S)
        assert np.array_equal(p, R.p) and np.array_equal(q, R.q)
        assert same_pattern(L, R.L) and same_pattern(U, R.U)
        assert max_rel(L.data, R.L.data) <= 1e-12
        assert max_rel(U.data, R.U.data) <= 1e-12
        As = OS.apply_row_scaling(A, B.Rs[m])
        rec = abs(As[p, :][:, q] - L @ U).max() / abs(As).max()
        assert rec <= 1e-13


def test_batch_mixed_refine_to_bwd_tol():
    """FP32 fronts + FP64 refinement, several right-hand sides per matrix."""
    k, nmat, nrhs = 10, 6, 4
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=5)
    n = P.shape[0]
    rng = np.random.default_rng(9)
    Bv = rng.standard_normal((nmat, n, nrhs))
    B = I.BatchedLU(P, nmat, precision="fp32")
    B.refactor(Ax)
    X = B.solve(Bv, refine=30, tol=1e-14)
    assert X.shape == Bv.shape
    for m in range(nmat):
        A = B.matrix(Ax[m])
        for j in range(nrhs):
            assert _bwd_err(A, X[m][:, j], Bv[m][:, j]) <= 1e-14
    assert B.last_report["backward_error"] <= 1e-14


def test_batch_many_rhs_separate_wave_and_workgroup_launches():
    """Enough matrices x right-hand sides that levels with one-wave and workgroup fronts run
    them in separate launches (the one-wave items are too many to share a launch); every
    matrix's FP64 solve must still be accurate."""
    k, nmat, nrhs = 14, 64, 64   # 14^3: a level of 50 one-wave and 2 workgroup fronts
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=41)
    n = P.shape[0]
    print(f"[case] batch fp64 k={k} nmat={nmat} nrhs={nrhs}", flush=True)
    Bv = np.random.default_rng(8).standard_normal((nmat, n, nrhs))
    B = I.BatchedLU(P, nmat, precision="fp64", use_graph=False)
    B.refactor(Ax)
    X = B.solve(Bv, refine=0)
    for m in range(0, nmat, 7):
        A = B.matrix(Ax[m])
        for j in (0, nrhs - 1):
            assert _bwd_err(A, X[m][:, j], Bv[m][:, j]) <= 1e-13


def test_batch_matches_single_matrix():
    """Matrix m of a batch solves exactly like a one-matrix plan on A_m (same schedule)."""
    k, nmat = 8, 4
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=21)
    n = P.shape[0]
    b = np.random.default_rng(3).standard_normal((nmat, n))
    B = I.BatchedLU(P, nmat, precision="fp64", use_graph=False)
    B.refactor(Ax)
    X = B.solve(b, refine=1)
    for m in range(nmat):
        A = B.matrix(Ax[m])
        F = I.splu(A, method="multifrontal", symbolic=B.S)
        x = I.solve(F, b[m], refine=1)
        assert np.abs(X[m] - x).max() <= 1e-13 * np.abs(x).max()


def test_batch_graph_replay_matches_eager():
    """hipGraph-captured refactorization, replayed for fresh values, equals eager launches."""
    k, nmat = 8, 4
    P, Ax1 = M.poisson3d_batch_values(k, nmat, seed=31)
    _, Ax2 = M.poisson3d_batch_values(k, nmat, seed=32)
    n = P.shape[0]
    b = np.random.default_rng(4).standard_normal((nmat, n, 2))
    G = I.BatchedLU(P, nmat, precision="fp32", use_graph=True)
    E = I.BatchedLU(P, nmat, precision="fp32", use_graph=False, symbolic=G.S)
    for Ax in (Ax1, Ax2, Ax1):
        G.refactor(Ax)
        E.refactor(Ax)
        Xg = G.solve(b, refine=2)
        Xe = E.solve(b, refine=2)
        assert np.array_equal(Xg, Xe)
        for m in range(nmat):
            Lg, Ug, _, _ = G.factors(m)
            Le, Ue, _, _ = E.factors(m)
            assert np.array_equal(Lg.data, Le.data) and np.array_equal(Ug.data, Ue.data)


def test_batch_smaller_than_plan():
    """A plan sized for nbatch matrices also refactors fewer."""
    P, Ax = M.poisson3d_batch_values(6, 4, seed=2)
    B = I.BatchedLU(P, 4, precision="fp64", use_graph=False)
    B.refactor(Ax[:2])
    x = B.solve(np.ones((2, P.shape[0])))
    for m in range(2):
        A = B.matrix(Ax[m])
        assert _bwd_err(A, x[m], np.ones(P.shape[0])) <= 1e-14
    with pytest.raises(I.ArgumentError):
        B.refactor(np.vstack([Ax, Ax[:1]]))


def test_batch_device_entry_points_match_host():
    """mf_refactor_batch_device / mf_solve_batch_device (inputs already in HBM) give the host
    entry points' results bit for bit."""
    import ctypes
    from pureumfpack_hip import _lib
    from _util import DeviceArrays

    k, nmat, nrhs = 6, 3, 2
    P, Ax = M.poisson3d_batch_values(k, nmat, seed=41)
    n = P.shape[0]
    B = I.BatchedLU(P, nmat, precision="fp32", use_graph=False)
    rhs = np.random.default_rng(8).standard_normal((nmat, n, nrhs))
    B.refactor(Ax)
    Xh = B.solve(rhs, refine=3)
    lib = _lib.load()
    dev = DeviceArrays()
    try:
        rt = np.ascontiguousarray(np.transpose(rhs, (0, 2, 1)))   # per matrix n x nrhs, column-major
        dAx, dRs, dB = dev.upload(Ax), dev.upload(B.Rs), dev.upload(rt)
        dX = dev.upload(np.zeros_like(rt))
        info = np.zeros(nmat, np.int64)
        h = B._plan.h
        assert lib.mf_refactor_batch_device(h, nmat, dAx, dRs, info.ctypes.data_as(_lib.c_i64p)) == 0
        rep = _lib.mf_solve_report()
        assert lib.mf_solve_batch_device(h, nmat, nrhs, dB, dX, 3, 0.0, ctypes.byref(rep)) == 0
        Xd = np.transpose(dev.download(dX, rt), (0, 2, 1))
    finally:
        dev.free()
    assert np.array_equal(Xd, Xh)
