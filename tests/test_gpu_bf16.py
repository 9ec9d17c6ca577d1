"""BF16 Schur complements (precision="bf16": FP32 fronts, getrf and trsm; the Schur gemms of the
fronts with nf >= 1024 with operands packed in BF16 tiles, each with its own power-of-two
scale, FP32 accumulation; smaller fronts' Schur gemms in FP32) refined in FP64."""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M

pytestmark = pytest.mark.gpu


def _bwd(A, x, b):
    return np.abs(b - A @ x).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


def _big_fronts(S):
    nf = np.diff(np.asarray(S.sstart)) + np.diff(np.asarray(S.upd_ptr))
    return int((nf >= 1024).sum())


@pytest.mark.parametrize("name,make,ordering", [("poisson3d-30", lambda: M.poisson3d(30), "amd"),
                                                ("convdiff-24", lambda: M.convdiff3d(24), "colamd")])
def test_bf16_refines_to_backward_error(name, make, ordering):
    A = make()
    S = I.symbolic_mf(A, ordering=ordering)
    print(f"[case] bf16 {name}: {_big_fronts(S)} fronts with nf >= 1024", flush=True)
    assert _big_fronts(S) > 0
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
    x = I.solve(F, b, refine=60, tol=1e-14)
    assert F.last_report["converged"] == 1
    assert _bwd(A, x, b) <= 1e-14
    # the BF16 Schur complements really ran: the factors differ from the FP32 ones by about
    # BF16's rounding (2^-9 relative to each tile's largest entry), not by an error
    G = I.splu(A, method="multifrontal", symbolic=S, precision="fp32")
    d = np.abs(F.U.data - G.U.data).max() / np.abs(G.U.data).max()
    assert 1e-6 < d < 5e-2


def test_bf16_matches_fp32_where_no_front_is_large():
    """Below nf = 1024 every Schur gemm runs in FP32, so the BF16 plan's factors are the FP32
    plan's, bit for bit."""
    A = M.convdiff3d(12)
    S = I.symbolic_mf(A, ordering="colamd")
    assert _big_fronts(S) == 0
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
    G = I.splu(A, method="multifrontal", symbolic=S, precision="fp32")
    assert np.array_equal(F.U.data, G.U.data) and np.array_equal(F.L.data, G.L.data)


def test_bf16_colamd_convdiff_to_1e12():
    """The conv-diff workload shape of the BF16 configuration at a small grid: COLAMD ordering,
    BF16 Schur complements, FP32 getrf/trsm, FP64 refinement to backward error <= 1e-12, on the
    reference's supernodes and on the relaxed amalgamation."""
    A = M.convdiff3d(24)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    for relax in (None, True):
        S = I.symbolic_mf(A, ordering="colamd", relax=relax)
        F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
        x = I.solve(F, b, refine=60, tol=1e-12)
        assert F.last_report["converged"] == 1
        assert _bwd(A, x, b) <= 1e-12


def test_bf16_tiles_with_tiny_entries():
    """Operand tiles whose largest magnitude is an FP32 subnormal keep a finite scale: a front whose
    Schur operands mix O(1) entries with 1e-40 ones factors without Inf or NaN."""
    A = M.convdiff3d(24).tolil()
    n = A.shape[0]
    rng = np.random.default_rng(2)
    for i in rng.choice(n, 200, replace=False):
        j = (i + 577) % n
        A[i, j] = 1e-40
        A[j, i] = 1e-40
    A = A.tocsc()
    S = I.symbolic_mf(A, ordering="colamd")
    b = rng.standard_normal(n)
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16")
    x = I.solve(F, b, refine=60, tol=1e-12)
    assert np.isfinite(x).all() and F.last_report["converged"] == 1
