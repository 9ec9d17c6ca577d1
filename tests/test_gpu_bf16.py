"""BF16 Schur complements (precision="bf16": FP32 fronts, getrf and trsm; Schur gemm operands
in BF16 with per-tile power-of-two scaling, FP32 accumulation) refined in FP64."""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M

pytestmark = pytest.mark.gpu


def _bwd(A, x, b):
    return np.abs(b - A @ x).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("name,make", [("poisson3d-12", lambda: M.poisson3d(12)),
                                       ("convdiff-12", lambda: M.convdiff3d(12))])
def test_bf16_refines_to_backward_error(name, make):
    A = make()
    print(f"[case] bf16 {name}", flush=True)
    S = I.symbolic_mf(A)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16", small_nf=1)
    x = I.solve(F, b, refine=60, tol=1e-14)
    assert F.last_report["converged"] == 1
    assert _bwd(A, x, b) <= 1e-14
    # the BF16 Schur complements really ran: the factors differ from the FP32 ones, by about
    # BF16's rounding, not by an error
    G = I.splu(A, method="multifrontal", symbolic=S, precision="fp32", small_nf=1)
    d = np.abs(F.U.data - G.U.data).max() / np.abs(G.U.data).max()
    assert 1e-6 < d < 5e-2


def test_bf16_colamd_convdiff_to_1e12():
    """The conv-diff workload shape of the BF16 configuration at a small grid: COLAMD ordering,
    BF16 Schur complements, FP32 getrf/trsm, FP64 refinement to backward error <= 1e-12."""
    A = M.convdiff3d(14)
    S = I.symbolic_mf(A, ordering="colamd")
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    F = I.splu(A, method="multifrontal", symbolic=S, precision="bf16", small_nf=1)
    x = I.solve(F, b, refine=60, tol=1e-12)
    assert F.last_report["converged"] == 1
    assert _bwd(A, x, b) <= 1e-12
