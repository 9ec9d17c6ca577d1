"""The device plan's storage layout, computed on the host by mf_plan_estimate (no GPU call).

The factor store holds exactly what the solves and the export read: per supernode the panel
[L11\\U11; L21] (nf x np) and U12 (np x nu), i.e. nnz(L) + nnz(U) - n values, plus the
inverses of pivot blocks with np > 128.  The workspace arena holds each front only from its
etree level through its parent's level (src/multifrontal.jl:180-187,319-332 keep a LIFO arena
for the same reason), so it needs about the peak of the fronts alive at one level, not the
sum over all fronts; the plan checks by itself that blocks alive together never overlap."""
import numpy as np
import pytest

from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M


def _live_peak(S, small_nf=64):
    """Max over levels of the elements alive at that level: nf^2 for a front factored by the
    large-front kernels, nu^2 for one factored in LDS, alive from its level through its
    parent's level."""
    ns = S.nsuper
    npv = np.diff(np.asarray(S.sstart))
    nu = np.diff(np.asarray(S.upd_ptr))
    nf = npv + nu
    cp, ci = np.asarray(S.child_ptr), np.asarray(S.child_idx)
    parent = np.full(ns, -1)
    level = np.zeros(ns, np.int64)
    for s in range(ns):
        for c in ci[cp[s]:cp[s + 1]]:
            parent[c] = s
            level[s] = max(level[s], level[c] + 1)
    end = np.where(parent >= 0, level[np.maximum(parent, 0)], level)
    need = np.where(nf > small_nf, nf.astype(float) ** 2, nu.astype(float) ** 2)
    diff = np.zeros(level.max() + 2)
    np.add.at(diff, level, need)
    np.add.at(diff, end + 1, -need)
    return np.cumsum(diff).max(), float((nf.astype(float) ** 2).sum()), npv


@pytest.mark.parametrize("name,make,ordering,prec,esize", [
    ("poisson3d-16", lambda: M.poisson3d(16), "amd", "fp64", 8),
    ("poisson3d-48-perturbed", lambda: M.poisson3d_perturbed(48), "amd", "fp32", 4),
    ("convdiff-24", lambda: M.convdiff3d(24), "colamd", "bf16", 4),
])
def test_store_and_arena_sizes(name, make, ordering, prec, esize):
    A = make()
    S = I.symbolic_mf(A, ordering=ordering)
    st = I.plan_estimate(A, S, precision=prec)
    peak, resident, npv = _live_peak(S)
    factor = S.predicted_fill() - A.shape[0]
    inv = float((2 * npv[npv > 128].astype(float) ** 2).sum())
    store = st["store_bytes"] / esize
    assert factor + inv <= store <= factor + 2 * inv
    arena = st["arena_bytes"] / esize
    assert peak <= arena <= 1.5 * peak + 1e6
    assert st["resident_front_bytes"] / esize == resident
    assert arena < resident


@pytest.mark.slow
def test_convdiff64_colamd_fits_one_gpu():
    """The 64^3 convection-diffusion workload under COLAMD (fill 1.09e9, max front 17106):
    every front resident would take 616 GB in FP32; the store plus the planned arena fit in a
    few GB of the 288 GB of HBM."""
    A = M.convdiff3d(64)
    S = I.symbolic_mf(A, ordering="colamd")
    st = I.plan_estimate(A, S, precision="bf16")
    assert st["max_nf"] == 17106 and S.predicted_fill() == 1089096332
    assert st["resident_front_bytes"] > 600e9
    assert st["store_bytes"] + st["arena_bytes"] < 10e9


def test_batch_scales_both_areas():
    A = M.poisson3d(8)
    S = I.symbolic_mf(A)
    one = I.plan_estimate(A, S, precision="fp32")
    four = I.plan_estimate(A, S, precision="fp32", nbatch=4)
    assert four["store_bytes"] == 4 * one["store_bytes"] and four["arena_bytes"] == 4 * one["arena_bytes"]


def test_estimate_argument_errors():
    A = M.poisson3d(4)
    S = I.symbolic_mf(A)
    with pytest.raises(I.ArgumentError):
        I.plan_estimate(A, S, precision="fp16")
    bad = I.SymbolicMF(S.n, S.qf.copy(), S.parent, S.sstart, S.upd_ptr, S.upd_idx, S.child_ptr, S.child_idx,
                       S.ncolstruct)
    bad.qf[0] = bad.qf[1]
    with pytest.raises(I.ArgumentError):
        I.plan_estimate(A, bad)
