"""mf_solve_device_noalloc (the allocation-free solve, scratch in a file-scope device array)
against mf_solve_device through the C ABI: identical results on two calls in a row, the second
with a smaller input, for an FP64 and an FP32 plan; sizes beyond the scratch are refused."""
import ctypes

import numpy as np
import pytest

from pureumfpack_hip import _lib
from pureumfpack_hip import interface as I
from pureumfpack_hip import matrices as M

from _util import DeviceArrays

pytestmark = pytest.mark.gpu


def _both(lib, hip, plan, n, B, refine, tol):
    nrhs = B.shape[1]
    dB = hip.upload(np.asfortranarray(B))
    dX1 = hip.upload(np.zeros((n, nrhs), order="F"))
    dX2 = hip.upload(np.full((n, nrhs), np.nan, order="F"))
    r1, r2 = _lib.mf_solve_report(), _lib.mf_solve_report()
    assert lib.mf_solve_device(plan.h, nrhs, dB, n, dX1, n, refine, tol, ctypes.byref(r1)) == 0
    rc = lib.mf_solve_device_noalloc(plan.h, nrhs, dB, n, dX2, n, refine, tol, ctypes.byref(r2))
    assert rc == 0, lib.mf_last_error()
    X1 = hip.download(dX1, np.zeros(n * nrhs)).reshape((n, nrhs), order="F")
    X2 = hip.download(dX2, np.zeros(n * nrhs)).reshape((n, nrhs), order="F")
    return X1, X2, r1, r2


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_solve_matches_allocating_solve(precision):
    lib = _lib.load()
    hip = DeviceArrays()
    rng = np.random.default_rng(21)
    try:
        # the second call is smaller (fewer unknowns, one RHS): the shared scratch keeps the
        # first call's contents and must not leak into the result
        for k, nrhs, refine, tol in ((24, 4, 3, 1e-14), (10, 1, 1, 0.0)):
            A = M.poisson3d_perturbed(k, seed=k)
            n = A.shape[0]
            print(f"[case] noalloc {precision} poisson3d-{k} nrhs={nrhs}", flush=True)
            S = I.symbolic_mf(A)
            plan = I._Plan(A, S, precision=precision)
            plan.factor(A.data, I.row_scaling(A))
            B = rng.standard_normal((n, nrhs))
            X1, X2, r1, r2 = _both(lib, hip, plan, n, B, refine, tol)
            assert np.array_equal(X1, X2)
            assert r1.iterations == r2.iterations and r1.backward_error == r2.backward_error
            # the first call refines to 1e-14; the second takes one fixed step (FP32: ~1e-13)
            assert r2.backward_error <= (1e-14 if tol > 0 else 1e-12)
            del plan
    finally:
        hip.free()


def test_noalloc_refuses_sizes_beyond_the_scratch():
    lib = _lib.load()
    hip = DeviceArrays()
    try:
        A = M.poisson3d(24)
        n = A.shape[0]
        plan = I._Plan(A, I.symbolic_mf(A))
        plan.factor(A.data, I.row_scaling(A))
        nrhs = 512
        dB = hip.upload(np.zeros((n, nrhs), order="F"))
        dX = hip.upload(np.zeros((n, nrhs), order="F"))
        rep = _lib.mf_solve_report()
        rc = lib.mf_solve_device_noalloc(plan.h, nrhs, dB, n, dX, n, 0, 0.0, ctypes.byref(rep))
        assert rc == _lib.MF_BADARG and b"bytes of scratch" in lib.mf_last_error()
    finally:
        hip.free()
