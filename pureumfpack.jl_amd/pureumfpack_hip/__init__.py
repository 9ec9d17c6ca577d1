"""pureumfpack_hip -- MI355X numeric phase of PureUMFPACK.jl's supernodal multifrontal LU.

The Python side mirrors the reference's multifrontal interface (splu, multifrontal_lu,
solve, row_scaling, symbolic_mf); all numeric work runs in libmfgpu.so (HIP, gfx950)
through its C ABI (include/mfgpu.h).
"""
from .interface import (SCALE_MAX, SCALE_NONE, SCALE_SUM, ArgumentError, BatchedLU, DeviceError, DimensionMismatch,
                        MFFactorization, PureLU, SingularException, SymbolicMF, ldiv, multifrontal_lu,
                        row_scaling, solve, splu, symbolic_mf)
from . import matrices

__all__ = ["SCALE_MAX", "SCALE_NONE", "SCALE_SUM", "ArgumentError", "BatchedLU", "DeviceError", "DimensionMismatch",
           "MFFactorization", "PureLU", "SingularException", "SymbolicMF", "ldiv", "multifrontal_lu",
           "row_scaling", "solve", "splu", "symbolic_mf", "matrices"]
