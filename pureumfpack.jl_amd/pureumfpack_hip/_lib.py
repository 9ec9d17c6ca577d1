"""ctypes binding of libmfgpu.so (the C ABI declared in include/mfgpu.h).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB_PATH

MF_OK, MF_SINGULAR, MF_DIM, MF_BADARG, MF_HIP, MF_OOM = range(6)
MF_FP64, MF_FP32, MF_BF16 = 0, 1, 2
MF_ORDER_AMD, MF_ORDER_COLAMD, MF_ORDER_GIVEN = 0, 1, 2
MF_PIVOT_AUTO, MF_PIVOT_PARTIAL = 0, 1
MF_SOLVE_AUTO, MF_SOLVE_FRONTS, MF_SOLVE_COLUMNS = 0, 1, 2

c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_f32p = ctypes.POINTER(ctypes.c_float)
c_void_pp = ctypes.POINTER(ctypes.c_void_p)


class mf_options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("small_nf", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("nbatch", ctypes.c_int32), ("pivoting", ctypes.c_int32), ("pivot_tol", ctypes.c_float),
                ("solve_path", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class mf_solve_report(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32), ("backward_error", ctypes.c_double),
                ("residual_2", ctypes.c_double), ("seconds", ctypes.c_double)]


class mf_plan_stats(ctypes.Structure):
    _fields_ = [("last_schedule", ctypes.c_int64), ("factor_flops", ctypes.c_double), ("gemm_flops", ctypes.c_double),
                ("gemm_flops_big", ctypes.c_double), ("front_bytes", ctypes.c_double),
                ("nsuper", ctypes.c_int64), ("nlevels", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("max_nf", ctypes.c_int64), ("store_bytes", ctypes.c_double), ("arena_bytes", ctypes.c_double),
                ("resident_front_bytes", ctypes.c_double), ("last_redone", ctypes.c_int64)]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "mf_analyze": (ctypes.c_int, [ctypes.c_int64, c_i64p, c_i64p, ctypes.c_int32, c_i64p, c_void_pp]),
    "mf_symbolic_sizes": (ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p, c_i64p, c_i64p]),
    "mf_symbolic_export": (ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p]),
    "mf_symbolic_relax": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, c_i64p]),
    "mf_symbolic_destroy": (None, [ctypes.c_void_p]),
    "mf_plan_create": (ctypes.c_int, [ctypes.c_int64, c_i64p, ctypes.c_int64, c_i64p, c_i64p, c_i64p, c_i64p,
                                      c_i64p, c_i64p, c_i64p, ctypes.POINTER(mf_options), c_void_pp]),
    "mf_factor": (ctypes.c_int, [ctypes.c_void_p, c_f64p, c_f64p, ctypes.c_int32, c_i64p]),
    "mf_factor_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, c_i64p]),
    "mf_refactor_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f64p, c_f64p, c_i64p]),
    "mf_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_f64p, ctypes.c_int64, c_f64p, ctypes.c_int64,
                                ctypes.c_int32, ctypes.c_double, ctypes.POINTER(mf_solve_report)]),
    "mf_solve_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                       ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                       ctypes.POINTER(mf_solve_report)]),
    "mf_solve_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                               ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double,
                                               ctypes.POINTER(mf_solve_report)]),
    "mf_solve_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, c_f64p, c_f64p,
                                      ctypes.c_int32, ctypes.c_double, ctypes.POINTER(mf_solve_report)]),
    "mf_solve_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                             ctypes.POINTER(mf_solve_report)]),
    "mf_refactor_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                                c_i64p]),
    "mf_nnz": (ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p]),
    "mf_export_factors": (ctypes.c_int, [ctypes.c_void_p, c_i64p, c_i64p, c_f64p, c_i64p, c_i64p, c_f64p, c_i64p,
                                         c_i64p]),
    "mf_export_factors_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, c_i64p, c_i64p, c_f64p, c_i64p, c_i64p, c_f64p, c_i64p,
                                         c_i64p]),
    "mf_set_gemm_tiles": (ctypes.c_int, [ctypes.c_int32, c_i32p]),
    "mf_get_gemm_tiles": (ctypes.c_int, [ctypes.c_int32, c_i32p]),
    "mf_gemm_tile_info": (ctypes.c_int, [c_i32p, c_i32p, c_i32p, c_i32p, c_i32p]),
    "mf_time_gemm": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int32, ctypes.c_int32, c_f64p]),
    "mf_time_gemm_list": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, c_i32p, ctypes.c_int32, ctypes.c_int32,
                                         c_f64p]),
    "mf_mlp_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "mf_mlp_destroy": (None, [ctypes.c_void_p]),
    "mf_mlp_num_params": (ctypes.c_int64, [ctypes.c_void_p]),
    "mf_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, c_f32p]),
    "mf_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, c_f32p]),
    "mf_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, c_f32p, ctypes.c_int64, c_f32p]),
    "mf_mlp_gradient": (ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p, ctypes.c_int64, c_f32p, c_f32p]),
    "mf_mlp_train": (ctypes.c_int, [ctypes.c_void_p, c_f32p, c_f32p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float,
                                    c_f32p]),
    "mf_mlp_last_error": (ctypes.c_char_p, []),
    "mf_copy_probe": (ctypes.c_int, [ctypes.c_int64, c_f64p]),
    "mf_plan_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(mf_plan_stats)]),
    "mf_plan_estimate": (ctypes.c_int, [ctypes.c_int64, c_i64p, ctypes.c_int64, c_i64p, c_i64p, c_i64p, c_i64p,
                                        c_i64p, c_i64p, c_i64p, ctypes.POINTER(mf_options),
                                        ctypes.POINTER(mf_plan_stats)]),
    "mf_plan_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mf_plan_get_profile": (ctypes.c_int, [ctypes.c_void_p, c_f64p, c_i64p, c_f64p, c_f64p]),
    "mf_last_error": (ctypes.c_char_p, []),
    "mf_plan_destroy": (None, [ctypes.c_void_p]),
}

_LIB = None


def load():
    """Load libmfgpu.so built in-tree; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libmfgpu.so not built ({LIB_PATH}); run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


def ptr(a, kind=c_i64p):
    return a.ctypes.data_as(kind) if a is not None else None
