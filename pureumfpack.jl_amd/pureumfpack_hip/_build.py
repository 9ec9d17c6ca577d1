"""Build of the C-ABI HIP library libmfgpu.so (gfx950) from pureumfpack.jl_amd/csrc.

The library is built in-tree next to this package so that the same file is what the
tests, the smoke check and the benchmark load.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                  # pureumfpack.jl_amd/
CSRC = os.path.join(ROOT, "csrc")
LIB_PATH = os.path.join(ROOT, "libmfgpu.so")
SOURCES = ["mf_analysis.cpp", "mf_plan.hip", "mf_factor_kernels.hip", "mf_solve_kernels.hip", "mf_mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # the 64-step register-resident substitutions are fully unrolled (constant register
         # indices); the default pragma threshold stops short of that
         "-mllvm", "-pragma-unroll-threshold=200000",
         "-Wall", "-Wno-unused-result"]
# per-source extra flags
EXTRA = {}


def build(verbose: bool = False) -> str:
    """Compile every source to an object, then link libmfgpu.so. Returns the library path."""
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC, *FLAGS, *EXTRA.get(src, []), *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the sources compile independently: run them side by side
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        procs.append(subprocess.Popen(cmd))
    failed = [c for c, p in zip(cmds, procs) if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB_PATH + ".tmp"
    subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
