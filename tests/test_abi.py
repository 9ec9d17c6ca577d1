"""The C-ABI library loads and exports every function include/mfgpu.h declares (no GPU call)."""
import ctypes
import os
import re

from pureumfpack_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?[\w ]+?\*?\s*\b(mf_\w+)\s*\(", txt, flags=re.M))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 15
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_declarations():
    assert _declared() <= set(_lib.SIGNATURES)


def test_last_error_without_gpu():
    lib = _lib.load()
    h = ctypes.c_void_p()
    import numpy as np
    Ap = np.array([1, 2], dtype=np.int64)
    Ai = np.array([1], dtype=np.int64)
    rc = lib.mf_analyze(1, _lib.ptr(Ap), _lib.ptr(Ai), 7, None, ctypes.byref(h))
    assert rc == _lib.MF_BADARG and b"ordering" in lib.mf_last_error()
