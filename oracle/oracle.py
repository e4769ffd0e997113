"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker: only tests/ and __graft_entry__.smoke() import it.
The product libraries never load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for name in ("oracle_dslash_f64", "oracle_dslash_3d_f64", "oracle_dslash_f32", "oracle_dslash_3d_f32"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_void_p]
        L.oracle_coords.restype = None
        L.oracle_coords.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_long, C.POINTER(C.c_int)]
        L.oracle_index.restype = C.c_long
        L.oracle_index.argtypes = [C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]
        L.oracle_set_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def dslash(u: np.ndarray, psi: np.ndarray, res: np.ndarray, latt, isign: int, cb: int, nd: int = 4,
           site_table: np.ndarray | None = None) -> np.ndarray:
    """sse_su3dslash_wilson semantics on host arrays (float32 or float64, same dtype)."""
    dt = psi.dtype
    if u.dtype != dt or res.dtype != dt or dt not in (np.float32, np.float64):
        raise TypeError("u, psi, res must share dtype float32 or float64")
    for a in (u, psi, res):
        if not a.flags.c_contiguous:
            raise ValueError("arrays must be C-contiguous")
    V = int(np.prod(latt))
    if u.size != V * 72 or psi.size != V * 24 or res.size != V * 24:
        raise ValueError("array sizes do not match the lattice")
    L = (C.c_int * 4)(*[int(v) for v in latt])
    tab = None
    if site_table is not None:
        site_table = np.ascontiguousarray(site_table, dtype=np.int32)
        tab = site_table.ctypes.data
    name = "oracle_dslash_" + ("3d_" if nd == 3 else "") + ("f64" if dt == np.float64 else "f32")
    getattr(lib(), name)(u.ctypes.data, psi.ctypes.data, res.ctypes.data, L, int(isign), int(cb), tab)
    return res
