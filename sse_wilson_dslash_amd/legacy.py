"""Python mirror of the reference's drop-in C API (include/sse_dslash.h,
include/sse_dslash_3d.h), bound by ctypes to libsse_dslash_f32.so /
libsse_dslash_f64.so (the MI355X libraries) or, with backend="cpu", to
libsse_dslash_cpu_f32.so / libsse_dslash_cpu_f64.so (the same API computed on
the host CPU with OpenMP; chosen explicitly, never as a fallback).

Same names, argument meaning and error behaviour as the reference:
``init_sse_su3dslash(latt_size, getSiteCoords, getLinearSiteIndex, nodeNumber)``
builds the site table from the caller's callbacks (refcounted; an odd
dimension prints a message and exits the process with status 1),
``sse_su3dslash_wilson(u, psi, res, isign, cb)`` overwrites the target
checkerboard 1-cb of ``res`` with D psi (isign=+1) or D^dagger psi (isign=-1)
and does nothing for other isign.  Arrays are numpy, full lattice, in the
caller's site order: u is [V][4][3][3][2] transposed links, psi/res [V][4][3][2].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

SITE_COORDS = C.CFUNCTYPE(None, C.POINTER(C.c_int), C.c_int, C.c_int)
LINEAR_INDEX = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_int))
NODE_NUMBER = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_int))

_vp, _i = C.c_void_p, C.c_int


def canonical_callbacks(latt, nd: int = 4):
    """The cb2 (nd=4) / cb3d (nd=3) layout callbacks exported by the core library,
    i.e. what QDP++'s Layout provides; returns C function pointers."""
    core = _lib.core()
    # An invalid size is left to init_sse_su3dslash, which reports it and exits
    # before any callback is used.
    core.dslash_layout_set(_lib.latt_arg(latt), nd)
    return (C.cast(core.dslash_layout_site_coords, C.c_void_p).value,
            C.cast(core.dslash_layout_linear_index, C.c_void_p).value,
            C.cast(core.dslash_layout_node_number, C.c_void_p).value)


def table_callbacks(latt, site_table: np.ndarray, nd: int = 4):
    """Python callbacks for a caller order given as site_table[my canonical] = caller index."""
    from .layout import coords_of_full, full_of_coords
    tab = np.asarray(site_table, dtype=np.int64)
    inv = np.empty_like(tab)
    inv[tab] = np.arange(tab.size)
    coords = coords_of_full(latt, nd, inv)  # coords of caller index q

    def get_site_coords(coord, node, linear):
        c = coords[linear]
        for k in range(4):
            coord[k] = int(c[k])

    def get_linear(coord):
        c = np.array([coord[0], coord[1], coord[2], coord[3]], dtype=np.int64)
        return int(tab[int(full_of_coords(latt, nd, c))])

    def node_number(coord):
        return 0

    return SITE_COORDS(get_site_coords), LINEAR_INDEX(get_linear), NODE_NUMBER(node_number)


class SSEDslash:
    """One precision of the drop-in library (precision = 32 or 64)."""

    def __init__(self, precision: int = 64, backend: str = "gpu"):
        if precision not in (32, 64):
            raise ValueError("precision must be 32 or 64")
        if backend not in ("gpu", "cpu"):
            raise ValueError("backend must be 'gpu' or 'cpu'")
        self.backend = backend
        if backend == "gpu":
            _lib.core()  # the legacy library resolves the core library from the same directory
        self.precision = precision
        self.dtype = np.float32 if precision == 32 else np.float64
        name = (_lib.LEGACY if backend == "gpu" else _lib.LEGACY_CPU)[precision]
        self.lib = C.CDLL(_lib.lib_path(name), mode=C.RTLD_LOCAL)
        L = self.lib
        for name in ("init_sse_su3dslash", "init_sse_su3dslash_3d"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [C.POINTER(C.c_int), _vp, _vp, _vp]
        for name in ("free_sse_su3dslash", "free_sse_su3dslash_3d", "sse_su3dslash_prepost_receives"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = []
        for name in ("sse_su3dslash_wilson", "sse_su3dslash_wilson_3d"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [_vp, _vp, _vp, _i, _i]
        for name in ("sse_su3dslash_last_error", "sse_su3dslash_3d_last_error",
                     "sse_su3dslash_gauge_release", "sse_su3dslash_3d_gauge_release"):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = []
        for name in ("sse_su3dslash_set_sloppy", "sse_su3dslash_3d_set_sloppy", "sse_su3dslash_set_gauge_reuse",
                     "sse_su3dslash_3d_set_gauge_reuse"):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = [_i]
        for name in ("sse_su3dslash_gauge_upload", "sse_su3dslash_3d_gauge_upload"):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = [_vp, _i]
        for name in ("sse_su3dslash_wilson_resident", "sse_su3dslash_wilson_3d_resident"):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = [_vp, _vp, _i, _i]
        for name in ("sse_su3dslash_wilson_device", "sse_su3dslash_wilson_3d_device"):
            getattr(L, name).restype = _i
            getattr(L, name).argtypes = [_vp, _vp, _i, _i, _vp]
        for name in ("pack_gauge_field", "unpack_gauge_field", "unpack_gauge_field_inverse"):
            getattr(L, name).restype = None
            getattr(L, name).argtypes = [_i, _vp, _vp]
        self._keep = {}
        self._ref = {4: 0, 3: 0}
        self._V = {4: 0, 3: 0}

    # ---- helpers -------------------------------------------------------
    def _arr(self, a: np.ndarray, n_reals: int, name: str) -> int:
        if not isinstance(a, np.ndarray) or a.dtype != self.dtype or not a.flags.c_contiguous:
            raise TypeError(f"{name} must be a C-contiguous {np.dtype(self.dtype).name} numpy array")
        if a.size != n_reals:
            raise ValueError(f"{name} must hold {n_reals} reals, got {a.size}")
        return a.ctypes.data

    def _init(self, fn, key, latt_size, getSiteCoords, getLinearSiteIndex, nodeNumber, nd):
        latt = (C.c_int * 4)(*[int(v) for v in latt_size])
        if getLinearSiteIndex is None:
            cbs = canonical_callbacks(latt_size, nd)
        else:
            cbs = tuple(C.cast(f, C.c_void_p).value if f is not None else None
                        for f in (getSiteCoords, getLinearSiteIndex, nodeNumber))
            self._keep[key] = (getSiteCoords, getLinearSiteIndex, nodeNumber)
        fn(latt, cbs[0], cbs[1], cbs[2])
        if self._ref[nd] == 0:  # a second init keeps the first size, as the library does
            self._V[nd] = int(np.prod(latt_size))
        self._ref[nd] += 1

    # ---- 4D ------------------------------------------------------------
    def init_sse_su3dslash(self, latt_size, getSiteCoords=None, getLinearSiteIndex=None, nodeNumber=None):
        self._init(self.lib.init_sse_su3dslash, 4, latt_size, getSiteCoords, getLinearSiteIndex, nodeNumber, 4)

    def free_sse_su3dslash(self):
        self.lib.free_sse_su3dslash()
        self._ref[4] = max(0, self._ref[4] - 1)

    def sse_su3dslash_prepost_receives(self):
        self.lib.sse_su3dslash_prepost_receives()

    def sse_su3dslash_wilson(self, u, psi, res, isign, cb):
        V = self._V[4]
        self.lib.sse_su3dslash_wilson(self._arr(u, V * 72, "u"), self._arr(psi, V * 24, "psi"),
                                      self._arr(res, V * 24, "res"), int(isign), int(cb))

    def last_error(self) -> int:
        return self.lib.sse_su3dslash_last_error()

    def gauge_upload(self, u, gauge_reals: int = 18) -> int:
        return self.lib.sse_su3dslash_gauge_upload(self._arr(u, self._V[4] * 72, "u"), int(gauge_reals))

    def gauge_release(self) -> int:
        return self.lib.sse_su3dslash_gauge_release()

    def wilson_resident(self, psi, res, isign, cb) -> int:
        V = self._V[4]
        return self.lib.sse_su3dslash_wilson_resident(self._arr(psi, V * 24, "psi"), self._arr(res, V * 24, "res"),
                                                      int(isign), int(cb))

    def wilson_device(self, d_res, d_psi, isign, cb, stream=None) -> int:
        return self.lib.sse_su3dslash_wilson_device(_lib.ptr(d_res), _lib.ptr(d_psi), int(isign), int(cb),
                                                    stream)

    def set_sloppy(self, on: bool = True) -> int:
        """FP64 library: FP32 links and arithmetic behind the FP64 interface."""
        return self.lib.sse_su3dslash_set_sloppy(int(bool(on)))

    def set_sloppy_3d(self, on: bool = True) -> int:
        return self.lib.sse_su3dslash_3d_set_sloppy(int(bool(on)))

    def set_gauge_reuse(self, mode: int) -> int:
        """0: upload the gauge every call; 1: reuse while pointer and checksum
        match; 2: reuse while the pointer matches."""
        return self.lib.sse_su3dslash_set_gauge_reuse(int(mode))

    def set_gauge_reuse_3d(self, mode: int) -> int:
        return self.lib.sse_su3dslash_3d_set_gauge_reuse(int(mode))

    # ---- 3D ------------------------------------------------------------
    def init_sse_su3dslash_3d(self, latt_size, getSiteCoords=None, getLinearSiteIndex=None, nodeNumber=None):
        self._init(self.lib.init_sse_su3dslash_3d, 3, latt_size, getSiteCoords, getLinearSiteIndex, nodeNumber, 3)

    def free_sse_su3dslash_3d(self):
        self.lib.free_sse_su3dslash_3d()
        self._ref[3] = max(0, self._ref[3] - 1)

    def sse_su3dslash_wilson_3d(self, u, psi, res, isign, cb):
        V = self._V[3]
        self.lib.sse_su3dslash_wilson_3d(self._arr(u, V * 72, "u"), self._arr(psi, V * 24, "psi"),
                                         self._arr(res, V * 24, "res"), int(isign), int(cb))

    def last_error_3d(self) -> int:
        return self.lib.sse_su3dslash_3d_last_error()

    def gauge_upload_3d(self, u, gauge_reals: int = 18) -> int:
        return self.lib.sse_su3dslash_3d_gauge_upload(self._arr(u, self._V[3] * 72, "u"), int(gauge_reals))

    def wilson_3d_resident(self, psi, res, isign, cb) -> int:
        V = self._V[3]
        return self.lib.sse_su3dslash_wilson_3d_resident(self._arr(psi, V * 24, "psi"),
                                                         self._arr(res, V * 24, "res"), int(isign), int(cb))

    # ---- C packer (lib/packer_nopad.c:29-69) ---------------------------
    def pack_gauge_field(self, volume: int, u: np.ndarray, u_tmp: np.ndarray) -> None:
        n = 2 * volume * 4 * 18
        self.lib.pack_gauge_field(int(volume), self._arr(u, n, "u"), self._arr(u_tmp, n, "u_tmp"))

    def unpack_gauge_field(self, volume: int, u_tmp: np.ndarray, u: np.ndarray) -> None:
        """The reference's unpacker (lib/packer_nopad.c:47-69): u becomes a copy of u_tmp."""
        n = 2 * volume * 4 * 18
        self.lib.unpack_gauge_field(int(volume), self._arr(u_tmp, n, "u_tmp"), self._arr(u, n, "u"))

    def unpack_gauge_field_inverse(self, volume: int, u_tmp: np.ndarray, u: np.ndarray) -> None:
        """Extension: the true inverse of pack_gauge_field."""
        n = 2 * volume * 4 * 18
        self.lib.unpack_gauge_field_inverse(int(volume), self._arr(u_tmp, n, "u_tmp"), self._arr(u, n, "u"))
