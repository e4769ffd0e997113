"""Device-resident Wilson Dslash on one MI355X (torch supplies memory and streams).

``WilsonDslash`` owns the packed link bundles of both target checkerboards in
HBM and applies the fused HIP kernel through the C ABI of libdslash_mi355x.so.
Fields are torch uint8 tensors holding the library's device layout (see
include/dslash_mi355x.h); ``pack_spinor``/``unpack_spinor`` convert from and to
the legacy full-lattice [V][4][3][2] arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import F32, F64, F16, BF16, F64S, check, latt_arg, ptr

_NP = {F32: np.float32, F64: np.float64}
_TORCH = {F32: torch.float32, F64: torch.float64}


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _prec_of_dtype(dtype) -> int:
    if dtype in (np.float32, torch.float32) or dtype == np.dtype(np.float32):
        return F32
    if dtype in (np.float64, torch.float64) or dtype == np.dtype(np.float64):
        return F64
    raise TypeError(f"legacy arrays must be float32 or float64, got {dtype}")


class WilsonDslash:
    """The Wilson hopping term on a periodic lattice, fields resident in HBM.

    nd = 4: sum over mu = 0..3 in the cb2 order; nd = 3: mu = 0..2 in the cb3d
    order.  prec: F32, F64, F16 or BF16 (16-bit formats store spinors with a
    per-site norm and links as raw values, arithmetic in FP32), or F64S (FP64
    spinors, FP32 links and arithmetic: the sloppy FP64 operator).  gauge_reals:
    18 (full links) or 12 (two rows, SU(3) links only).
    """

    def __init__(self, latt, nd: int = 4, prec: int = F32, gauge_reals: int = 18, device: str = "cuda"):
        self.lib = _lib.core()
        self.latt = tuple(int(v) for v in latt)
        self.nd = int(nd)
        self.prec = int(prec)
        self.G = int(gauge_reals)
        self.device = torch.device(device)
        self._L = latt_arg(self.latt)
        check(self.lib.dslash_check_lattice(self._L, self.nd))
        self.V = int(np.prod(self.latt))
        self.vh = self.V // 2
        self.sb = int(self.lib.dslash_spinor_cb_bytes(self._L, self.nd, self.prec))
        self.gb = int(self.lib.dslash_gauge_cb_bytes(self._L, self.nd, self.prec, self.G))
        if self.sb == 0 or self.gb == 0:
            check(self.lib.dslash_last_error() or _lib.ERR_INVALID)
        self.gauge = None  # allocated by load_gauge
        self.has_gauge = False

    # ---- sizes and allocation ------------------------------------------
    def spinor_bytes(self, prec: int | None = None) -> int:
        p = self.prec if prec is None else prec
        return int(self.lib.dslash_spinor_cb_bytes(self._L, self.nd, p))

    def new_spinor(self, full: bool = True, prec: int | None = None, zero: bool = False) -> torch.Tensor:
        n = self.spinor_bytes(prec) * (2 if full else 1)
        f = torch.zeros if zero else torch.empty
        return f(n, dtype=torch.uint8, device=self.device)

    def cb_view(self, full: torch.Tensor, cb: int, prec: int | None = None) -> torch.Tensor:
        sb = self.spinor_bytes(prec)
        return full[cb * sb:(cb + 1) * sb]

    # ---- gauge ---------------------------------------------------------
    def load_gauge(self, u_packed, site_table=None) -> None:
        """Pack a legacy [V][4][3][3][2] gauge (numpy or torch, f32/f64) for both target cbs."""
        stream = current_stream()
        if isinstance(u_packed, np.ndarray):
            pin = _prec_of_dtype(u_packed.dtype)
            d_u = torch.from_numpy(np.ascontiguousarray(u_packed).reshape(-1)).to(self.device)
        else:
            pin = _prec_of_dtype(u_packed.dtype)
            d_u = u_packed.reshape(-1).to(self.device).contiguous()
        if d_u.numel() != self.V * 72:
            raise ValueError(f"gauge must hold V*72 reals ({self.V * 72}), got {d_u.numel()}")
        d_tab = None
        if site_table is not None:
            d_tab = torch.as_tensor(np.asarray(site_table, dtype=np.int32)).to(self.device)
        if self.gauge is None:
            self.gauge = torch.empty(2 * self.gb, dtype=torch.uint8, device=self.device)
        for t in (0, 1):
            check(self.lib.dslash_pack_gauge(ptr(self.gauge) + t * self.gb, t, ptr(d_u), pin, self._L, self.nd,
                                             self.prec, self.G, ptr(d_tab), stream))
        torch.cuda.current_stream().synchronize()
        self.has_gauge = True

    def gen_gauge(self, seed: int = 1) -> None:
        """Generate the seeded SU(3) gauge directly in HBM and pack it."""
        pin = F64 if self.prec in (F64, F64S) else F32
        d_u = torch.empty(self.V * 72, dtype=_TORCH[pin], device=self.device)
        check(self.lib.dslash_gen_gauge_device(ptr(d_u), pin, self._L, self.nd, seed, current_stream()))
        self.load_gauge(d_u)

    # ---- spinors -------------------------------------------------------
    def pack_spinor(self, aos, out: torch.Tensor | None = None, prec: int | None = None,
                    site_table=None) -> torch.Tensor:
        """Legacy full-lattice [V][4][3][2] spinor -> device layout (both cbs)."""
        p = self.prec if prec is None else prec
        if isinstance(aos, np.ndarray):
            pin = _prec_of_dtype(aos.dtype)
            d = torch.from_numpy(np.ascontiguousarray(aos).reshape(-1)).to(self.device)
        else:
            pin = _prec_of_dtype(aos.dtype)
            d = aos.reshape(-1).to(self.device).contiguous()
        if d.numel() != self.V * 24:
            raise ValueError("spinor must hold V*24 reals")
        if out is None:
            out = self.new_spinor(True, p)
        d_tab = None
        if site_table is not None:
            d_tab = torch.as_tensor(np.asarray(site_table, dtype=np.int32)).to(self.device)
        sb = self.spinor_bytes(p)
        for cb in (0, 1):
            check(self.lib.dslash_pack_spinor(ptr(out) + cb * sb, cb, ptr(d), pin, self._L, self.nd, p,
                                              ptr(d_tab), current_stream()))
        return out

    def unpack_spinor(self, dev: torch.Tensor, dtype=np.float64, prec: int | None = None,
                      cbs=(0, 1)) -> np.ndarray:
        """Device layout (full lattice) -> legacy [V][4][3][2] numpy array (other cbs zero)."""
        p = self.prec if prec is None else prec
        pout = _prec_of_dtype(np.dtype(dtype))
        d = torch.zeros(self.V * 24, dtype=_TORCH[pout], device=self.device)
        sb = self.spinor_bytes(p)
        for cb in cbs:
            check(self.lib.dslash_unpack_spinor(ptr(d), cb, ptr(dev) + cb * sb, p, self._L, self.nd, pout,
                                                None, 0, current_stream()))
        return d.cpu().numpy().reshape(self.V, 4, 3, 2)

    def pack_cb(self, aos, cb: int, prec: int | None = None) -> torch.Tensor:
        """One checkerboard of a legacy full-lattice spinor -> a single-cb device field."""
        p = self.prec if prec is None else prec
        if isinstance(aos, np.ndarray):
            pin = _prec_of_dtype(aos.dtype)
            d = torch.from_numpy(np.ascontiguousarray(aos).reshape(-1)).to(self.device)
        else:
            pin = _prec_of_dtype(aos.dtype)
            d = aos.reshape(-1).to(self.device).contiguous()
        out = self.new_spinor(False, p)
        check(self.lib.dslash_pack_spinor(ptr(out), int(cb), ptr(d), pin, self._L, self.nd, p, None,
                                          current_stream()))
        return out

    def unpack_cb(self, dev_cb: torch.Tensor, cb: int, dtype=np.float64, prec: int | None = None) -> np.ndarray:
        """A single-cb device field -> legacy full-lattice numpy array (other cb zero)."""
        p = self.prec if prec is None else prec
        pout = _prec_of_dtype(np.dtype(dtype))
        d = torch.zeros(self.V * 24, dtype=_TORCH[pout], device=self.device)
        check(self.lib.dslash_unpack_spinor(ptr(d), int(cb), ptr(dev_cb), p, self._L, self.nd, pout, None, 0,
                                            current_stream()))
        return d.cpu().numpy().reshape(self.V, 4, 3, 2)

    def gen_spinor(self, seed: int = 1, field: int = 0, prec: int | None = None) -> torch.Tensor:
        """Seeded Gaussian spinor generated in HBM, returned in device layout."""
        pin = F64 if (self.prec if prec is None else prec) in (F64, F64S) else F32
        d = torch.empty(self.V * 24, dtype=_TORCH[pin], device=self.device)
        check(self.lib.dslash_gen_spinor_device(ptr(d), pin, self._L, self.nd, seed, field, current_stream()))
        return self.pack_spinor(d, prec=prec)

    # ---- the operator --------------------------------------------------
    def apply(self, res: torch.Tensor, psi: torch.Tensor, isign: int, cb: int) -> torch.Tensor:
        """Legacy semantics on full-lattice device fields: cb is the source checkerboard."""
        if not self.has_gauge:
            raise RuntimeError("load_gauge() first")
        check(self.lib.dslash_apply_full(ptr(res), ptr(psi), ptr(self.gauge), self._L, self.nd, self.prec,
                                         self.G, int(isign), int(cb), current_stream()))
        return res

    def apply_cb(self, res_cb: torch.Tensor, psi_cb: torch.Tensor, isign: int, target_cb: int) -> torch.Tensor:
        """Single-checkerboard form: res_cb(target) = D psi_cb(1 - target)."""
        if not self.has_gauge:
            raise RuntimeError("load_gauge() first")
        g = ptr(self.gauge) + int(target_cb) * self.gb
        check(self.lib.dslash_apply(ptr(res_cb), ptr(psi_cb), g, self._L, self.nd, self.prec, self.G,
                                    int(isign), int(target_cb), current_stream()))
        return res_cb

    def time_apply(self, res: torch.Tensor, psi: torch.Tensor, n: int, alternate: bool = True) -> float:
        """Average ms per launch of n back-to-back applications (HIP events on the launch stream)."""
        ms = C.c_float(0.0)
        check(self.lib.dslash_time_apply(ptr(res), ptr(psi), ptr(self.gauge), self._L, self.nd, self.prec, self.G,
                                         int(n), 1 if alternate else 0, current_stream(), C.byref(ms)))
        return float(ms.value)

    # ---- accounting ----------------------------------------------------
    def flops_per_call(self) -> int:
        """1320 flop/site in 4D, 984 in 3D (8 or 6 directions x 168 - 24)."""
        return (1320 if self.nd == 4 else 984) * self.vh

    def bytes_per_call(self) -> int:
        """Compulsory HBM bytes: read one source spinor and 2*nd links, write one spinor, per target site."""
        s = {F32: 4, F64: 8, F16: 2, BF16: 2, F64S: 8}[self.prec]
        sl = 4 if self.prec == F64S else s  # F64S: FP32 links
        norm = 8 if self.prec in (F16, BF16) else 0
        return self.vh * (48 * s + 2 * self.nd * self.G * sl + norm)


def gen_gauge_host(latt, nd: int = 4, seed: int = 1, dtype=np.float64) -> np.ndarray:
    """Seeded SU(3) gauge in the legacy packed layout, canonical order of nd."""
    lib = _lib.core()
    V = int(np.prod(latt))
    u = np.empty((V, 4, 3, 3, 2), dtype=dtype)
    check(lib.dslash_gen_gauge_host(u.ctypes.data, _prec_of_dtype(u.dtype), latt_arg(latt), nd, seed))
    return u


def gen_spinor_host(latt, nd: int = 4, seed: int = 1, field: int = 0, dtype=np.float64) -> np.ndarray:
    """Seeded Gaussian spinor in the legacy layout, canonical order of nd."""
    lib = _lib.core()
    V = int(np.prod(latt))
    p = np.empty((V, 4, 3, 2), dtype=dtype)
    check(lib.dslash_gen_spinor_host(p.ctypes.data, _prec_of_dtype(p.dtype), latt_arg(latt), nd, seed, field))
    return p
