"""Batched Dslash: many right-hand sides on one gauge, and one (optionally
persistent) launch over several lattices of uneven volumes
(dslash_apply_mrhs / dslash_tune_mrhs in include/dslash_mi355x.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import MrhsProblem, check, ptr
from .device import WilsonDslash, current_stream


class RhsBatch:
    """nrhs single-checkerboard fields of one lattice, back to back in HBM."""

    def __init__(self, op: WilsonDslash, nrhs: int, zero: bool = False):
        self.op = op
        self.nrhs = int(nrhs)
        self.stride = op.spinor_bytes()
        f = torch.zeros if zero else torch.empty
        self.buf = f(self.nrhs * self.stride, dtype=torch.uint8, device=op.device)

    def field(self, k: int) -> torch.Tensor:
        return self.buf[k * self.stride:(k + 1) * self.stride]

    def set_from_full(self, k: int, full: torch.Tensor, cb: int) -> None:
        self.field(k).copy_(self.op.cb_view(full, cb))


@dataclass
class BatchItem:
    op: WilsonDslash       # lattice + packed gauge (F32 or F64, 4D)
    psi: RhsBatch          # source fields (checkerboard 1 - target_cb)
    res: RhsBatch          # target fields
    isign: int = 1
    target_cb: int = 0


def _problems(items) -> C.Array:
    arr = (MrhsProblem * len(items))()
    for i, it in enumerate(items):
        if it.op.nd != 4:
            raise ValueError("the batched Dslash is 4D")
        if it.psi.nrhs != it.res.nrhs:
            raise ValueError("psi and res batches differ in size")
        p = arr[i]
        p.latt = (C.c_int * 4)(*it.op.latt)
        p.d_gauge_cb = ptr(it.op.gauge) + it.target_cb * it.op.gb
        p.d_psi = ptr(it.psi.buf)
        p.d_res = ptr(it.res.buf)
        p.field_stride = it.psi.stride
        p.nrhs = it.psi.nrhs
        p.isign = it.isign
        p.target_cb = it.target_cb
    return arr


def _common(items):
    precs = {it.op.prec for it in items}
    gs = {it.op.G for it in items}
    if len(precs) != 1 or len(gs) != 1:
        raise ValueError("all lattices of one launch share the format and link reals")
    return precs.pop(), gs.pop()


def apply_batch(items, rhs_block: int = 0, block_threads: int = 0, grid_blocks: int = 0, stream=None) -> None:
    """One launch over every field of every item (grid_blocks > 0: persistent grid)."""
    prec, G = _common(items)
    arr = _problems(items)
    check(_lib.core().dslash_apply_mrhs(arr, len(items), prec, G, rhs_block, block_threads, grid_blocks,
                                        current_stream() if stream is None else stream))


def tune_batch(items, persistent: bool = False) -> dict:
    """Time the launch shapes on these items; the best is used by apply_batch(..., 0, 0, ...)."""
    prec, G = _common(items)
    arr = _problems(items)
    R, B, Gr, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
    check(_lib.core().dslash_tune_mrhs(arr, len(items), prec, G, 1 if persistent else 0, current_stream(),
                                       C.byref(R), C.byref(B), C.byref(Gr), C.byref(ms)))
    out = {"rhs_block": R.value, "block_threads": B.value, "grid_blocks": Gr.value, "ms": ms.value}
    out["t_block"] = get_launch(prec, G, persistent)["t_block"]
    return out


def get_launch(prec: int, G: int, persistent: bool = False) -> dict:
    """The batched launch shape apply_batch uses for (format, link reals, persistent)."""
    R, B, Gr, T = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    check(_lib.core().dslash_get_mrhs_launch(prec, G, 1 if persistent else 0, C.byref(R), C.byref(B),
                                             C.byref(Gr), C.byref(T)))
    return {"rhs_block": R.value, "block_threads": B.value, "grid_blocks": Gr.value, "t_block": T.value}


def set_launch(prec: int, G: int, persistent: bool, rhs_block: int, block_threads: int, grid_blocks: int,
               t_block: int) -> None:
    check(_lib.core().dslash_set_mrhs_launch(prec, G, 1 if persistent else 0, rhs_block, block_threads,
                                             grid_blocks, t_block))


def tune_apply(op: WilsonDslash, res: torch.Tensor, psi: torch.Tensor) -> dict:
    """Tune the single-RHS launch shape of op's (nd, format, link reals)."""
    B, Rm, M, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
    check(op.lib.dslash_tune_apply(ptr(res), ptr(psi), ptr(op.gauge), op._L, op.nd, op.prec, op.G,
                                   current_stream(), C.byref(B), C.byref(Rm), C.byref(ms)))
    check(op.lib.dslash_get_launch_march(op.nd, op.prec, op.G, C.byref(M)))
    return {"block_threads": B.value, "xcd_remap": Rm.value, "march": M.value, "ms": ms.value}


def flops(items) -> int:
    return sum(1320 * it.op.vh * it.psi.nrhs for it in items)


def compulsory_bytes(items) -> int:
    """Links once per lattice, one source spinor read and one target spinor write per field."""
    tot = 0
    for it in items:
        s = 8 if it.op.prec == _lib.F64 else 4
        tot += it.op.vh * (8 * it.op.G * s + it.psi.nrhs * 48 * s)
    return tot


def gen_batch(op: WilsonDslash, nrhs: int, cb: int, seed: int = 1, field0: int = 0) -> tuple[RhsBatch, np.ndarray]:
    """nrhs seeded Gaussian fields on checkerboard cb (device) and their legacy host copies."""
    from .device import gen_spinor_host
    b = RhsBatch(op, nrhs)
    host = []
    dt = np.float64 if op.prec == _lib.F64 else np.float32
    for k in range(nrhs):
        h = gen_spinor_host(op.latt, 4, seed=seed, field=field0 + k, dtype=dt)
        b.field(k).copy_(op.pack_cb(h, cb))
        host.append(h)
    return b, np.stack(host)
