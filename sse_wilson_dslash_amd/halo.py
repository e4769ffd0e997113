"""The domain-decomposition (parscalar) split of the Wilson Dslash.

Mirrors the stages of the reference's multi-node path
(lib/sse_su3dslash_32bit_parscalar.c:108-820): ``pack`` is decomp/decomp_hvv of
the boundary source sites into face buffers, ``interior`` the fused kernel over
every hop that stays inside the local lattice, ``exterior`` mvv_recons/recons
of the received faces on the boundary target sites.  With one GPU the exchange
is the identity (``apply`` passes the send buffer as the receive buffer); a
multi-node caller copies faces between neighbours with ``face`` views.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, ptr
from .device import WilsonDslash, current_stream


class HaloDslash:
    """Partition directions ``dir_mask`` (bit mu) of a WilsonDslash operator."""

    def __init__(self, op: WilsonDslash, dir_mask: int):
        self.op = op
        self.mask = int(dir_mask)
        off = (C.c_size_t * 8)()
        fb = (C.c_size_t * 4)()
        tot = C.c_size_t(0)
        check(op.lib.dslash_face_layout(op._L, op.nd, op.prec, self.mask, off, fb, C.byref(tot)))
        self.offsets = [int(v) for v in off]
        self.face_bytes = [int(v) for v in fb]
        self.total_bytes = int(tot.value)

    def new_faces(self) -> torch.Tensor:
        return torch.empty(max(self.total_bytes, 256), dtype=torch.uint8, device=self.op.device)

    def face(self, buf: torch.Tensor, mu: int, forward: bool) -> torch.Tensor:
        """Byte view of one face: forward = sites on x_mu = 0 (for the -mu side's
        forward hops), backward = sites on x_mu = L-1 (for the +mu side)."""
        o = self.offsets[2 * mu + (0 if forward else 1)]
        return buf[o:o + self.face_bytes[mu]]

    def _gauge(self):
        if not self.op.has_gauge:
            raise RuntimeError("load_gauge() first")
        return ptr(self.op.gauge)

    def pack(self, send: torch.Tensor, psi: torch.Tensor, isign: int, cb: int) -> torch.Tensor:
        op = self.op
        check(op.lib.dslash_halo_pack(ptr(send), ptr(psi), self._gauge(), op._L, op.nd, op.prec, op.G,
                                      int(isign), int(cb), self.mask, current_stream()))
        return send

    def interior(self, res: torch.Tensor, psi: torch.Tensor, isign: int, cb: int) -> torch.Tensor:
        op = self.op
        check(op.lib.dslash_halo_interior(ptr(res), ptr(psi), self._gauge(), op._L, op.nd, op.prec, op.G,
                                          int(isign), int(cb), self.mask, current_stream()))
        return res

    def exterior(self, res: torch.Tensor, recv: torch.Tensor, isign: int, cb: int) -> torch.Tensor:
        op = self.op
        check(op.lib.dslash_halo_exterior(ptr(res), ptr(recv), self._gauge(), op._L, op.nd, op.prec, op.G,
                                          int(isign), int(cb), self.mask, current_stream()))
        return res

    def apply(self, res: torch.Tensor, psi: torch.Tensor, isign: int, cb: int,
              faces: torch.Tensor | None = None) -> torch.Tensor:
        """D psi through the split with the one-GPU (periodic) exchange."""
        f = self.new_faces() if faces is None else faces
        self.pack(f, psi, isign, cb)
        self.interior(res, psi, isign, cb)
        self.exterior(res, f, isign, cb)
        return res
