"""torch.autograd wrapper of the Dslash and its hand-written backward kernel.

``wilson_dslash(u, psi, latt, isign, cb)`` returns D psi on the target
checkerboard 1-cb (zeros on cb), with u the legacy packed link array
[V][4][3][3][2] and psi a legacy spinor [V][4][3][2], both float64 CUDA
tensors in the canonical cb2 order.  Its backward, for an upstream gradient
chi, is
  d/dpsi = D^dagger chi   (the Dslash kernel with -isign, onto checkerboard cb)
  d/du   = the spin-traced outer products of dslash_force (dS/d packed reals)
i.e. the vector-Jacobian product of S = Re <chi, D psi>.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import F64, check, latt_arg, ptr
from .device import WilsonDslash, current_stream


class _Dslash(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, psi, latt, isign, cb):
        op = WilsonDslash(latt, 4, F64, 18)
        op.load_gauge(u.detach())
        d_psi = op.pack_spinor(psi.detach())
        d_res = op.new_spinor(zero=True)
        op.apply(d_res, d_psi, isign, cb)
        out = _unpack(op, d_res)
        ctx.save_for_backward(u, psi)
        ctx.latt, ctx.isign, ctx.cb = latt, isign, cb
        ctx.op, ctx.d_psi = op, d_psi
        return out

    @staticmethod
    def backward(ctx, grad_out):
        u, psi = ctx.saved_tensors
        op, d_psi = ctx.op, ctx.d_psi
        d_chi = op.pack_spinor(grad_out.detach().to(torch.float64).contiguous())
        grad_u = grad_psi = None
        if ctx.needs_input_grad[0]:
            grad_u = torch.empty_like(u, dtype=torch.float64)
            check(op.lib.dslash_force(ptr(grad_u), None, None, ptr(d_chi), ptr(d_psi), latt_arg(ctx.latt),
                                      ctx.isign, ctx.cb, current_stream()))
        if ctx.needs_input_grad[1]:
            d_g = op.new_spinor(zero=True)
            op.apply(d_g, d_chi, -ctx.isign, 1 - ctx.cb)  # D^dagger chi lands on cb
            grad_psi = _unpack(op, d_g)
        return grad_u, grad_psi, None, None, None


def _unpack(op: WilsonDslash, dev: torch.Tensor) -> torch.Tensor:
    d = torch.zeros(op.V * 24, dtype=torch.float64, device=op.device)
    for cb in (0, 1):
        check(op.lib.dslash_unpack_spinor(ptr(d), cb, ptr(dev) + cb * op.sb, F64, op._L, 4, F64, None, 0,
                                          current_stream()))
    return d.view(op.V, 4, 3, 2)


def wilson_dslash(u: torch.Tensor, psi: torch.Tensor, latt, isign: int = 1, cb: int = 0) -> torch.Tensor:
    return _Dslash.apply(u, psi, tuple(int(v) for v in latt), int(isign), int(cb))


def su3_force(u: torch.Tensor, chi: torch.Tensor, psi: torch.Tensor, latt, isign: int = 1, cb: int = 0):
    """(grad_packed, force): dS/d packed reals and the su(3) projection TA((U G)^dagger)."""
    op = WilsonDslash(latt, 4, F64, 18)
    d_chi = op.pack_spinor(chi.detach())
    d_psi = op.pack_spinor(psi.detach())
    grad = torch.empty(op.V * 72, dtype=torch.float64, device=op.device)
    force = torch.empty(op.V * 72, dtype=torch.float64, device=op.device)
    uu = u.detach().reshape(-1).to(op.device).contiguous()
    check(_lib.core().dslash_force(ptr(grad), ptr(force), ptr(uu), ptr(d_chi), ptr(d_psi), latt_arg(latt),
                                   int(isign), int(cb), current_stream()))
    return grad.view(op.V, 4, 3, 3, 2), force.view(op.V, 4, 3, 3, 2)
