"""Mixed-precision even-odd Wilson CG (dslash_cg_* in include/dslash_mi355x.h).

Solves A x = b on the even checkerboard with A = Mhat^dagger Mhat,
Mhat = 1 - kappa^2 D_eo D_oe, M = 1 - kappa D.  The inner iteration runs in a
16-bit storage format (or FP32 / FP64); the residual is recomputed in FP64 at
reliable updates and convergence is declared on that FP64 true residual.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import F16, F32, F64, CgParams, CgResult, check, latt_arg, ptr


class WilsonCG:
    def __init__(self, latt, kappa: float = 0.12, tol: float = 1e-10, prec_sloppy: int = F16,
                 gauge_reals_sloppy: int = 12, delta: float = 0.1, max_iter: int = 20000, use_graph: bool = True,
                 gauge_reals_precise: int = 18):
        self.lib = _lib.core()
        self.latt = tuple(int(v) for v in latt)
        self.params = CgParams(kappa, tol, delta, max_iter, prec_sloppy, gauge_reals_sloppy, 1 if use_graph else 0, 1,
                               gauge_reals_precise)
        h = self.lib.dslash_cg_create(latt_arg(self.latt), C.byref(self.params))
        if not h:
            check(self.lib.dslash_last_error() or _lib.ERR_INVALID)
        self.h = h
        self.prec_sloppy = prec_sloppy

    def load_gauge(self, u_packed) -> None:
        if isinstance(u_packed, np.ndarray):
            d = torch.from_numpy(np.ascontiguousarray(u_packed).reshape(-1)).cuda()
        else:
            d = u_packed.reshape(-1).cuda().contiguous()
        pin = F64 if d.dtype == torch.float64 else F32
        check(self.lib.dslash_cg_load_gauge(self.h, ptr(d), pin, torch.cuda.current_stream().cuda_stream))

    def solve(self, x: torch.Tensor, b: torch.Tensor, zero_guess: bool = True, raise_on_fail: bool = True) -> dict:
        """x, b: even-cb F64 device fields.  Returns the solver statistics."""
        if bool(self.params.zero_guess) != bool(zero_guess):
            self.params.zero_guess = 1 if zero_guess else 0
            check(self.lib.dslash_cg_set_params(self.h, C.byref(self.params)))
        torch.cuda.current_stream().synchronize()
        res = CgResult()
        st = self.lib.dslash_cg_solve(self.h, ptr(x), ptr(b), C.byref(res))
        out = {"iterations": res.iterations, "reliable_updates": res.reliable_updates,
               "converged": bool(res.converged), "true_residual": res.true_residual, "seconds": res.seconds,
               "status": st}
        if st != 0 and raise_on_fail:
            check(st)
        return out

    def set_params(self, **kw) -> None:
        for k, v in kw.items():
            setattr(self.params, k, v)
        check(self.lib.dslash_cg_set_params(self.h, C.byref(self.params)))

    def apply_op(self, out: torch.Tensor, inp: torch.Tensor, sloppy: bool = False) -> None:
        torch.cuda.current_stream().synchronize()
        check(self.lib.dslash_cg_apply_op(self.h, ptr(out), ptr(inp), 1 if sloppy else 0))

    def close(self) -> None:
        if self.h:
            self.lib.dslash_cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
