"""dslash_normal_op_noalloc (the normal operator A = Mhat^dagger Mhat with its
work fields in the library's fixed device array) against dslash_cg_apply_op
(the same operator with a CG solver's workspace), through the C ABI: the same
bytes for FP64, FP32 and F16 fields, on two calls in a row where the second
lattice is smaller, no device memory allocated by repeated calls, and a clear
error for a lattice whose work fields exceed the array."""
import ctypes as C

import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import _lib
from sse_wilson_dslash_amd._lib import F16, F32, F64
from sse_wilson_dslash_amd.device import WilsonDslash, gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.layout import volume
from sse_wilson_dslash_amd.solver import WilsonCG

pytestmark = pytest.mark.gpu
KAPPA = 0.12


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def noalloc(out, inp, gauge, latt, prec, G, kappa=KAPPA):
    return _lib.core().dslash_normal_op_noalloc(_lib.ptr(out), _lib.ptr(inp), _lib.ptr(gauge), _lib.latt_arg(latt),
                                                 kappa, prec, G, torch.cuda.current_stream().cuda_stream)


def existing(latt, u, d_in, prec, G):
    """dslash_cg_apply_op: FP64 fields use the solver's FP64 operator (links of
    gauge_reals_precise), other formats its sloppy operator."""
    if prec == F64:
        cg = WilsonCG(latt, kappa=KAPPA, prec_sloppy=F64, gauge_reals_sloppy=G, gauge_reals_precise=G)
        sloppy = False
    else:
        cg = WilsonCG(latt, kappa=KAPPA, prec_sloppy=prec, gauge_reals_sloppy=G)
        sloppy = True
    cg.load_gauge(u)
    out = torch.zeros_like(d_in)
    cg.apply_op(out, d_in, sloppy=sloppy)
    cg.close()
    return out


@pytest.mark.parametrize("prec,G,lattices", [(F64, 18, [(16, 16, 16, 32), (8, 8, 8, 8)]),
                                             (F32, 12, [(24, 24, 24, 48), (4, 6, 8, 4)]),
                                             (F16, 12, [(16, 16, 16, 16), (8, 4, 4, 8)])])
def test_noalloc_matches_cg_apply_op(prec, G, lattices):
    for k, latt in enumerate(lattices):  # the second call has the smaller input
        print(f"[no-alloc] prec={prec} G={G} latt={latt} call {k + 1}", flush=True)
        u = gen_gauge_host(latt, 4, seed=3 + k)
        b = gen_spinor_host(latt, 4, seed=3 + k, field=2)
        op = WilsonDslash(latt, 4, prec, G)
        op.load_gauge(u)
        d_in = op.pack_cb(b, 0)
        want = existing(latt, u, d_in, prec, G)
        got = torch.full_like(d_in, 0x5A)
        assert noalloc(got, d_in, op.gauge, latt, prec, G) == 0, _lib.core().dslash_last_error_message()
        torch.cuda.synchronize()
        assert torch.equal(got, want), f"{latt}: differs from dslash_cg_apply_op"
        # the result is A in, not a copy or zero: <in, A in> = |Mhat in|^2 > 0
        if prec == F64:
            x = op.unpack_cb(d_in, 0).reshape(-1)
            y = op.unpack_cb(got, 0).reshape(-1)
            assert float(x @ y) > 0.1 * float(x @ x)
        del op, d_in, want, got
        torch.cuda.empty_cache()


def test_noalloc_allocates_nothing_and_repeats():
    latt = (8, 8, 8, 16)
    op = WilsonDslash(latt, 4, F64, 18)
    op.load_gauge(gen_gauge_host(latt, 4, seed=9))
    d_in = op.pack_cb(gen_spinor_host(latt, 4, seed=9), 0)
    outs = [torch.zeros_like(d_in) for _ in range(3)]
    assert noalloc(outs[0], d_in, op.gauge, latt, F64, 18) == 0
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for o in outs[1:]:
        assert noalloc(o, d_in, op.gauge, latt, F64, 18) == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert torch.equal(outs[1], outs[0]) and torch.equal(outs[2], outs[0])


def test_noalloc_rejects_what_does_not_fit():
    """FP64 at 24^3 x 48 needs 2 x 63.7 MB of work fields, more than the
    64 MiB array: a clear error before any launch."""
    lib = _lib.core()
    latt = (24, 24, 24, 48)
    t = torch.zeros(16, dtype=torch.uint8, device="cuda")
    st = lib.dslash_normal_op_noalloc(_lib.ptr(t), _lib.ptr(t) + 4096, _lib.ptr(t), _lib.latt_arg(latt), KAPPA,
                                      F64, 18, None)
    assert st == _lib.ERR_INVALID
    msg = lib.dslash_last_error_message()
    assert b"fixed array" in msg and b"bytes" in msg
    assert noalloc(t, t, t, (8, 8, 8, 8), F64, 18) == _lib.ERR_INVALID  # out == in
    assert lib.dslash_normal_op_noalloc(_lib.ptr(t), _lib.ptr(t) + 16, _lib.ptr(t), _lib.latt_arg((8, 8, 8, 8)),
                                        KAPPA, F64, 13, None) == _lib.ERR_INVALID
