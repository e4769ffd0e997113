"""The Dslash at a lattice 16x the headline volume (64^3 x 128, 33.5M sites):
full-lattice parity with the oracle, so site indexing, plane strides and the
XCD block order are exercised far beyond the 32^3 x 64 configs.  Needs ~42 GB of
HBM and ~20 GB of host memory."""
import time

import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import F32
from sse_wilson_dslash_amd._lib import check, latt_arg, ptr
from sse_wilson_dslash_amd.device import WilsonDslash, current_stream
from tests import refimpl as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def test_64c128_full_lattice_against_oracle(oracle):
    latt = (64, 64, 64, 128)
    t0 = time.time()
    op = WilsonDslash(latt, 4, F32, 18)
    d_u = torch.empty(op.V * 72, dtype=torch.float32, device="cuda")
    check(op.lib.dslash_gen_gauge_device(ptr(d_u), F32, latt_arg(latt), 4, 5, current_stream()))
    op.load_gauge(d_u)
    d_psi_aos = torch.empty(op.V * 24, dtype=torch.float32, device="cuda")
    check(op.lib.dslash_gen_spinor_device(ptr(d_psi_aos), F32, latt_arg(latt), 4, 5, 0, current_stream()))
    d_psi = op.pack_spinor(d_psi_aos)
    u = d_u.cpu().numpy().reshape(op.V, 4, 3, 3, 2)
    psi = d_psi_aos.cpu().numpy().reshape(op.V, 4, 3, 2)
    del d_u, d_psi_aos
    torch.cuda.empty_cache()
    print(f"[large 64^3x128] fields generated and copied to host in {time.time() - t0:.1f} s", flush=True)
    d_res = op.new_spinor(zero=True)
    vh = op.vh
    want = np.empty_like(psi)
    for isign, cb in ((1, 0), (-1, 1)):
        op.apply(d_res, d_psi, isign, cb)
        got = op.unpack_spinor(d_res, dtype=np.float32, cbs=(1 - cb,))
        t1 = time.time()
        oracle.dslash(u, psi, want, latt, isign, cb)
        t = slice((1 - cb) * vh, (2 - cb) * vh)
        g, w = got.reshape(-1, 24)[t], want.reshape(-1, 24)[t]
        e2, einf = R.rel_err(g, w), R.max_rel_err(g, w)
        print(f"  isign={isign:+d} cb={cb}: oracle {time.time() - t1:.1f} s, rel L2 {e2:.2e}, rel max {einf:.2e}",
              flush=True)
        assert e2 < 1e-5 and einf < 1e-5
