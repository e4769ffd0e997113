"""Per-direction checks of the HIP kernels with delta sources, in place of the
reference's per-site-kernel unit tests (tests/testSiteDslash32bit.cc,
tests/testSiteDslash64bit.cc).

With every link the identity, a source that is one unit vector (spin s,
colour c) at site x reaches exactly 2*nd target sites: x - mu^ through the
forward hop, which applies (1 - isign gamma_mu), and x + mu^ through the
backward hop, which applies (1 + isign gamma_mu) (the operator of
lib/sse_su3dslash_32bit_scalar.c:39-64 with the projectors of
lib/site_dslash_64bit_scalar.c:34-166, 638-926, 1228-1549, 2171-2474).  So each
of those sites must hold column s of that projector, in colour c, and every
other target site must be zero; the values are 0, +-1, +-i, exact in every
format.  Many sources sit in one field, far enough apart that no two responses
meet, at the places where the kernels change code path: tile and box edges,
march chunk boundaries and the periodic wrap.  Expected values come from the
explicit DeGrand-Rossi gamma matrices (tests/refimpl.py), not from the oracle,
and a mismatch names the source, mu, hop and component that is wrong.  Each
case runs in a chosen launch shape (one site per thread, marching, LDS box
tiles, batched box tiles), checked to be the one that actually ran.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import _lib
from sse_wilson_dslash_amd._lib import BF16, F16, F32, F64, F64S
from sse_wilson_dslash_amd.device import WilsonDslash
from sse_wilson_dslash_amd.layout import full_of_coords, volume
from tests import refimpl as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def _dist(a, b, latt, nd):
    d = 0
    for mu in range(nd):
        k = abs(int(a[mu]) - int(b[mu])) % latt[mu]
        d += min(k, latt[mu] - k)
    if nd == 3:  # 3D: different t slices never meet
        d += 0 if a[3] == b[3] else 100
    return d


def place_sources(latt, nd, cb, count=48):
    """Source sites on checkerboard cb at tile / box / chunk edges and at the
    periodic wrap, pairwise >= 3 apart (so responses at +-1 never coincide),
    each with its own (spin, colour) component, cycling through all 12."""
    L = latt
    cand_axis = [sorted({0, 1, 3, 4, 7, 8, L[m] // 2, L[m] - 2, L[m] - 1} & set(range(L[m]))) for m in range(4)]
    rng = np.random.default_rng(1234 + cb + 10 * nd)
    picks = []
    tries = 0
    while len(picks) < count and tries < 20000:
        tries += 1
        c = np.array([rng.choice(cand_axis[m]) for m in range(4)], dtype=np.int64)
        par = (c[0] + c[1] + c[2] + (c[3] if nd == 4 else 0)) & 1
        if par != cb:
            c[0] = (c[0] + 1) % L[0]
        if all(_dist(c, p, L, nd) >= 3 for p in picks):
            picks.append(c)
    assert len(picks) >= 24, "could not place enough separated sources"
    return [(p, k % 4, (k // 4) % 3) for k, p in enumerate(picks)]


def expected(latt, nd, isign, sources):
    """Full-lattice complex (V, 4, 3) response and, per target site, what put it there."""
    V = volume(latt)
    g = R.gammas()
    out = np.zeros((V, 4, 3), dtype=np.complex128)
    who = {}
    for n, (x, s, c) in enumerate(sources):
        for mu in range(nd):
            for hop, sign in (("forward", -1), ("backward", +1)):
                y = x.copy()
                y[mu] = (y[mu] + sign) % latt[mu]  # forward hop: y = x - mu^ sees x as y + mu^
                proj = np.eye(4) + sign * isign * g[mu]
                iy = int(full_of_coords(latt, nd, y))
                assert iy not in who
                out[iy, :, c] += proj[:, s]
                who[iy] = (n, tuple(int(v) for v in x), s, c, mu, hop)
    return out, who


def unit_gauge(latt, dtype):
    V = volume(latt)
    u = np.zeros((V, 4, 3, 3, 2), dtype=dtype)
    for i in range(3):
        u[:, :, i, i, 0] = 1.0
    return u


def delta_field(latt, nd, sources, dtype):
    V = volume(latt)
    psi = np.zeros((V, 4, 3, 2), dtype=dtype)
    for x, s, c in sources:
        psi[int(full_of_coords(latt, nd, x)), s, c, 0] = 1.0
    return psi


def compare(got_full, latt, nd, isign, cb, sources, label):
    vh = volume(latt) // 2
    want, who = expected(latt, nd, isign, sources)
    tcb = 1 - cb
    got = R.to_complex_spinor(got_full).reshape(-1, 4, 3)
    bad = np.nonzero(np.any(np.abs(got[tcb * vh:(tcb + 1) * vh] - want[tcb * vh:(tcb + 1) * vh]) > 1e-6,
                            axis=(1, 2)))[0]
    if bad.size:
        msgs = []
        for b in bad[:6]:
            iy = int(b) + tcb * vh
            w = who.get(iy)
            place = (f"source {w[0]} at {w[1]} spin {w[2]} colour {w[3]}: mu={w[4]} {w[5]} hop"
                     if w else "a site no hop should reach")
            msgs.append(f"target {iy}: {place}; got {np.round(got[iy], 3).tolist()} "
                        f"want {np.round(want[iy], 3).tolist()}")
        pytest.fail(f"{label} isign={isign} cb={cb}: {bad.size} wrong target sites\n" + "\n".join(msgs))


class Shape:
    """Sets the single-RHS marching length of (nd, prec, G) and restores it."""

    def __init__(self, nd, prec, G, march):
        self.key, self.march, self.lib = (nd, prec, G), march, _lib.core()

    def __enter__(self):
        m = C.c_int()
        _lib.check(self.lib.dslash_get_launch_march(*self.key, C.byref(m)))
        self.saved = m.value
        _lib.check(self.lib.dslash_set_launch_march(*self.key, self.march))
        return self

    def __exit__(self, *exc):
        _lib.check(self.lib.dslash_set_launch_march(*self.key, self.saved))


CASES = [
    # (nd, prec, G, march, lattice): march 1 one site per thread, 4/8 marching,
    # 16 LDS-staged box tiles, 32 box tiles with the LDS halo (at 24 x 48^3 boxes of
    # 12 x 3 x 6 sites, halo faces of unequal size; boxes that span y or z have halo
    # faces in the other direction only); 32^3 x 64 is large enough for every shape to run
    (4, F32, 18, 1, (32, 32, 32, 64)),
    (4, F32, 18, 8, (32, 32, 32, 64)),
    (4, F32, 18, 16, (32, 32, 32, 64)),
    (4, F32, 12, 16, (32, 32, 32, 64)),
    (4, F32, 18, 32, (32, 32, 32, 64)),
    (4, F32, 12, 32, (32, 32, 32, 64)),
    (4, F32, 18, 32, (24, 48, 48, 48)),
    (4, F32, 18, 32, (64, 4, 128, 64)),   # boxes span y: z faces only (halo 2 x 128)
    (4, F32, 12, 32, (32, 64, 4, 256)),   # boxes span z: y faces only (halo 2 x 64)
    (4, F64, 18, 16, (32, 32, 32, 64)),
    (4, F64, 18, 4, (32, 32, 32, 64)),
    (4, F64S, 18, 16, (32, 32, 32, 64)),
    (4, F64S, 18, 32, (32, 32, 32, 64)),
    (4, F16, 12, 1, (16, 16, 16, 16)),
    (4, BF16, 18, 1, (16, 16, 16, 16)),
    (3, F32, 18, 1, (16, 16, 16, 8)),
    (3, F32, 18, 8, (32, 32, 32, 64)),
    (3, F64, 18, 4, (32, 32, 32, 64)),
]


@pytest.mark.parametrize("nd,prec,G,march,latt", CASES)
def test_each_hop_matches_projector(nd, prec, G, march, latt):
    label = f"nd={nd} prec={prec} G={G} march={march} latt={latt}"
    print(f"[directions] {label}", flush=True)
    pin = np.float64 if prec in (F64, F64S) else np.float32
    op = WilsonDslash(latt, nd, prec, G)
    op.load_gauge(unit_gauge(latt, pin))
    with Shape(nd, prec, G, march):
        m = C.c_int()
        _lib.check(op.lib.dslash_launch_march_for(_lib.latt_arg(latt), nd, prec, G, C.byref(m)))
        assert m.value == march, f"{label}: the launch would run march {m.value}"
        for cb in (0, 1):
            sources = place_sources(latt, nd, cb)
            d_psi = op.pack_spinor(delta_field(latt, nd, sources, pin))
            for isign in (1, -1):
                d_res = op.new_spinor(zero=True)
                op.apply(d_res, d_psi, isign, cb)
                torch.cuda.synchronize()
                got = op.unpack_spinor(d_res, dtype=np.float64, cbs=(1 - cb,))
                compare(got, latt, nd, isign, cb, sources, label)
    del op
    torch.cuda.empty_cache()


@pytest.mark.parametrize("t_block", [1, 4, 16])
def test_each_hop_batched(t_block):
    """The batched kernel (4 fields, each with its own source placement, on one
    gauge): one site per thread, register t-marching and LDS box tiles."""
    from sse_wilson_dslash_amd.batch import BatchItem, RhsBatch, apply_batch, get_launch, set_launch
    latt = (24, 24, 24, 48)
    nrhs = 4
    print(f"[directions batched] t_block={t_block}", flush=True)
    op = WilsonDslash(latt, 4, F32, 18)
    op.load_gauge(unit_gauge(latt, np.float32))
    saved = get_launch(F32, 18, False)
    try:
        R_blk, block = (2, 192) if t_block == 16 else (1, 128)
        set_launch(F32, 18, False, R_blk, block, 0, t_block)
        for cb in (0, 1):
            srcs = [place_sources(latt, 4, cb, count=24 + 4 * k) for k in range(nrhs)]
            psi = RhsBatch(op, nrhs)
            for k in range(nrhs):
                psi.field(k).copy_(op.pack_cb(delta_field(latt, 4, srcs[k], np.float32), cb))
            for isign in (1, -1):
                res = RhsBatch(op, nrhs, zero=True)
                apply_batch([BatchItem(op, psi, res, isign, 1 - cb)])
                torch.cuda.synchronize()
                for k in range(nrhs):
                    got = op.unpack_cb(res.field(k), 1 - cb)
                    compare(got, latt, 4, isign, cb, srcs[k], f"batched t_block={t_block} field {k}")
    finally:
        set_launch(F32, 18, False, saved["rhs_block"], saved["block_threads"], saved["grid_blocks"],
                   saved["t_block"])
    del op
    torch.cuda.empty_cache()
