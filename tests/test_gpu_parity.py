"""GPU parity of the fused HIP Dslash with the oracle, through the C ABI.

Tolerances (relative L2 over the target checkerboard, and max-abs relative to
the largest entry): FP64 1e-13, FP32 1e-5.  The 16-bit formats are a storage
compression of the operator used inside the mixed-precision solver; their
error against the FP64 oracle is bounded and reported, not a parity claim.
"""
import glob
import os

import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import F16, F32, F64, BF16, F64S
from sse_wilson_dslash_amd.device import WilsonDslash, gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.layout import volume
from tests import refimpl as R

pytestmark = pytest.mark.gpu

TOL = {F64: 1e-13, F32: 1e-5}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def run_gpu(u, psi, latt, nd, prec, G, isign, cb, sentinel=None):
    op = WilsonDslash(latt, nd=nd, prec=prec, gauge_reals=G)
    op.load_gauge(u)
    d_psi = op.pack_spinor(psi)
    d_res = op.pack_spinor(sentinel if sentinel is not None else np.zeros_like(psi))
    op.apply(d_res, d_psi, isign, cb)
    torch.cuda.synchronize()
    dt = np.float64 if prec in (F64, F64S) else np.float32
    return op.unpack_spinor(d_res, dtype=dt)


def check(got, want, latt, cb, prec, label=""):
    vh = volume(latt) // 2
    t = slice((1 - cb) * vh, (2 - cb) * vh)
    g = got.reshape(-1, 24)[t]
    w = want.reshape(-1, 24)[t]
    e2, einf = R.rel_err(g, w), R.max_rel_err(g, w)
    tol = TOL[prec]
    assert e2 < tol and einf < tol, f"{label}: rel L2 {e2:.3e}, rel max {einf:.3e} (tol {tol})"
    return e2


@pytest.mark.parametrize("nd,latt", [(4, (4, 4, 4, 4)), (4, (2, 6, 4, 2)), (4, (6, 2, 2, 10)),
                                     (3, (4, 4, 4, 4)), (3, (4, 2, 6, 3)), (3, (2, 2, 2, 5))])
@pytest.mark.parametrize("prec", [F64, F32])
@pytest.mark.parametrize("G", [18, 12])
def test_small_lattices(oracle, nd, latt, prec, G):
    dt = np.float64 if prec == F64 else np.float32
    print(f"[parity small] nd={nd} latt={latt} prec={prec} G={G}", flush=True)
    u = gen_gauge_host(latt, nd, seed=1, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=1, field=0, dtype=dt)
    for isign in (1, -1):
        for cb in (0, 1):
            sentinel = np.full_like(psi, 5.5)
            got = run_gpu(u, psi, latt, nd, prec, G, isign, cb, sentinel)
            want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
            check(got, want, latt, cb, prec, f"nd={nd} {latt} isign={isign} cb={cb}")
            vh = volume(latt) // 2
            assert np.all(got.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == 5.5), "source cb of res modified"


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.rsplit("/", 1)[-1][:-4])
@pytest.mark.parametrize("prec,G", [(F64, 18), (F32, 18), (F64, 12), (F32, 12)])
def test_golden_fixtures(path, prec, G):
    """The HIP path against the committed oracle fixtures (tests/golden), without
    calling the oracle: both isign, both source checkerboards, 18- and 12-real
    links; the source checkerboard of res keeps its sentinel."""
    f = np.load(path)
    nd, latt = int(f["nd"]), tuple(int(v) for v in f["latt"])
    tag, dt = ("f64", np.float64) if prec == F64 else ("f32", np.float32)
    print(f"[golden] {path.rsplit('/', 1)[-1]} prec={prec} G={G}", flush=True)
    u, psi = f["u"].astype(dt), f["psi"].astype(dt)
    vh = volume(latt) // 2
    for isign in (1, -1):
        for cb in (0, 1):
            got = run_gpu(u, psi, latt, nd, prec, G, isign, cb, np.full_like(psi, 2.5))
            check(got, f[f"res_{tag}_{'p' if isign > 0 else 'm'}{cb}"], latt, cb, prec,
                  f"{path} isign={isign} cb={cb}")
            assert np.all(got.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == 2.5), "source cb of res modified"


@pytest.mark.parametrize("prec", [F64, F32])
def test_non_unitary_gauge_full_links(oracle, prec):
    """With 18-real links the operator is exact for any complex matrices (the
    reference never assumes unitarity); the force checks rely on this."""
    latt = (4, 2, 4, 6)
    rng = np.random.default_rng(4)
    V = volume(latt)
    dt = np.float64 if prec == F64 else np.float32
    u = rng.normal(size=(V, 4, 3, 3, 2)).astype(dt)
    psi = rng.normal(size=(V, 4, 3, 2)).astype(dt)
    for nd in (4, 3):
        for isign in (1, -1):
            for cb in (0, 1):
                got = run_gpu(u, psi, latt, nd, prec, 18, isign, cb)
                want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
                check(got, want, latt, cb, prec)


@pytest.mark.parametrize("nd", [4, 3])
@pytest.mark.parametrize("prec,G", [(F32, 18), (F64, 18), (F32, 12)])
def test_full_size_32c64(oracle, nd, prec, G):
    """Config: 32^3 x 64, both parities and both isign, against the oracle."""
    latt = (32, 32, 32, 64)
    dt = np.float64 if prec == F64 else np.float32
    print(f"[parity 32^3x64] nd={nd} prec={prec} G={G}", flush=True)
    u = gen_gauge_host(latt, nd, seed=1, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=1, field=0, dtype=dt)
    op = WilsonDslash(latt, nd=nd, prec=prec, gauge_reals=G)
    op.load_gauge(u)
    d_psi = op.pack_spinor(psi)
    d_res = op.new_spinor(zero=True)
    want = np.empty_like(psi)
    for isign in (1, -1):
        for cb in (0, 1):
            op.apply(d_res, d_psi, isign, cb)
            got = op.unpack_spinor(d_res, dtype=dt, cbs=(1 - cb,))
            oracle.dslash(u, psi, want, latt, isign, cb, nd=nd)
            e = check(got, want, latt, cb, prec, f"isign={isign} cb={cb}")
            print(f"  isign={isign:+d} cb={cb} rel_err={e:.3e}", flush=True)


@pytest.mark.parametrize("nd,latt", [(4, (4, 4, 4, 4)), (4, (6, 2, 2, 10)), (3, (4, 2, 6, 3)),
                                     (4, (16, 16, 16, 32)), (3, (16, 16, 16, 8))])
@pytest.mark.parametrize("G", [18, 12])
def test_sloppy_fp64(oracle, nd, latt, G):
    """F64S = FP64 spinors at the interface, FP32 links and FP32 arithmetic: the
    reference's SSEDSLASH_SLOPPY FP64 build (include/types64.h:26-31), which its
    test accepts at diff_norm < 1e-7 (tests/testDslashFull.cc:50-53, 103-109).
    Bound here: relative L2 and max error against the FP64 oracle <= 1e-6, and
    equal to the F32 kernel on the same (rounded) inputs to FP32 rounding."""
    print(f"[sloppy fp64] nd={nd} latt={latt} G={G}", flush=True)
    u = gen_gauge_host(latt, nd, seed=2)
    psi = gen_spinor_host(latt, nd, seed=2, field=1)
    vh = volume(latt) // 2
    for isign in (1, -1):
        for cb in (0, 1):
            sentinel = np.full_like(psi, 5.5)
            got = run_gpu(u, psi, latt, nd, F64S, G, isign, cb, sentinel)
            assert got.dtype == np.float64
            want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
            t = slice((1 - cb) * vh, (2 - cb) * vh)
            g, w = got.reshape(-1, 24)[t], want.reshape(-1, 24)[t]
            e2, einf = R.rel_err(g, w), R.max_rel_err(g, w)
            dn = np.sqrt(np.sum((g - w) ** 2)) / (24 * vh)
            print(f"  isign={isign:+d} cb={cb} rel L2 {e2:.2e} rel max {einf:.2e} diff_norm {dn:.2e}", flush=True)
            assert e2 < 1e-6 and einf < 1e-6 and dn < 1e-7
            assert e2 > 1e-12, "sloppy result is suspiciously exact: FP32 path not taken"
            assert np.all(got.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == 5.5), "source cb of res modified"
            f32 = run_gpu(u.astype(np.float32), psi.astype(np.float32), latt, nd, F32, G, isign, cb)
            d = np.max(np.abs(g - f32.reshape(-1, 24)[t].astype(np.float64))) / np.max(np.abs(w))
            assert d < 1e-6, f"F64S vs F32 kernel: {d:.2e}"


@pytest.mark.parametrize("prec,tol", [(F16, 2e-3), (BF16, 1.6e-2)])
def test_half_storage_operator_error(oracle, prec, tol):
    """16-bit spinors (per-site norm) and 12-real 16-bit links, FP32 math:
    error against the FP64 oracle is bounded by the storage rounding."""
    latt = (8, 8, 8, 8)
    u = gen_gauge_host(latt, 4, seed=1)
    psi = gen_spinor_host(latt, 4, seed=1)
    for isign in (1, -1):
        for cb in (0, 1):
            got = run_gpu(u, psi, latt, 4, prec, 12, isign, cb)
            want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb)
            vh = volume(latt) // 2
            t = slice((1 - cb) * vh, (2 - cb) * vh)
            e = R.rel_err(got.reshape(-1, 24)[t], want.reshape(-1, 24)[t])
            print(f"  prec={prec} isign={isign:+d} cb={cb} rel_err vs FP64 oracle = {e:.3e}", flush=True)
            assert e < tol


def test_device_generators_match_host():
    latt = (8, 4, 6, 4)
    op = WilsonDslash(latt, nd=4, prec=F64)
    for nd in (4, 3):
        lib = op.lib
        from sse_wilson_dslash_amd._lib import latt_arg, ptr, check as ck
        V = volume(latt)
        d = torch.empty(V * 72, dtype=torch.float64, device="cuda")
        ck(lib.dslash_gen_gauge_device(ptr(d), F64, latt_arg(latt), nd, 1, None))
        h = gen_gauge_host(latt, nd, seed=1)
        assert np.max(np.abs(d.cpu().numpy() - h.ravel())) < 1e-13
        s = torch.empty(V * 24, dtype=torch.float64, device="cuda")
        ck(lib.dslash_gen_spinor_device(ptr(s), F64, latt_arg(latt), nd, 1, 3, None))
        hs = gen_spinor_host(latt, nd, seed=1, field=3)
        assert np.max(np.abs(s.cpu().numpy() - hs.ravel())) < 1e-12


@pytest.mark.parametrize("block,remap", [(64, 0), (128, 1), (192, 1), (256, 0)])
def test_launch_shapes_agree(oracle, block, remap):
    from sse_wilson_dslash_amd import _lib
    latt = (8, 6, 4, 10)
    u = gen_gauge_host(latt, 4, seed=2, dtype=np.float32)
    psi = gen_spinor_host(latt, 4, seed=2, dtype=np.float32)
    import ctypes as C
    b, r = C.c_int(), C.c_int()
    _lib.check(_lib.core().dslash_get_launch(4, F32, 18, C.byref(b), C.byref(r)))
    _lib.check(_lib.core().dslash_set_launch(4, F32, 18, block, remap))
    try:
        got = run_gpu(u, psi, latt, 4, F32, 18, 1, 0)
    finally:
        _lib.check(_lib.core().dslash_set_launch(4, F32, 18, b.value, r.value))
    want = oracle.dslash(u, psi, np.zeros_like(psi), latt, 1, 0)
    check(got, want, latt, 0, F32)


MARCH_TOL = {F64: 1e-13, F32: 1e-5, F64S: 1e-6, F16: 2e-3, BF16: 1.6e-2}
# Largest difference between two marching lengths, relative to max |res|: FP32
# arithmetic may contract differently in the two kernels, which for 16-bit
# results can move a stored value by one unit of its format (2^-10, 2^-7 of the
# site's norm).
MARCH_SAME = {F64: 1e-13, F32: 1e-6, F64S: 1e-6, F16: 2.0 ** -10, BF16: 2.0 ** -7}


def _march_for(latt, nd, prec, G):
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    m = C.c_int()
    _lib.check(_lib.core().dslash_launch_march_for(_lib.latt_arg(latt), nd, prec, G, C.byref(m)))
    return m.value


class _March:
    """Sets the marching length of (nd, prec, G) and restores it on exit."""

    def __init__(self, nd, prec, G):
        import ctypes as C
        from sse_wilson_dslash_amd import _lib
        self.lib, self.check, self.key = _lib.core(), _lib.check, (nd, prec, G)
        m = C.c_int()
        self.check(self.lib.dslash_get_launch_march(nd, prec, G, C.byref(m)))
        self.saved = m.value

    def set(self, march):
        self.check(self.lib.dslash_set_launch_march(*self.key, march))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.set(self.saved)


@pytest.mark.parametrize("nd,latt", [(4, (4, 6, 4, 10)), (4, (2, 2, 2, 2)), (3, (4, 2, 6, 3)), (3, (2, 4, 10, 2))])
@pytest.mark.parametrize("prec", [F64, F32, F16])
def test_marching_falls_back_on_small_lattices(oracle, nd, latt, prec):
    """A marching length set for the format is not used where the launch would
    have too few threads to fill the device: these lattices run one site per
    thread whatever is set, and agree with the oracle."""
    print(f"[march small] nd={nd} latt={latt} prec={prec}", flush=True)
    u = gen_gauge_host(latt, nd, seed=3)
    psi = gen_spinor_host(latt, nd, seed=3, field=2)
    if prec != F64:
        u, psi = u.astype(np.float32), psi.astype(np.float32)
    vh = volume(latt) // 2
    with _March(nd, prec, 12) as m:
        for march in (4, 8):
            m.set(march)
            assert _march_for(latt, nd, prec, 12) == 1
            for isign, cb in ((1, 0), (-1, 1)):
                got = run_gpu(u, psi, latt, nd, prec, 12, isign, cb)
                want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
                t = slice((1 - cb) * vh, (2 - cb) * vh)
                assert R.rel_err(got.reshape(-1, 24)[t], want.reshape(-1, 24)[t]) < MARCH_TOL[prec]


# Lattices on which marching runs: >= 131072 threads at march 8.  t = 66 and
# z = 34 leave a 2-slice last chunk at march 4 and 8; t = 2 and z = 2 are shorter
# than the march; t = 65 is an odd time extent of the 3D operator.
MARCH_CASES = [(4, (32, 32, 32, 66), F64, 18), (4, (32, 32, 32, 66), F32, 12), (4, (32, 32, 32, 66), BF16, 12),
               (4, (24, 48, 48, 48), F32, 18),  # box tiles of 12 x 3 x 6 sites: 40 idle lanes per block
               (4, (48, 48, 48, 24), F64, 18),  # box tiles of 24 x 3 x 3 sites: 40 idle lanes per block
               (4, (64, 64, 64, 2), F64S, 18), (4, (64, 64, 64, 2), F16, 12),
               (4, (32, 128, 256, 2), F32, 18),  # halo tiles (16 x 4 x 4 boxes) on a 2-slice t extent
               (4, (32, 2, 3280, 10), F64S, 18),  # 16 x 2 x 8 boxes span y: z faces only; 8 + 2 slices
               (3, (32, 32, 34, 65), F64, 12), (3, (32, 32, 34, 65), F32, 18), (3, (32, 32, 34, 65), F16, 18),
               (3, (64, 64, 2, 64), BF16, 18), (3, (64, 64, 2, 64), F64S, 12)]


@pytest.mark.parametrize("nd,latt,prec,G", MARCH_CASES)
def test_marching_lengths(oracle, nd, latt, prec, G):
    """res = D psi with a thread walking 1, 4 or 8 slices of the slowest
    coordinate (t in 4D, z in 3D), and for the 4D FP32, FP64 and sloppy-FP64
    operators with LDS-staged box tiles (march 16, 256 threads): every shape agrees with the oracle and
    with the one-site-per-thread result, and leaves the source checkerboard of
    res untouched."""
    print(f"[march] nd={nd} latt={latt} prec={prec} G={G}", flush=True)
    u = gen_gauge_host(latt, nd, seed=3)
    psi = gen_spinor_host(latt, nd, seed=3, field=2)
    if prec == F32:
        u, psi = u.astype(np.float32), psi.astype(np.float32)
    dt = np.float32 if prec in (F32, F16, BF16) else np.float64
    vh = volume(latt) // 2
    op = WilsonDslash(latt, nd=nd, prec=prec, gauge_reals=G)
    op.load_gauge(u)
    d_psi = op.pack_spinor(psi)
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    lib = _lib.core()
    b0, r0 = C.c_int(), C.c_int()
    _lib.check(lib.dslash_get_launch(nd, prec, G, C.byref(b0), C.byref(r0)))
    marches = (1, 4, 8, 16, 32) if (nd == 4 and prec in (F32, F64S)) else \
        (1, 4, 8, 16) if (nd == 4 and prec == F64) else (1, 4, 8)
    with _March(nd, prec, G) as m:
        for isign, cb in ((1, 0), (-1, 1)):
            want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
            t = slice((1 - cb) * vh, (2 - cb) * vh)
            w = want.reshape(-1, 24)[t]
            first = None
            for march in marches:
                m.set(march)
                _lib.check(lib.dslash_set_launch(nd, prec, G, 256 if march >= 16 else b0.value, r0.value))
                ran = _march_for(latt, nd, prec, G)
                # 32 (box tiles with the LDS halo) falls back to 16 where the halo would
                # need more than one entry per thread (64^3 x 2: 384 > 256)
                assert ran == march or (march == 32 and ran == 16 and latt == (64, 64, 64, 2)), (march, ran)
                d_res = op.pack_spinor(np.full_like(psi, 3.25))
                op.apply(d_res, d_psi, isign, cb)
                got = op.unpack_spinor(d_res, dtype=dt).reshape(-1, 24)
                g = got[t]
                e2, einf = R.rel_err(g, w), R.max_rel_err(g, w)
                print(f"  isign={isign:+d} cb={cb} march={march}: rel L2 {e2:.2e} rel max {einf:.2e}", flush=True)
                assert e2 < MARCH_TOL[prec], (march, isign, cb, e2)
                if prec in (F64, F32, F64S):
                    assert einf < MARCH_TOL[prec], (march, isign, cb, einf)
                assert np.all(got[cb * vh:(cb + 1) * vh] == 3.25), "source cb of res modified"
                if first is None:
                    first = g
                else:
                    assert np.max(np.abs(g - first)) <= MARCH_SAME[prec] * np.max(np.abs(first)), \
                        (march, isign, cb)
    _lib.check(lib.dslash_set_launch(nd, prec, G, b0.value, r0.value))
    del op, d_psi
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", [F64, F32])
def test_xpay_epilogue(oracle, prec):
    """dslash_apply_xpay: res = x + a D psi in one pass (and in place, x aliasing res)."""
    from sse_wilson_dslash_amd._lib import check, ptr
    latt = (4, 6, 4, 8)
    dt = np.float64 if prec == F64 else np.float32
    u = gen_gauge_host(latt, 4, seed=3, dtype=dt)
    psi = gen_spinor_host(latt, 4, seed=3, field=1, dtype=dt)
    x = gen_spinor_host(latt, 4, seed=3, field=2, dtype=dt)
    a = -0.0144
    op = WilsonDslash(latt, 4, prec, 18)
    op.load_gauge(u)
    vh = op.vh
    for isign in (1, -1):
        for tcb in (0, 1):
            d_psi = op.pack_cb(psi, 1 - tcb)
            d_x = op.pack_cb(x, tcb)
            d_res = op.new_spinor(False)
            g = ptr(op.gauge) + tcb * op.gb
            check(op.lib.dslash_apply_xpay(ptr(d_res), ptr(d_psi), g, ptr(d_x), a, op._L, prec, 18, isign, tcb,
                                           torch.cuda.current_stream().cuda_stream))
            d_in = d_x.clone()
            check(op.lib.dslash_apply_xpay(ptr(d_in), ptr(d_psi), g, ptr(d_in), a, op._L, prec, 18, isign, tcb,
                                           torch.cuda.current_stream().cuda_stream))
            got = op.unpack_cb(d_res, tcb, dtype=dt)
            got_inplace = op.unpack_cb(d_in, tcb, dtype=dt)
            d = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, 1 - tcb)
            want = x + a * d
            t = slice(tcb * vh, (tcb + 1) * vh)
            assert R.rel_err(got[t], want[t]) < TOL[prec]
            assert np.array_equal(got_inplace, got)


# t = 66 leaves a 2-slice last tile; 24 x 48 x 48 boxes (12 x 3 x 6 sites) leave
# 40 lanes of each 256-thread block idle; t = 2 is shorter than a tile.
XPAY_BOX_CASES = [((32, 32, 32, 66), F64, 18), ((24, 48, 48, 48), F32, 18), ((64, 64, 64, 2), F64S, 12),
                  ((32, 32, 32, 66), F32, 12)]


@pytest.mark.parametrize("latt,prec,G", XPAY_BOX_CASES)
def test_xpay_epilogue_box_tiles(oracle, latt, prec, G):
    """dslash_apply_xpay on lattices where the 4D operator runs LDS-staged box
    tiles (march 16; for FP32 also with the LDS halo, march 32): res = x + a D psi
    agrees with the oracle and with the one-site-per-thread kernel (march 1), and
    the in-place form (x aliasing res) gives the same bits as the out-of-place one."""
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    from sse_wilson_dslash_amd._lib import ptr
    print(f"[xpay box] latt={latt} prec={prec} G={G}", flush=True)
    dt = np.float32 if prec == F32 else np.float64
    u = gen_gauge_host(latt, 4, seed=4, dtype=dt)
    psi = gen_spinor_host(latt, 4, seed=4, field=1, dtype=dt)
    x = gen_spinor_host(latt, 4, seed=4, field=2, dtype=dt)
    a = -0.0144
    op = WilsonDslash(latt, 4, prec, G)
    op.load_gauge(u)
    vh = op.vh
    lib = _lib.core()
    b0, r0 = C.c_int(), C.c_int()
    _lib.check(lib.dslash_get_launch(4, prec, G, C.byref(b0), C.byref(r0)))
    stream = torch.cuda.current_stream().cuda_stream
    try:
        with _March(4, prec, G) as m:
            for isign, tcb in ((1, 0), (-1, 1)):
                d_psi = op.pack_cb(psi, 1 - tcb)
                d_x = op.pack_cb(x, tcb)
                g = ptr(op.gauge) + tcb * op.gb
                t = slice(tcb * vh, (tcb + 1) * vh)
                want = (x + a * oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, 1 - tcb)).reshape(-1, 24)[t]
                outs = {}
                for march in ((32, 16, 1) if prec in (F32, F64S) else (16, 1)):
                    m.set(march)
                    _lib.check(lib.dslash_set_launch(4, prec, G, 256, 1))
                    ran = _march_for(latt, 4, prec, G)
                    assert ran == march or (march == 32 and ran == 16 and latt == (64, 64, 64, 2)), (march, ran)
                    d_res = op.new_spinor(False)
                    _lib.check(lib.dslash_apply_xpay(ptr(d_res), ptr(d_psi), g, ptr(d_x), a, op._L, prec, G, isign,
                                                     tcb, stream))
                    d_in = d_x.clone()
                    _lib.check(lib.dslash_apply_xpay(ptr(d_in), ptr(d_psi), g, ptr(d_in), a, op._L, prec, G, isign,
                                                     tcb, stream))
                    got = op.unpack_cb(d_res, tcb, dtype=dt).reshape(-1, 24)[t]
                    got_in = op.unpack_cb(d_in, tcb, dtype=dt).reshape(-1, 24)[t]
                    e2, einf = R.rel_err(got, want), R.max_rel_err(got, want)
                    print(f"  isign={isign:+d} tcb={tcb} march={march}: rel L2 {e2:.2e} rel max {einf:.2e}",
                          flush=True)
                    assert e2 < MARCH_TOL[prec] and einf < MARCH_TOL[prec], (march, isign, tcb, e2, einf)
                    assert np.array_equal(got_in, got), (march, isign, tcb)
                    outs[march] = got
                for mb in outs:
                    assert np.max(np.abs(outs[mb] - outs[1])) <= MARCH_SAME[prec] * np.max(np.abs(outs[1])), mb
    finally:
        _lib.check(lib.dslash_set_launch(4, prec, G, b0.value, r0.value))
    del op
    torch.cuda.empty_cache()


@pytest.mark.parametrize("p16,tol", [(F16, 6e-4), (BF16, 4e-3)])
def test_precision_conversion_round_trip(p16, tol):
    """F64 -> 16-bit (per-site norm) -> F64 keeps every site to the format's rounding
    relative to that site's largest component; F64 -> F32 -> F64 to 6e-8."""
    from sse_wilson_dslash_amd._lib import check, ptr
    latt = (8, 4, 4, 6)
    psi = gen_spinor_host(latt, 4, seed=9)
    psi[5] *= 1e-30  # a tiny site: the per-site norm keeps its relative precision
    psi[7] = 0.0     # an all-zero site
    op = WilsonDslash(latt, 4, F64)
    d64 = op.pack_cb(psi, 0)
    for pl, t in ((p16, tol), (F32, 6e-8)):
        d_lo = op.new_spinor(False, prec=pl)
        back = op.new_spinor(False)
        s = torch.cuda.current_stream().cuda_stream
        check(op.lib.dslash_convert_spinor(ptr(d_lo), pl, ptr(d64), F64, op._L, 4, s))
        check(op.lib.dslash_convert_spinor(ptr(back), F64, ptr(d_lo), pl, op._L, 4, s))
        got = op.unpack_cb(back, 0).reshape(-1, 24)[: op.vh]
        ref = psi.reshape(-1, 24)[: op.vh]
        scale = np.maximum(np.max(np.abs(ref), axis=1, keepdims=True), 1e-300)
        assert np.max(np.abs(got - ref) / scale) < t
        assert np.all(got[7] == 0.0)


@pytest.mark.parametrize("nd,latt,prec", [(4, (8, 4, 6, 10), F64), (4, (8, 4, 6, 10), F32), (3, (4, 6, 8, 5), F32),
                                          (4, (8, 8, 8, 8), F16)])
def test_site_ranges(nd, latt, prec):
    """dslash_apply_sites / dslash_pack_spinor_sites / dslash_unpack_spinor_sites
    (the pieces of the drop-in slab pipeline): pieces over consecutive site
    ranges, including empty and single-site ones, give the same bytes as the
    whole-checkerboard calls, and sites outside a range are not written."""
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    lib, L = _lib.core(), _lib.latt_arg(latt)
    u = gen_gauge_host(latt, nd, seed=8)
    psi = gen_spinor_host(latt, nd, seed=8)
    op = WilsonDslash(latt, nd, prec, 18)
    op.load_gauge(u)
    vh = op.vh
    st = torch.cuda.current_stream().cuda_stream
    d_aos = torch.from_numpy(psi.reshape(-1)).cuda()
    cuts = sorted({0, 1, 1, vh // 3, vh // 3, vh // 2 + 1, vh - 1, vh})
    for cb in (0, 1):
        tcb = 1 - cb
        whole = op.new_spinor(full=False)
        _lib.check(lib.dslash_pack_spinor(_lib.ptr(whole), cb, _lib.ptr(d_aos), F64, L, nd, prec, None, st))
        parts = torch.full_like(whole, 0x3C)
        for a, b in zip(cuts[:-1], cuts[1:]):
            _lib.check(lib.dslash_pack_spinor_sites(_lib.ptr(parts), cb, _lib.ptr(d_aos), F64, L, nd, prec, None,
                                                    a, b, st))
        torch.cuda.synchronize()
        assert torch.equal(parts, whole)
        g = _lib.ptr(op.gauge) + tcb * op.gb
        for isign in (1, -1):
            ref = op.new_spinor(full=False, zero=True)
            _lib.check(lib.dslash_apply_sites(_lib.ptr(ref), _lib.ptr(whole), g, L, nd, prec, 18, isign, tcb, 0, vh,
                                              st))
            got = torch.full_like(ref, 0x5A)
            sentinel = got.clone()
            a, b = cuts[1], cuts[3]
            _lib.check(lib.dslash_apply_sites(_lib.ptr(got), _lib.ptr(whole), g, L, nd, prec, 18, isign, tcb, a, b,
                                              st))
            torch.cuda.synchronize()
            # the planes of sites [a, b) hold the result, the rest is untouched
            nvec = {F64: 12, F32: 6, F16: 3}[prec]
            pl = got[:nvec * 16 * vh].view(nvec, vh, 16)
            pr = ref[:nvec * 16 * vh].view(nvec, vh, 16)
            ps = sentinel[:nvec * 16 * vh].view(nvec, vh, 16)
            assert torch.equal(pl[:, a:b], pr[:, a:b])
            assert torch.equal(pl[:, :a], ps[:, :a]) and torch.equal(pl[:, b:], ps[:, b:])
            for a2, b2 in zip(cuts[:-1], cuts[1:]):
                _lib.check(lib.dslash_apply_sites(_lib.ptr(got), _lib.ptr(whole), g, L, nd, prec, 18, isign, tcb, a2,
                                                  b2, st))
            torch.cuda.synchronize()
            assert torch.equal(got, ref)
            # the full-lattice call (any launch shape) agrees to rounding
            full = op.new_spinor(full=False, zero=True)
            op.apply_cb(full, whole, isign, tcb)
            torch.cuda.synchronize()
            x = op.unpack_cb(full, tcb)
            y = op.unpack_cb(ref, tcb)
            assert R.rel_err(y, x) < (1e-13 if prec == F64 else 1e-5 if prec == F32 else 1e-3)
            out_whole = torch.zeros(op.V * 24, dtype=torch.float64, device="cuda")
            out_parts = torch.zeros_like(out_whole)
            _lib.check(lib.dslash_unpack_spinor(_lib.ptr(out_whole), tcb, _lib.ptr(ref), prec, L, nd, F64, None, 0,
                                                st))
            for a2, b2 in zip(cuts[:-1], cuts[1:]):
                _lib.check(lib.dslash_unpack_spinor_sites(_lib.ptr(out_parts), tcb, _lib.ptr(ref), prec, L, nd, F64,
                                                          None, 0, a2, b2, st))
            torch.cuda.synchronize()
            assert torch.equal(out_parts, out_whole)
    # ranges outside the checkerboard are rejected before any launch
    t = op.new_spinor(full=False)
    assert lib.dslash_apply_sites(_lib.ptr(t), _lib.ptr(t) + 16, _lib.ptr(op.gauge), L, nd, prec, 18, 1, 0, 0,
                                  vh + 1, st) == _lib.ERR_INVALID
    assert lib.dslash_pack_spinor_sites(_lib.ptr(t), 0, _lib.ptr(d_aos), F64, L, nd, prec, None, 5, 4,
                                        st) == _lib.ERR_INVALID


def test_ceiling_kernel_runs():
    """dslash_time_ceiling moves the headline kernel's 768 B/site: it runs, its
    output planes are the xor of the 6 + 36 input planes at each site, and bad
    shapes are rejected."""
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    lib = _lib.core()
    latt = (8, 8, 8, 16)
    vh = volume(latt) // 2
    g = torch.Generator(device="cpu").manual_seed(3)
    a = torch.randint(0, 2**31 - 1, (6, vh, 4), generator=g, dtype=torch.int64).to(torch.int32).cuda()
    links = torch.randint(0, 2**31 - 1, (36, vh, 4), generator=g, dtype=torch.int64).to(torch.int32).cuda()
    out = torch.zeros_like(a)
    ms = C.c_float()
    _lib.check(lib.dslash_time_ceiling(_lib.ptr(out), _lib.ptr(a), _lib.ptr(links), _lib.latt_arg(latt), 64, 1, 3,
                                       torch.cuda.current_stream().cuda_stream, C.byref(ms)))
    assert ms.value > 0
    want = a.clone()
    for k in range(36):
        want[k % 6] ^= links[k]
    assert torch.equal(out, want)
    assert lib.dslash_time_ceiling(_lib.ptr(out), _lib.ptr(a), _lib.ptr(links), _lib.latt_arg(latt), 96, 1, 3, None,
                                   C.byref(ms)) == _lib.ERR_INVALID
