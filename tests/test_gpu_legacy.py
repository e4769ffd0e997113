"""The drop-in API (sse_dslash.h / sse_dslash_3d.h) through libsse_dslash_f32.so /
libsse_dslash_f64.so, as tests/testDslashFull.cc:61-112 and
tests/testDslash3D.cc:64-113 exercise the reference: a reunitarised Gaussian
gauge, a Gaussian source, both isign and both source checkerboards, compared
against an independent Dslash (here the oracle)."""
import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.layout import volume
from sse_wilson_dslash_amd.legacy import SSEDslash, table_callbacks
from tests import refimpl as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def libs():
    return {32: SSEDslash(32), 64: SSEDslash(64)}


def diff_norm(a, b, latt):
    """tests/testDslashFull.cc:103-104: sqrt(norm2(diff)) / (4*3*2*V/2)."""
    V = volume(latt)
    return np.sqrt(np.sum((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)) / (24 * V / 2)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nd", [4, 3])
def test_legacy_canonical(oracle, libs, prec, nd):
    lg = libs[prec]
    dt = lg.dtype
    latt = (8, 8, 8, 8) if nd == 4 else (8, 4, 6, 5)
    print(f"[legacy canonical] prec={prec} nd={nd}", flush=True)
    u = gen_gauge_host(latt, nd, seed=1, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=1, dtype=dt)
    init = lg.init_sse_su3dslash if nd == 4 else lg.init_sse_su3dslash_3d
    free = lg.free_sse_su3dslash if nd == 4 else lg.free_sse_su3dslash_3d
    apply = lg.sse_su3dslash_wilson if nd == 4 else lg.sse_su3dslash_wilson_3d
    init(latt)
    try:
        vh = volume(latt) // 2
        for isign in (1, -1):
            for cb in (0, 1):
                res = np.full_like(psi, -2.0)
                apply(u, psi, res, isign, cb)
                want = oracle.dslash(u, psi, np.full_like(psi, -2.0), latt, isign, cb, nd=nd)
                # the reference's own acceptance (1e-7 FP32, 1e-16 FP64 diff_norm) ...
                assert diff_norm(res, want, latt) < (1e-7 if prec == 32 else 1e-16)
                # ... and this project's relative bound
                assert R.rel_err(res, want) < (1e-5 if prec == 32 else 1e-13)
                assert np.all(res.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == -2.0)
        err = lg.last_error() if nd == 4 else lg.last_error_3d()
        assert err == 0
    finally:
        free()


@pytest.mark.parametrize("nd", [4, 3])
def test_legacy_sloppy_fp64(oracle, libs, nd):
    """The FP64 library in sloppy mode (FP32 links and arithmetic behind FP64
    arrays) meets the reference's SSEDSLASH_SLOPPY acceptance, diff_norm < 1e-7
    (tests/testDslashFull.cc:50-53), on the host, resident and 12-real paths;
    switching back restores FP64 parity."""
    lg = libs[64]
    latt = (8, 8, 8, 8) if nd == 4 else (8, 4, 6, 5)
    print(f"[legacy sloppy] nd={nd}", flush=True)
    u = gen_gauge_host(latt, nd, seed=3)
    psi = gen_spinor_host(latt, nd, seed=3)
    init = lg.init_sse_su3dslash if nd == 4 else lg.init_sse_su3dslash_3d
    free = lg.free_sse_su3dslash if nd == 4 else lg.free_sse_su3dslash_3d
    apply = lg.sse_su3dslash_wilson if nd == 4 else lg.sse_su3dslash_wilson_3d
    sloppy = lg.set_sloppy if nd == 4 else lg.set_sloppy_3d
    upload = lg.gauge_upload if nd == 4 else lg.gauge_upload_3d
    resident = lg.wilson_resident if nd == 4 else lg.wilson_3d_resident
    init(latt)
    try:
        assert sloppy(True) == 0
        for path in ("host", "resident18", "resident12"):
            if path != "host":
                assert upload(u, 18 if path == "resident18" else 12) == 0
            for isign in (1, -1):
                for cb in (0, 1):
                    res = np.zeros_like(psi)
                    if path == "host":
                        apply(u, psi, res, isign, cb)
                    else:
                        assert resident(psi, res, isign, cb) == 0
                    want = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd)
                    e = R.rel_err(res, want)
                    assert diff_norm(res, want, latt) < 1e-7 and 1e-12 < e < 1e-6, (path, isign, cb, e)
        assert sloppy(False) == 0
        # the resident gauge was packed for the sloppy format: it is dropped
        assert resident(psi, np.zeros_like(psi), 1, 0) != 0
        assert upload(u, 18) == 0
        res = np.zeros_like(psi)
        assert resident(psi, res, 1, 0) == 0
        want = oracle.dslash(u, psi, np.zeros_like(psi), latt, 1, 0, nd=nd)
        assert R.rel_err(res, want) < 1e-13
    finally:
        sloppy(False)
        free()


@pytest.mark.parametrize("nd", [4, 3])
def test_legacy_caller_order_callbacks(oracle, libs, nd):
    """A non-canonical caller order supplied through the callbacks."""
    lg = libs[64]
    latt = (4, 4, 2, 6) if nd == 4 else (4, 2, 4, 3)
    V = volume(latt)
    perm = np.random.default_rng(5).permutation(V)  # my canonical index -> caller index
    cbs = table_callbacks(latt, perm, nd)
    u = gen_gauge_host(latt, nd, seed=3)
    psi = gen_spinor_host(latt, nd, seed=3)
    u_c = np.empty_like(u); u_c[perm] = u
    p_c = np.empty_like(psi); p_c[perm] = psi
    init = lg.init_sse_su3dslash if nd == 4 else lg.init_sse_su3dslash_3d
    free = lg.free_sse_su3dslash if nd == 4 else lg.free_sse_su3dslash_3d
    apply = lg.sse_su3dslash_wilson if nd == 4 else lg.sse_su3dslash_wilson_3d
    init(latt, *cbs)
    try:
        for isign in (1, -1):
            for cb in (0, 1):
                res_c = np.full_like(psi, 9.0)
                apply(u_c, p_c, res_c, isign, cb)
                want = oracle.dslash(u, psi, np.full_like(psi, 9.0), latt, isign, cb, nd=nd)
                assert R.rel_err(res_c[perm], want) < 1e-13
    finally:
        free()


@pytest.mark.parametrize("prec,nd,latt,order", [(32, 4, (32, 32, 32, 66), "canonical"),
                                                (64, 3, (32, 32, 34, 33), "canonical"),
                                                (64, 4, (16, 16, 40, 36), "permuted")])
def test_legacy_staged_transfers(oracle, libs, prec, nd, latt, order):
    """Host arrays of 64 MiB and more move through the pinned double-buffered
    staging (32 MiB chunks; these sizes leave a partial last chunk: gauge
    594 / 631 / 202.5 MiB, source half 99 / 105 MiB, permuted full source
    67.5 MiB; the permuted case's 34 MiB result takes the direct copy)."""
    lg = libs[prec]
    dt = lg.dtype
    V = volume(latt)
    print(f"[legacy staged] prec={prec} nd={nd} latt={latt} {order}", flush=True)
    u = gen_gauge_host(latt, nd, seed=4, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=4, dtype=dt)
    init = lg.init_sse_su3dslash if nd == 4 else lg.init_sse_su3dslash_3d
    free = lg.free_sse_su3dslash if nd == 4 else lg.free_sse_su3dslash_3d
    apply = lg.sse_su3dslash_wilson if nd == 4 else lg.sse_su3dslash_wilson_3d
    if order == "permuted":
        perm = np.random.default_rng(6).permutation(V)
        init(latt, *table_callbacks(latt, perm, nd))
        u_c = np.empty_like(u); u_c[perm] = u
        p_c = np.empty_like(psi); p_c[perm] = psi
    else:
        perm = None
        init(latt)
        u_c, p_c = u, psi
    try:
        vh = V // 2
        for isign, cb in ((1, 0), (-1, 1)):
            res = np.full_like(psi, -3.0)
            apply(u_c, p_c, res, isign, cb)
            if perm is not None:
                res = res[perm]
            want = oracle.dslash(u, psi, np.full_like(psi, -3.0), latt, isign, cb, nd=nd)
            e = R.rel_err(res, want)
            print(f"  isign={isign:+d} cb={cb}: rel L2 {e:.2e}", flush=True)
            assert e < (1e-5 if prec == 32 else 1e-13)
            assert R.max_rel_err(res, want) < (1e-5 if prec == 32 else 1e-13)
            assert np.all(res.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == -3.0)
        assert (lg.last_error() if nd == 4 else lg.last_error_3d()) == 0
    finally:
        free()


def test_legacy_exit_without_free():
    """A caller that initialises the library and exits without free_sse_su3dslash
    (the reference allows it) ends with status 0: the staging workers are stopped
    at exit, not left running (which would terminate the process)."""
    import os
    import subprocess
    import sys
    code = ("import numpy as np\n"
            "from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host\n"
            "from sse_wilson_dslash_amd.legacy import SSEDslash\n"
            "lg = SSEDslash(32)\n"
            "latt = (8, 8, 8, 8)\n"
            "u = gen_gauge_host(latt, 4, seed=1, dtype=np.float32)\n"
            "psi = gen_spinor_host(latt, 4, seed=1, dtype=np.float32)\n"
            "lg.init_sse_su3dslash(latt)\n"
            "lg.init_sse_su3dslash_3d(latt)\n"
            "res = np.zeros_like(psi)\n"
            "lg.sse_su3dslash_wilson(u, psi, res, 1, 0)\n"
            "assert lg.last_error() == 0 and np.abs(res).max() > 0\n"
            "print('applied', flush=True)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "applied" in r.stdout, (r.returncode, r.stdout, r.stderr[-2000:])


def test_legacy_semantics(oracle, libs):
    """Aliasing psi == res, isign no-op, refcounting, resident-gauge paths."""
    lg = libs[32]
    latt = (4, 4, 4, 4)
    u = gen_gauge_host(latt, 4, seed=1, dtype=np.float32)
    psi = gen_spinor_host(latt, 4, seed=1, dtype=np.float32)
    lg.init_sse_su3dslash(latt)
    lg.init_sse_su3dslash((6, 6, 6, 6))  # refcount 2: size ignored (lib/sse_su3dslash_32bit_scalar.c:95-99)
    try:
        want = oracle.dslash(u, psi, psi.copy(), latt, 1, 0)
        inplace = psi.copy()
        lg.sse_su3dslash_wilson(u, inplace, inplace, 1, 0)
        assert R.rel_err(inplace, want) < 1e-5
        before = inplace.copy()
        lg.sse_su3dslash_wilson(u, psi, inplace, 0, 0)  # isign not +-1: nothing happens
        lg.sse_su3dslash_wilson(u, psi, inplace, 3, 1)
        assert np.array_equal(inplace, before)
        lg.sse_su3dslash_prepost_receives()
        lg.free_sse_su3dslash()  # refcount 1: still usable
        res = np.zeros_like(psi)
        lg.sse_su3dslash_wilson(u, psi, res, -1, 1)
        assert R.rel_err(res, oracle.dslash(u, psi, np.zeros_like(psi), latt, -1, 1)) < 1e-5
        # resident gauge, 18 and 12 reals
        for G in (18, 12):
            assert lg.gauge_upload(u, G) == 0
            for isign in (1, -1):
                for cb in (0, 1):
                    r2 = np.zeros_like(psi)
                    assert lg.wilson_resident(psi, r2, isign, cb) == 0
                    w = oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb)
                    assert R.rel_err(r2, w) < 1e-5
        # device-resident spinors in the library layout
        from sse_wilson_dslash_amd import F32
        from sse_wilson_dslash_amd.device import WilsonDslash
        op = WilsonDslash(latt, 4, F32, 12)
        d_psi = op.pack_spinor(psi)
        d_res = op.new_spinor(zero=True)
        assert lg.wilson_device(d_res, d_psi, 1, 1, torch.cuda.current_stream().cuda_stream) == 0
        torch.cuda.synchronize()
        got = op.unpack_spinor(d_res, dtype=np.float32)
        assert R.rel_err(got, oracle.dslash(u, psi, np.zeros_like(psi), latt, 1, 1)) < 1e-5
    finally:
        lg.free_sse_su3dslash()
    # after the last free the API reports use-before-init instead of crashing
    assert lg.wilson_resident(psi, np.zeros_like(psi), 1, 0) != 0


def test_device_call_then_upload_on_side_stream(libs):
    """sse_su3dslash_wilson_device on a non-default stream returns before its
    kernels run; a gauge upload issued right after must not rewrite the
    resident links under them.  Many queued applications of gauge A, then an
    upload of gauge B: the result still equals D(A) psi bit for bit, and the
    next application uses B."""
    from sse_wilson_dslash_amd import F32
    from sse_wilson_dslash_amd.device import WilsonDslash
    lg = libs[32]
    latt = (32, 32, 32, 64)
    print(f"[legacy side stream] latt={latt}", flush=True)
    ua = gen_gauge_host(latt, 4, seed=1, dtype=np.float32)
    ub = gen_gauge_host(latt, 4, seed=2, dtype=np.float32)
    psi = gen_spinor_host(latt, 4, seed=1, dtype=np.float32)
    op = WilsonDslash(latt, 4, F32, 18)
    d_psi = op.pack_spinor(psi)
    lg.init_sse_su3dslash(latt)
    try:
        assert lg.gauge_upload(ua, 18) == 0
        want_a = op.new_spinor(zero=True)
        assert lg.wilson_device(want_a, d_psi, 1, 0, torch.cuda.current_stream().cuda_stream) == 0
        torch.cuda.synchronize()
        side = torch.cuda.Stream()
        d_res = op.new_spinor(zero=True)
        torch.cuda.synchronize()
        for _ in range(200):
            assert lg.wilson_device(d_res, d_psi, 1, 0, side.cuda_stream) == 0
        assert lg.gauge_upload(ub, 18) == 0  # must wait for the 200 launches above
        side.synchronize()
        assert torch.equal(d_res, want_a)
        assert lg.wilson_device(d_res, d_psi, 1, 0, side.cuda_stream) == 0
        side.synchronize()
        op.load_gauge(ub)
        want_b = op.new_spinor(zero=True)
        op.apply(want_b, d_psi, 1, 0)
        torch.cuda.synchronize()
        assert torch.equal(d_res, want_b) and not torch.equal(want_a, want_b)
    finally:
        lg.free_sse_su3dslash()


@pytest.mark.parametrize("prec", [32, 64])
def test_gauge_reuse_modes(oracle, libs, prec):
    """sse_su3dslash_set_gauge_reuse: mode 1 reuses the uploaded links while the
    pointer and checksum are unchanged and notices an in-place change; mode 2
    reuses while the pointer is unchanged (an in-place change is seen only after
    gauge_release, as documented); a different array is always uploaded; mode 0
    is the reference's upload on every call.  16^3 x 32 FP64 (75 MB of links)
    takes the multi-threaded checksum, 8^4 the single-thread one."""
    lg = libs[prec]
    dt = lg.dtype
    tol = 1e-5 if prec == 32 else 1e-13
    for latt in ((8, 8, 8, 8), (16, 16, 16, 32)):
        print(f"[legacy reuse] prec={prec} latt={latt}", flush=True)
        u = gen_gauge_host(latt, 4, seed=1, dtype=dt)
        u2 = gen_gauge_host(latt, 4, seed=2, dtype=dt)
        psi = gen_spinor_host(latt, 4, seed=1, dtype=dt)
        d1 = oracle.dslash(u, psi, np.zeros_like(psi), latt, 1, 0)
        d2 = oracle.dslash(u2, psi, np.zeros_like(psi), latt, 1, 0)
        lg.init_sse_su3dslash(latt)

        def call(arr, isign=1, cb=0):
            res = np.zeros_like(psi)
            lg.sse_su3dslash_wilson(arr, psi, res, isign, cb)
            assert lg.last_error() == 0
            return res

        try:
            assert lg.set_gauge_reuse(3) != 0 and lg.set_gauge_reuse(-1) != 0
            assert lg.set_gauge_reuse(1) == 0
            work = u.copy()
            assert R.rel_err(call(work), d1) < tol
            assert R.rel_err(call(work), d1) < tol  # reused
            r = call(work, -1, 1)                   # both checkerboards are resident
            assert R.rel_err(r, oracle.dslash(u, psi, np.zeros_like(psi), latt, -1, 1)) < tol
            work[...] = u2                          # in place: the checksum changes
            assert R.rel_err(call(work), d2) < tol
            work.reshape(-1)[5] += 0.25             # a single real changed in place
            d3 = oracle.dslash(work, psi, np.zeros_like(psi), latt, 1, 0)
            assert R.rel_err(call(work), d3) < tol
            assert R.rel_err(call(u), d1) < tol     # another pointer
            assert lg.set_gauge_reuse(2) == 0
            work[...] = u
            assert R.rel_err(call(work), d1) < tol
            work[...] = u2                          # pointer unchanged: still the old links
            assert R.rel_err(call(work), d1) < tol
            assert lg.gauge_release() == 0          # the caller says it changed
            assert R.rel_err(call(work), d2) < tol
            assert lg.set_gauge_reuse(0) == 0
            work[...] = u
            assert R.rel_err(call(work), d1) < tol
        finally:
            lg.set_gauge_reuse(0)
            lg.free_sse_su3dslash()


@pytest.mark.parametrize("prec,nd", [(32, 4), (64, 4), (32, 3)])
def test_legacy_slab_pipeline_large(oracle, libs, prec, nd):
    """At 32^3 x 64 the host-pointer call moves the source checkerboard up and
    the result back in slabs of whole t (4D) / z (3D) slices, computing each
    target slab once its neighbouring source slabs are in: every isign and
    source checkerboard, the gauge uploaded per call and reused, and psi
    aliasing res, against the oracle."""
    lg = libs[prec]
    dt = lg.dtype
    latt = (32, 32, 32, 64)
    print(f"[legacy slabs] prec={prec} nd={nd}", flush=True)
    u = gen_gauge_host(latt, nd, seed=6, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=6, dtype=dt)
    init = lg.init_sse_su3dslash if nd == 4 else lg.init_sse_su3dslash_3d
    free = lg.free_sse_su3dslash if nd == 4 else lg.free_sse_su3dslash_3d
    apply = lg.sse_su3dslash_wilson if nd == 4 else lg.sse_su3dslash_wilson_3d
    reuse = lg.set_gauge_reuse if nd == 4 else lg.set_gauge_reuse_3d
    tol = 1e-5 if prec == 32 else 1e-13
    vh = volume(latt) // 2
    init(latt)
    try:
        for mode in (0, 2):
            assert reuse(mode) == 0
            for isign, cb in ((1, 0), (-1, 1), (-1, 0), (1, 1)):
                res = np.full_like(psi, 3.0)
                apply(u, psi, res, isign, cb)
                want = oracle.dslash(u, psi, np.full_like(psi, 3.0), latt, isign, cb, nd=nd)
                assert R.rel_err(res, want) < tol, (mode, isign, cb)
                assert np.all(res.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == 3.0)
        inplace = psi.copy()
        apply(u, inplace, inplace, 1, 0)
        assert R.rel_err(inplace, oracle.dslash(u, psi, psi.copy(), latt, 1, 0, nd=nd)) < tol
    finally:
        reuse(0)
        free()
