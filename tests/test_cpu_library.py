"""The host-CPU build of the drop-in API (libsse_dslash_cpu_f32.so /
libsse_dslash_cpu_f64.so, sse_wilson_dslash_amd/csrc/cpu_dslash.cpp) on the
CPU: parity with the oracle and the committed fixtures, the reference's
semantics (lib/sse_su3dslash_32bit_scalar.c:67-180: target checkerboard
overwritten, source checkerboard untouched, psi/res aliasing, isign no-op,
refcount, exit(1) on an odd dimension), caller site orders through the
neighbour table, and results that do not depend on the thread count or the
work-item shape."""
import glob
import os
import subprocess
import sys

import numpy as np
import pytest

from sse_wilson_dslash_amd import _lib
from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.layout import from_order, to_order, volume
from sse_wilson_dslash_amd.legacy import SSEDslash, table_callbacks
from tests import refimpl as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))
TOL = {32: 1e-5, 64: 1e-13}


@pytest.fixture(scope="module")
def cpu():
    return {32: SSEDslash(32, backend="cpu"), 64: SSEDslash(64, backend="cpu")}


def _fns(lg, nd):
    if nd == 4:
        return lg.init_sse_su3dslash, lg.sse_su3dslash_wilson, lg.free_sse_su3dslash
    return lg.init_sse_su3dslash_3d, lg.sse_su3dslash_wilson_3d, lg.free_sse_su3dslash_3d


def test_cpu_library_exports():
    """Both CPU builds export every symbol the drop-in headers declare, like
    the GPU builds, and need no HIP runtime."""
    names = ["init_sse_su3dslash", "free_sse_su3dslash", "sse_su3dslash_wilson", "sse_su3dslash_prepost_receives",
             "init_sse_su3dslash_3d", "free_sse_su3dslash_3d", "sse_su3dslash_wilson_3d", "pack_gauge_field",
             "unpack_gauge_field", "unpack_gauge_field_inverse", "sse_su3dslash_last_error",
             "sse_su3dslash_3d_last_error", "sse_su3dslash_gauge_upload", "sse_su3dslash_gauge_release",
             "sse_su3dslash_wilson_resident", "sse_su3dslash_wilson_device", "sse_su3dslash_set_sloppy",
             "sse_su3dslash_set_gauge_reuse", "sse_su3dslash_3d_gauge_upload", "sse_su3dslash_3d_gauge_release",
             "sse_su3dslash_wilson_3d_resident", "sse_su3dslash_wilson_3d_device", "sse_su3dslash_3d_set_sloppy",
             "sse_su3dslash_3d_set_gauge_reuse", "sse_su3dslash_cpu_threads"]
    import ctypes as C
    for prec in (32, 64):
        path = _lib.lib_path(_lib.LEGACY_CPU[prec])
        lib = C.CDLL(path, mode=C.RTLD_LOCAL)
        for n in names:
            assert hasattr(lib, n), (prec, n)
        deps = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
        assert "amdhip" not in deps and "dslash_mi355x" not in deps, deps
    # the thread query follows OMP_NUM_THREADS in a fresh process
    code = (f"import ctypes; L = ctypes.CDLL({_lib.lib_path(_lib.LEGACY_CPU[32])!r}); "
            "print(L.sse_su3dslash_cpu_threads())")
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, OMP_NUM_THREADS="3"),
                       capture_output=True, text=True, timeout=60)
    assert r.stdout.strip() == "3", r.stdout + r.stderr


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nd,latt", [(4, (4, 4, 4, 4)), (4, (2, 6, 4, 2)), (4, (8, 4, 6, 10)), (4, (6, 2, 2, 4)),
                                     (3, (4, 4, 4, 3)), (3, (8, 4, 6, 5)), (3, (2, 6, 4, 2))])
def test_cpu_matches_oracle(oracle, cpu, prec, nd, latt):
    lg = cpu[prec]
    init, apply, free = _fns(lg, nd)
    u = gen_gauge_host(latt, nd, seed=3, dtype=lg.dtype)
    psi = gen_spinor_host(latt, nd, seed=3, dtype=lg.dtype)
    vh = volume(latt) // 2
    init(latt)
    try:
        for isign in (1, -1):
            for cb in (0, 1):
                res = np.full_like(psi, -2.0)
                apply(u, psi, res, isign, cb)
                want = oracle.dslash(u, psi, np.full_like(psi, -2.0), latt, isign, cb, nd=nd)
                assert R.rel_err(res, want) < TOL[prec], (isign, cb)
                assert np.all(res.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == -2.0)
        assert (lg.last_error() if nd == 4 else lg.last_error_3d()) == 0
    finally:
        free()


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.rsplit("/", 1)[-1][:-4])
def test_cpu_matches_golden_fixtures(cpu, path):
    """The committed oracle fixtures (tests/golden/make_fixtures.py), without the oracle."""
    f = np.load(path)
    nd, latt = int(f["nd"]), tuple(int(v) for v in f["latt"])
    for prec, tag in ((64, "f64"), (32, "f32")):
        lg = cpu[prec]
        init, apply, free = _fns(lg, nd)
        u, psi = f["u"].astype(lg.dtype), f["psi"].astype(lg.dtype)
        init(latt)
        try:
            for isign in (1, -1):
                for cb in (0, 1):
                    res = np.zeros_like(psi)
                    apply(u, psi, res, isign, cb)
                    assert R.rel_err(res, f[f"res_{tag}_{'p' if isign > 0 else 'm'}{cb}"]) < TOL[prec]
        finally:
            free()


@pytest.mark.parametrize("nd,latt", [(4, (4, 6, 2, 4)), (3, (4, 2, 6, 3))])
def test_cpu_caller_order(oracle, cpu, nd, latt):
    """A non-canonical caller order (random permutation through Python
    callbacks) goes through the neighbour table built at init."""
    lg = cpu[64]
    init, apply, free = _fns(lg, nd)
    V = volume(latt)
    rng = np.random.default_rng(7)
    tab = rng.permutation(V).astype(np.int32)
    cbs = table_callbacks(latt, tab, nd)
    u = gen_gauge_host(latt, nd, seed=4)
    psi = gen_spinor_host(latt, nd, seed=4)
    init(latt, *cbs)
    try:
        u_c, psi_c = to_order(u, tab), to_order(psi, tab)
        for isign in (1, -1):
            for cb in (0, 1):
                res = np.full_like(psi_c, 5.0)
                apply(u_c, psi_c, res, isign, cb)
                want = oracle.dslash(u, psi, np.full_like(psi, 5.0), latt, isign, cb, nd=nd)
                assert R.rel_err(from_order(res, tab), want) < 1e-13
    finally:
        free()


def test_cpu_semantics(oracle, cpu):
    """psi == res aliasing, isign other than +-1 is a no-op, refcount (a second
    init keeps the first size), the resident-pointer extension, the device
    and sloppy extensions report unsupported, use after the last free reports
    not-initialised instead of crashing."""
    lg = cpu[32]
    latt = (4, 4, 4, 4)
    u = gen_gauge_host(latt, 4, seed=1, dtype=np.float32)
    psi = gen_spinor_host(latt, 4, seed=1, dtype=np.float32)
    lg.init_sse_su3dslash(latt)
    lg.init_sse_su3dslash((6, 6, 6, 6))
    try:
        want = oracle.dslash(u, psi, psi.copy(), latt, 1, 0)
        inplace = psi.copy()
        lg.sse_su3dslash_wilson(u, inplace, inplace, 1, 0)
        assert R.rel_err(inplace, want) < 1e-5
        before = inplace.copy()
        lg.sse_su3dslash_wilson(u, psi, inplace, 0, 0)
        lg.sse_su3dslash_wilson(u, psi, inplace, 2, 1)
        assert np.array_equal(inplace, before)
        lg.sse_su3dslash_prepost_receives()
        lg.free_sse_su3dslash()  # refcount 1: still usable
        assert lg.gauge_upload(u, 18) == 0
        r2 = np.zeros_like(psi)
        assert lg.wilson_resident(psi, r2, -1, 1) == 0
        assert R.rel_err(r2, oracle.dslash(u, psi, np.zeros_like(psi), latt, -1, 1)) < 1e-5
        assert lg.set_gauge_reuse(1) == 0 and lg.set_gauge_reuse(7) == _lib.ERR_INVALID
        assert lg.set_sloppy(True) == _lib.ERR_UNSUPPORTED and lg.set_sloppy(False) == 0
        assert lg.lib.sse_su3dslash_wilson_device(None, None, 1, 0, None) == _lib.ERR_UNSUPPORTED
    finally:
        lg.free_sse_su3dslash()
    assert lg.wilson_resident(psi, np.zeros_like(psi), 1, 0) != 0
    res = np.zeros_like(psi)
    lg.sse_su3dslash_wilson(u, psi, res, 1, 0)
    assert lg.last_error() == _lib.ERR_NOT_INIT and not res.any()


@pytest.mark.parametrize("nd,latt,msg", [(4, "[4,4,3,4]", b"4 dimensions"), (3, "[4,3,4,5]", b"3 dimensions")])
def test_cpu_odd_dimension_exits_1(nd, latt, msg):
    fn = "init_sse_su3dslash" if nd == 4 else "init_sse_su3dslash_3d"
    code = (f"import sys; sys.path.insert(0, {ROOT!r})\n"
            "from sse_wilson_dslash_amd.legacy import SSEDslash\n"
            f"SSEDslash(32, backend='cpu').{fn}({latt})\n"
            "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=120)
    assert r.returncode == 1
    assert msg in r.stderr and b"not reached" not in r.stdout


_CHILD = """
import sys; sys.path.insert(0, {root!r})
import numpy as np
from sse_wilson_dslash_amd.legacy import SSEDslash
from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host
lg = SSEDslash({prec}, backend='cpu')
latt = {latt}
u = gen_gauge_host(latt, {nd}, seed=5, dtype=lg.dtype)
psi = gen_spinor_host(latt, {nd}, seed=5, dtype=lg.dtype)
init, apply = (lg.init_sse_su3dslash, lg.sse_su3dslash_wilson) if {nd} == 4 else \\
    (lg.init_sse_su3dslash_3d, lg.sse_su3dslash_wilson_3d)
init(latt)
out = []
for isign, cb in ((1, 0), (-1, 1)):
    res = np.zeros_like(psi)
    apply(u, psi, res, isign, cb)
    out.append(res)
np.save(sys.argv[1], np.stack(out))
"""


@pytest.mark.parametrize("prec,nd,latt", [(32, 4, (8, 6, 10, 12)), (64, 3, (6, 4, 10, 9))])
def test_cpu_result_independent_of_threads_and_blocking(tmp_path, prec, nd, latt):
    """Each target site is computed by one thread in a fixed order, so the
    result is bit-identical for any thread count and any work-item shape
    (SSE_DSLASH_CPU_BLOCK), including shapes that do not divide the lattice."""
    code = _CHILD.format(root=ROOT, prec=prec, nd=nd, latt=latt)
    runs = [("1", None), ("3", None), ("4", "3,5,4"), ("2", "1,1,1"), ("5", "4,7,3")]
    outs = []
    for k, (threads, block) in enumerate(runs):
        env = dict(os.environ, OMP_NUM_THREADS=threads)
        env.pop("SSE_DSLASH_CPU_BLOCK", None)
        if block:
            env["SSE_DSLASH_CPU_BLOCK"] = block
        path = str(tmp_path / f"r{k}.npy")
        subprocess.run([sys.executable, "-c", code, path], check=True, env=env, timeout=240)
        outs.append(np.load(path))
    for o in outs[1:]:
        assert o.tobytes() == outs[0].tobytes()


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nd,latt", [(4, (8, 4, 6, 8)), (3, (4, 6, 4, 6))])
def test_c_caller_on_cpu_build(oracle, tmp_path, prec, nd, latt):
    """tests/dropin_check.c (a C caller shaped like the reference's full-lattice
    tests, tests/testDslashFull.cc:24-117) built against the CPU library of each
    precision: every isign / source checkerboard and the in-place call against
    the oracle, source checkerboard of res untouched."""
    exe = tmp_path / "dropin_cpu"
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "include", f"f{prec}"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "dropin_check.c"), "-L", _lib.LIB_DIR,
                    f"-l{_lib.LEGACY_CPU[prec][3:-3]}", f"-Wl,-rpath,{_lib.LIB_DIR}"],
                   check=True, capture_output=True, timeout=120)
    env = dict(os.environ, OMP_NUM_THREADS="2")
    out = subprocess.run([str(exe), str(nd), *[str(x) for x in latt], str(tmp_path)], capture_output=True,
                         text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    dt = np.float32 if prec == 32 else np.float64
    V = volume(latt)
    vh = V // 2
    u = np.fromfile(tmp_path / "u.bin", dtype=dt)
    psi = np.fromfile(tmp_path / "psi.bin", dtype=dt)
    for isign in (1, -1):
        for cb in (0, 1):
            got = np.fromfile(tmp_path / f"res_{isign}_{cb}.bin", dtype=dt)
            want = oracle.dslash(u, psi, np.full(V * 24, 7.0, dtype=dt), latt, isign, cb, nd=nd)
            t = slice((1 - cb) * vh * 24, (2 - cb) * vh * 24)
            assert np.all(got[cb * vh * 24:(cb + 1) * vh * 24] == 7.0)
            assert R.rel_err(got[t], want[t]) < TOL[prec], (isign, cb)
    got = np.fromfile(tmp_path / "inplace_1_0.bin", dtype=dt)
    want = oracle.dslash(u, psi, psi.copy(), latt, 1, 0, nd=nd)
    assert np.array_equal(got[:vh * 24], psi[:vh * 24])
    assert R.rel_err(got[vh * 24:], want[vh * 24:]) < TOL[prec]
