"""The drop-in API driven from compiled C, as Chroma or the reference's own full-
lattice tests (tests/testDslashFull.cc, tests/testDslash3D.cc) drive it: a C
program (tests/dropin_check.c) built with gcc against libsse_dslash_f32.so or
libsse_dslash_f64.so registers canonical site-order callbacks, applies
sse_su3dslash_wilson(_3d) for both isign and both source checkerboards to host
arrays, once in place, and frees the library.  Its arrays are checked against the
oracle: the target checkerboard to 1e-5 (FP32) / 1e-13 (FP64) relative and the
reference's diff_norm acceptance (testDslashFull.cc:45-56, 103-109), the source
checkerboard of res left as it was, bit for bit."""
import os
import subprocess

import numpy as np
import pytest
import torch

from tests import refimpl as R

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "sse_wilson_dslash_amd", "_lib")


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


@pytest.fixture(scope="module")
def programs(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    exes = {}
    for prec in (32, 64):
        exe = d / f"dropin_check_f{prec}"
        cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(ROOT, "include", f"f{prec}"), "-o", str(exe),
               os.path.join(ROOT, "tests", "dropin_check.c"), "-L", LIBDIR, f"-lsse_dslash_f{prec}",
               f"-Wl,-rpath,{LIBDIR}"]
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
        exes[prec] = exe
    return exes


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("nd,latt", [(4, (8, 4, 6, 8)), (4, (16, 16, 16, 16)), (3, (4, 6, 4, 6)),
                                     (3, (8, 8, 8, 16))])
def test_c_caller(oracle, programs, tmp_path, prec, nd, latt):
    print(f"[drop-in C] f{prec} nd={nd} latt={latt}", flush=True)
    out = subprocess.run([str(programs[prec]), str(nd), *[str(x) for x in latt], str(tmp_path)],
                         capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split()[-2:] == ["SSEREAL", str(prec // 8)]
    dt = np.float32 if prec == 32 else np.float64
    V = int(np.prod(latt))
    vh = V // 2
    u = np.fromfile(tmp_path / "u.bin", dtype=dt)
    psi = np.fromfile(tmp_path / "psi.bin", dtype=dt)
    assert u.size == V * 72 and psi.size == V * 24
    tol = 1e-5 if prec == 32 else 1e-13
    small = 1e-7 if prec == 32 else 1e-16
    for isign in (1, -1):
        for cb in (0, 1):
            got = np.fromfile(tmp_path / f"res_{isign}_{cb}.bin", dtype=dt)
            want = oracle.dslash(u, psi, np.full(V * 24, 7.0, dtype=dt), latt, isign, cb, nd=nd)
            t = slice((1 - cb) * vh * 24, (2 - cb) * vh * 24)
            s = slice(cb * vh * 24, (cb + 1) * vh * 24)
            assert np.all(got[s] == 7.0), "source checkerboard of res written"
            e = R.rel_err(got[t], want[t])
            diff_norm = float(np.linalg.norm((got[t] - want[t]).astype(np.float64))) / (4 * 3 * 2 * V / 2)
            print(f"  isign={isign:+d} cb={cb}: rel err {e:.2e}, diff_norm {diff_norm:.2e}", flush=True)
            assert e < tol and diff_norm < small, (isign, cb, e, diff_norm)
    got = np.fromfile(tmp_path / "inplace_1_0.bin", dtype=dt)
    want = oracle.dslash(u, psi, psi.copy(), latt, 1, 0, nd=nd)
    assert np.array_equal(got[:vh * 24], psi[:vh * 24]), "in place: the source checkerboard changed"
    assert R.rel_err(got[vh * 24:], want[vh * 24:]) < tol
