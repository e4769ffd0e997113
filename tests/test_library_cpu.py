"""CPU-side checks of the HIP libraries: they load, export every declared
symbol, and their host-only entry points (generators, layout callbacks, the C
gauge packer, argument validation) behave as specified.  No kernel runs here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from sse_wilson_dslash_amd import _lib
from sse_wilson_dslash_amd.layout import coords_of_full, full_of_coords, lex_permutation, volume
from tests import refimpl as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def test_core_exports_every_declared_symbol():
    lib = _lib.core()
    names = declared("dslash_mi355x.h")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert all(n in _lib.PROTOTYPES for n in names), [n for n in names if n not in _lib.PROTOTYPES]


@pytest.mark.parametrize("prec", [32, 64])
def test_legacy_exports_every_declared_symbol(prec):
    _lib.core()
    lib = C.CDLL(_lib.lib_path(_lib.LEGACY[prec]), mode=C.RTLD_LOCAL)
    names = declared("sse_dslash.h") + declared("sse_dslash_3d.h")
    for ref in ("init_sse_su3dslash", "free_sse_su3dslash", "sse_su3dslash_wilson",
                "sse_su3dslash_prepost_receives", "init_sse_su3dslash_3d", "free_sse_su3dslash_3d",
                "sse_su3dslash_wilson_3d", "pack_gauge_field"):
        assert ref in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32 with 10 rounds."""
    lib = _lib.core()

    def ph(ctr, key):
        o = (C.c_uint32 * 4)()
        lib.dslash_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), o)
        return list(o)

    assert ph([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert ph([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert ph([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


@pytest.mark.parametrize("nd", [4, 3])
def test_generated_gauge_is_su3_and_layout_independent(nd):
    from sse_wilson_dslash_amd.device import gen_gauge_host
    latt = (4, 4, 2, 6)
    u = gen_gauge_host(latt, nd, seed=1)
    U = R.links(u)
    V = volume(latt)
    eye = np.eye(3)
    UU = U @ U.conj().transpose(0, 1, 3, 2)
    assert np.max(np.abs(UU - eye)) < 1e-14
    assert np.max(np.abs(np.linalg.det(U) - 1)) < 1e-14
    # the generator follows tests/reunit.cc:184-241: column 2 = conj(col0 x col1)
    c0, c1, c2 = U[..., :, 0], U[..., :, 1], U[..., :, 2]
    assert np.max(np.abs(c2 - np.cross(c0, c1).conj())) < 1e-14
    # keyed by the lexicographic site: the 4D and 3D orders hold the same field
    other = gen_gauge_host(latt, 7 - nd, seed=1)
    lex_a = lex_permutation(latt, nd)
    lex_b = lex_permutation(latt, 7 - nd)
    a = np.empty_like(u); a[lex_a] = u
    b = np.empty_like(u); b[lex_b] = other
    assert np.array_equal(a, b)
    # a different seed gives a different field
    assert not np.array_equal(gen_gauge_host(latt, nd, seed=2), u)
    # f32 generation is the f64 field rounded
    assert np.array_equal(gen_gauge_host(latt, nd, seed=1, dtype=np.float32), u.astype(np.float32))
    assert u.shape == (V, 4, 3, 3, 2)


def test_generated_spinor_is_standard_gaussian():
    from sse_wilson_dslash_amd.device import gen_spinor_host
    latt = (8, 8, 8, 8)
    p = gen_spinor_host(latt, 4, seed=1, field=0)
    x = p.ravel()
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1) < 0.01
    q = gen_spinor_host(latt, 4, seed=1, field=1)
    assert abs(np.corrcoef(x, q.ravel())[0, 1]) < 0.01


@pytest.mark.parametrize("nd,latt", [(4, (6, 10, 2, 12)), (4, (48, 2, 2, 6)), (3, (4, 6, 2, 96)),
                                     (3, (24, 2, 4, 9)), (4, (2, 2, 2, 2))])
def test_layout_callbacks_match_canonical_order(nd, latt):
    """The exported cb2/cb3d callbacks (which use the kernels' fast division)
    agree with the defining formulas for every site."""
    lib = _lib.core()
    _lib.check(lib.dslash_layout_set(_lib.latt_arg(latt), nd))
    V = volume(latt)
    c = coords_of_full(latt, nd, np.arange(V))
    coord = (C.c_int * 4)()
    for i in range(V):
        lib.dslash_layout_site_coords(coord, 0, i)
        assert list(coord) == list(c[i]), i
        assert lib.dslash_layout_linear_index(coord) == i
    assert np.array_equal(full_of_coords(latt, nd, c), np.arange(V))


def test_sizes_and_validation():
    lib = _lib.core()
    L = _lib.latt_arg((32, 32, 32, 64))
    vh = 32 * 32 * 32 * 64 // 2
    assert lib.dslash_spinor_cb_bytes(L, 4, _lib.F32) == vh * 96
    assert lib.dslash_spinor_cb_bytes(L, 4, _lib.F64) == vh * 192
    assert lib.dslash_spinor_cb_bytes(L, 4, _lib.F16) == vh * 52
    assert lib.dslash_gauge_cb_bytes(L, 4, _lib.F32, 18) == vh * 8 * 18 * 4
    assert lib.dslash_gauge_cb_bytes(L, 4, _lib.F32, 12) == vh * 8 * 12 * 4
    assert lib.dslash_gauge_cb_bytes(L, 3, _lib.F64, 18) == vh * 6 * 18 * 8
    assert lib.dslash_gauge_cb_bytes(L, 4, _lib.F16, 12) == vh * 8 * 12 * 2
    # sloppy FP64: FP64 spinor storage, FP32 link bundles
    assert lib.dslash_spinor_cb_bytes(L, 4, _lib.F64S) == vh * 192
    assert lib.dslash_gauge_cb_bytes(L, 4, _lib.F64S, 18) == vh * 8 * 18 * 4
    assert lib.dslash_gauge_cb_bytes(L, 3, _lib.F64S, 12) == vh * 6 * 12 * 4
    assert lib.dslash_spinor_cb_bytes(L, 4, 5) == 0
    assert lib.dslash_check_lattice(_lib.latt_arg((4, 4, 4, 3)), 4) == _lib.ERR_INVALID
    assert b"even" in lib.dslash_last_error_message()
    assert lib.dslash_check_lattice(_lib.latt_arg((4, 4, 4, 3)), 3) == 0
    assert lib.dslash_check_lattice(_lib.latt_arg((4, 4, 4, 4)), 5) == _lib.ERR_INVALID
    # invalid isign / cb are rejected before any device access
    st = lib.dslash_apply(1, 2, 3, _lib.latt_arg((4, 4, 4, 4)), 4, _lib.F32, 18, 0, 0, None)
    assert st == _lib.ERR_INVALID and b"isign" in lib.dslash_last_error_message()
    st = lib.dslash_apply(1, 2, 3, _lib.latt_arg((4, 4, 4, 4)), 4, _lib.F32, 18, 1, 2, None)
    assert st == _lib.ERR_INVALID
    st = lib.dslash_apply(1, 2, 3, _lib.latt_arg((4, 4, 4, 4)), 4, 9, 18, 1, 0, None)
    assert st == _lib.ERR_INVALID
    st = lib.dslash_apply(1, 2, 3, _lib.latt_arg((4, 4, 4, 4)), 4, _lib.F32, 10, 1, 0, None)
    assert st == _lib.ERR_INVALID
    assert lib.dslash_set_launch(0, 0, 0, 100, 1) == _lib.ERR_INVALID
    b, r = C.c_int(), C.c_int()
    assert lib.dslash_get_launch(4, _lib.F32, 18, C.byref(b), C.byref(r)) == 0
    saved = (b.value, r.value)
    assert lib.dslash_set_launch(4, _lib.F32, 18, 128, 0) == 0
    lib.dslash_get_launch(4, _lib.F32, 18, C.byref(b), C.byref(r))
    assert (b.value, r.value) == (128, 0)
    lib.dslash_get_launch(3, _lib.F32, 18, C.byref(b), C.byref(r))
    assert b.value in (64, 128, 192, 256)
    assert lib.dslash_set_launch(4, _lib.F32, 18, *saved) == 0


def test_march_setting():
    lib = _lib.core()
    m = C.c_int()
    assert lib.dslash_get_launch_march(3, _lib.F32, 18, C.byref(m)) == 0
    saved = m.value
    assert saved in (1, 4, 8, 16, 32)
    for bad in (0, 2, 3, 24, 64, -1):
        assert lib.dslash_set_launch_march(3, _lib.F32, 18, bad) == _lib.ERR_INVALID
    assert b"march" in lib.dslash_last_error_message()
    assert lib.dslash_set_launch_march(5, _lib.F32, 18, 4) == _lib.ERR_INVALID
    assert lib.dslash_set_launch_march(3, _lib.F32, 18, 4) == 0
    b, r = C.c_int(), C.c_int()
    lib.dslash_get_launch(3, _lib.F32, 18, C.byref(b), C.byref(r))
    assert lib.dslash_set_launch(3, _lib.F32, 18, 64, 0) == 0  # leaves the march alone
    lib.dslash_get_launch_march(3, _lib.F32, 18, C.byref(m))
    assert m.value == 4
    assert lib.dslash_set_launch(3, _lib.F32, 18, b.value, r.value) == 0
    assert lib.dslash_set_launch_march(3, _lib.F32, 18, saved) == 0


def test_march_falls_back_below_the_thread_floor():
    lib = _lib.core()
    m = C.c_int()
    lib.dslash_get_launch_march(4, _lib.F64, 18, C.byref(m))
    saved4 = m.value
    lib.dslash_get_launch_march(3, _lib.F64, 18, C.byref(m))
    saved3 = m.value

    def eff(latt, nd):
        assert lib.dslash_launch_march_for(_lib.latt_arg(latt), nd, _lib.F64, 18, C.byref(m)) == 0
        return m.value
    try:
        assert lib.dslash_set_launch_march(4, _lib.F64, 18, 8) == 0
        assert lib.dslash_set_launch_march(3, _lib.F64, 18, 8) == 0
        # 4D: 16384 sites per t-slice; threads = 16384 * ceil(Lt / march) >= 131072
        assert eff((32, 32, 32, 64), 4) == 8   # 8 chunks
        assert eff((32, 32, 32, 66), 4) == 8   # 9 chunks, the last of 2 slices
        assert eff((32, 32, 32, 32), 4) == 4   # 8: 65536 threads; 4: 131072
        assert eff((32, 32, 32, 8), 4) == 1
        assert eff((4, 4, 4, 4), 4) == 1
        assert eff((64, 64, 64, 2), 4) == 8    # 131072 sites per slice
        # 3D: z is the slowest coordinate; 32768 sites per z-slice at 32^2 x 64
        assert eff((32, 32, 32, 64), 3) == 8
        assert eff((32, 32, 16, 64), 3) == 4
        assert eff((16, 16, 16, 64), 3) == 1
        assert lib.dslash_set_launch_march(4, _lib.F64, 18, 1) == 0
        assert eff((32, 32, 32, 64), 4) == 1
        # box tiles (16) exist for the 4D FP32 / FP64 operators; elsewhere they fall back to 8
        b, r = C.c_int(), C.c_int()
        lib.dslash_get_launch(4, _lib.F64, 18, C.byref(b), C.byref(r))
        assert lib.dslash_set_launch_march(4, _lib.F64, 18, 16) == 0
        assert lib.dslash_set_launch(4, _lib.F64, 18, 256, 1) == 0
        assert eff((32, 32, 32, 64), 4) == 16
        assert lib.dslash_set_launch(4, _lib.F64, 18, 64, 1) == 0
        assert eff((32, 32, 32, 64), 4) == 8
        assert lib.dslash_set_launch(4, _lib.F64, 18, b.value, r.value) == 0
        assert lib.dslash_set_launch_march(3, _lib.F64, 18, 16) == 0
        assert eff((32, 32, 32, 64), 3) == 8
        assert lib.dslash_launch_march_for(_lib.latt_arg((3, 4, 4, 4)), 4, _lib.F64, 18, C.byref(m)) == \
            _lib.ERR_INVALID
    finally:
        lib.dslash_set_launch_march(4, _lib.F64, 18, saved4)
        lib.dslash_set_launch_march(3, _lib.F64, 18, saved3)


def test_box_tiles_single_rhs_fallbacks():
    """march 16 (box tiles) runs where the 4D FP32 launch has >= 128 threads per
    block, a whole x row per block and >= 131072 threads; otherwise march 8 (then
    the thread floor applies as for any march)."""
    lib = _lib.core()
    b, r, m = C.c_int(), C.c_int(), C.c_int()
    lib.dslash_get_launch(4, _lib.F32, 18, C.byref(b), C.byref(r))
    lib.dslash_get_launch_march(4, _lib.F32, 18, C.byref(m))
    saved = (b.value, r.value, m.value)

    def eff(latt):
        assert lib.dslash_launch_march_for(_lib.latt_arg(latt), 4, _lib.F32, 18, C.byref(m)) == 0
        return m.value
    try:
        assert lib.dslash_set_launch_march(4, _lib.F32, 18, 16) == 0
        assert lib.dslash_set_launch(4, _lib.F32, 18, 256, 1) == 0
        assert eff((32, 32, 32, 64)) == 16   # 64 boxes of 16 x 4 x 4 x 8 chunks = 512 blocks of 256
        assert eff((48, 48, 48, 96)) == 16
        assert eff((8, 8, 8, 8)) == 1        # one box: below the thread floor, and so is march 8
        assert eff((600, 8, 8, 64)) == 8     # Lx/2 = 300 > 256 threads
        assert lib.dslash_set_launch(4, _lib.F32, 18, 64, 1) == 0
        assert eff((32, 32, 32, 64)) == 8    # 64 threads per block: no box tiles
        # 32: box tiles with the LDS halo, where the halo takes at most one entry per thread
        assert lib.dslash_set_launch_march(4, _lib.F32, 18, 32) == 0
        assert lib.dslash_set_launch(4, _lib.F32, 18, 256, 1) == 0
        assert eff((32, 32, 32, 64)) == 32   # 16 x 4 x 4 boxes: 2 x 64 y + 2 x 64 z halo sites = 256
        assert eff((32, 32, 32, 66)) == 32
        assert eff((48, 48, 48, 96)) == 16   # 24 x 3 x 3 boxes: 288 halo sites > 256 threads
        assert eff((8, 8, 8, 8)) == 1        # as for 16
        assert eff((600, 8, 8, 64)) == 8
        assert lib.dslash_set_launch(4, _lib.F32, 18, 64, 1) == 0
        assert eff((32, 32, 32, 64)) == 8
        lib.dslash_get_launch(4, _lib.F64, 18, C.byref(b), C.byref(r))
        lib.dslash_get_launch_march(4, _lib.F64, 18, C.byref(m))
        f64 = (b.value, r.value, m.value)
        assert lib.dslash_set_launch_march(4, _lib.F64, 18, 32) == 0  # FP64: the plain box tiles
        assert lib.dslash_set_launch(4, _lib.F64, 18, 256, 1) == 0
        assert lib.dslash_launch_march_for(_lib.latt_arg((32, 32, 32, 64)), 4, _lib.F64, 18, C.byref(m)) == 0
        assert m.value == 16
        assert lib.dslash_set_launch(4, _lib.F64, 18, f64[0], f64[1]) == 0
        assert lib.dslash_set_launch_march(4, _lib.F64, 18, f64[2]) == 0
    finally:
        lib.dslash_set_launch(4, _lib.F32, 18, saved[0], saved[1])
        lib.dslash_set_launch_march(4, _lib.F32, 18, saved[2])


def test_tune_file_round_trip(tmp_path):
    lib = _lib.core()
    b, r, m = C.c_int(), C.c_int(), C.c_int()
    lib.dslash_get_launch(3, _lib.F64, 12, C.byref(b), C.byref(r))
    lib.dslash_get_launch_march(3, _lib.F64, 12, C.byref(m))
    saved = (b.value, r.value)
    saved_m = m.value
    assert lib.dslash_set_launch(3, _lib.F64, 12, 192, 0) == 0
    assert lib.dslash_set_launch_march(3, _lib.F64, 12, 8) == 0
    path = str(tmp_path / "tune.txt").encode()
    assert lib.dslash_tune_save(path) == 0
    assert "apply 3 1 12 192 0 8\n" in open(path).read()
    assert lib.dslash_set_launch(3, _lib.F64, 12, 64, 1) == 0
    assert lib.dslash_set_launch_march(3, _lib.F64, 12, 1) == 0
    assert lib.dslash_tune_load(path) == 0
    lib.dslash_get_launch(3, _lib.F64, 12, C.byref(b), C.byref(r))
    lib.dslash_get_launch_march(3, _lib.F64, 12, C.byref(m))
    assert (b.value, r.value, m.value) == (192, 0, 8)
    old = tmp_path / "old.txt"  # files from before the march field: march 1
    old.write_text("apply 3 1 12 128 1\n")
    assert lib.dslash_tune_load(str(old).encode()) == 0
    lib.dslash_get_launch(3, _lib.F64, 12, C.byref(b), C.byref(r))
    lib.dslash_get_launch_march(3, _lib.F64, 12, C.byref(m))
    assert (b.value, r.value, m.value) == (128, 1, 1)
    assert lib.dslash_set_launch(3, _lib.F64, 12, *saved) == 0
    assert lib.dslash_set_launch_march(3, _lib.F64, 12, saved_m) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("apply 4 0 18 100 1\n")
    assert lib.dslash_tune_load(str(bad).encode()) == _lib.ERR_INVALID
    bad.write_text("apply 4 0 18 128 1 3\n")
    assert lib.dslash_tune_load(str(bad).encode()) == _lib.ERR_INVALID
    assert lib.dslash_tune_load(b"/nonexistent/tune.txt") == _lib.ERR_INVALID


def test_mrhs_launch_table_and_tune_file(tmp_path):
    lib = _lib.core()
    v = [C.c_int() for _ in range(4)]
    assert lib.dslash_get_mrhs_launch(_lib.F32, 18, 0, *[C.byref(x) for x in v]) == 0
    saved = [x.value for x in v]
    lib.dslash_get_mrhs_launch(_lib.F32, 12, 1, *[C.byref(x) for x in v])
    saved12 = [x.value for x in v]
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 4, 128, 0, 4) == _lib.ERR_INVALID  # no R=4 marching
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 128, 512, 1) == _lib.ERR_INVALID  # grid only if persistent
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 192, 0, 3) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 64, 0, 8) == _lib.ERR_INVALID  # LDS tiles: 128 threads
    assert lib.dslash_set_mrhs_launch(_lib.F64, 18, 0, 2, 128, 0, 8) == _lib.ERR_INVALID  # LDS tiles: F32 only
    assert lib.dslash_set_mrhs_launch(_lib.F64, 18, 0, 2, 192, 0, 16) == _lib.ERR_INVALID  # box tiles: F32 only
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 4, 192, 0, 16) == _lib.ERR_INVALID  # rhs_block <= 2
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 64, 0, 16) == _lib.ERR_INVALID  # >= 128 threads
    # wide box tiles: rhs_block x threads <= 768, F32 box tiles only
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 768, 0, 16) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 384, 0, 16) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 512, 0, 16) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 832, 0, 16) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 512, 0, 8) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 512, 0, 1) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F64, 18, 0, 1, 512, 0, 16) == _lib.ERR_INVALID
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 300, 0, 16) == _lib.ERR_INVALID  # multiple of 64
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 2, 192, 0, 16) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 12, 1, 2, 128, 1024, 8) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, 1, 192, 0, 4) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 12, 0, 1, 768, 0, 16) == 0
    path = str(tmp_path / "tune.txt")
    assert lib.dslash_tune_save(path.encode()) == 0
    assert "mrhs 0 18 0 1 192 0 4\n" in open(path).read()
    assert "mrhs 0 12 0 1 768 0 16\n" in open(path).read()
    saved12np = [2, 192, 0, 16]
    assert lib.dslash_set_mrhs_launch(_lib.F32, 12, 0, *saved12np) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, *saved) == 0
    assert "mrhs 0 12 1 2 128 1024 8\n" in open(path).read()
    assert lib.dslash_tune_load(path.encode()) == 0
    lib.dslash_get_mrhs_launch(_lib.F32, 18, 0, *[C.byref(x) for x in v])
    assert [x.value for x in v] == [1, 192, 0, 4]
    lib.dslash_get_mrhs_launch(_lib.F32, 12, 0, *[C.byref(x) for x in v])
    assert [x.value for x in v] == [1, 768, 0, 16]  # a wide shape reloads
    assert lib.dslash_set_mrhs_launch(_lib.F32, 12, 0, *saved12np) == 0
    old = tmp_path / "old.txt"  # a file from before t-marching tiles: 6 fields, t_block 1
    old.write_text("mrhs 0 18 0 2 128 0\n")
    assert lib.dslash_tune_load(str(old).encode()) == 0
    lib.dslash_get_mrhs_launch(_lib.F32, 18, 0, *[C.byref(x) for x in v])
    assert [x.value for x in v] == [2, 128, 0, 1]
    assert lib.dslash_set_mrhs_launch(_lib.F32, 18, 0, *saved) == 0
    assert lib.dslash_set_mrhs_launch(_lib.F32, 12, 1, *saved12) == 0


def test_face_layout():
    """Face buffer of the domain-decomposition stages: per partitioned mu a
    forward then a backward face of Vh / L_mu half-spinors (12 reals), each
    direction's pair 256-byte aligned; invalid directions and formats rejected."""
    lib = _lib.core()
    latt = (4, 6, 8, 10)
    vh = 4 * 6 * 8 * 10 // 2
    off, fb, tot = (C.c_size_t * 8)(), (C.c_size_t * 4)(), C.c_size_t()
    assert lib.dslash_face_layout(_lib.latt_arg(latt), 4, _lib.F64, 0b1010, off, fb, C.byref(tot)) == 0
    a256 = lambda b: (b + 255) // 256 * 256  # noqa: E731
    assert fb[1] == vh // 6 * 12 * 8 and fb[3] == vh // 10 * 12 * 8 and fb[0] == fb[2] == 0
    assert off[2] == 0 and off[3] == fb[1]
    assert off[6] == a256(2 * fb[1]) and off[7] == off[6] + fb[3]
    assert tot.value == off[6] + a256(2 * fb[3])
    assert lib.dslash_face_layout(_lib.latt_arg(latt), 4, _lib.F32, 0, None, None, C.byref(tot)) == 0
    assert tot.value == 0
    assert lib.dslash_face_layout(_lib.latt_arg((4, 4, 4, 3)), 3, _lib.F32, 0b1000, None, None, None) == _lib.ERR_INVALID
    assert lib.dslash_face_layout(_lib.latt_arg(latt), 4, _lib.F16, 1, None, None, None) == _lib.ERR_INVALID
    st = lib.dslash_halo_interior(1, 2, 3, _lib.latt_arg(latt), 4, _lib.F32, 18, 1, 0, 16, None)
    assert st == _lib.ERR_INVALID and b"dir_mask" in lib.dslash_last_error_message()


def test_sloppy_switch_host_logic():
    """sse_su3dslash_set_sloppy exists in both libraries; only the FP64 one
    accepts it (no device work is done by the switch itself)."""
    from sse_wilson_dslash_amd.legacy import SSEDslash
    l32, l64 = SSEDslash(32), SSEDslash(64)
    assert l32.set_sloppy(True) == _lib.ERR_UNSUPPORTED and l32.last_error() == _lib.ERR_UNSUPPORTED
    assert l32.set_sloppy(False) == 0 and l32.last_error() == 0
    assert l32.set_sloppy_3d(True) == _lib.ERR_UNSUPPORTED
    assert l64.set_sloppy(True) == 0 and l64.set_sloppy(False) == 0
    assert l64.set_sloppy_3d(True) == 0 and l64.set_sloppy_3d(False) == 0


@pytest.mark.parametrize("prec", [32, 64])
def test_qdp_packer_headers(tmp_path, prec):
    """include/sse_dslash_qdp_packer*.h compile against a minimal QDP++-shaped
    stub (tests/qdp_stub) and write the transposed [site][mu] wire format."""
    exe = tmp_path / f"qpc{prec}"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", f"-DSSE_PRECISION={prec}",
           "-I", os.path.join(ROOT, "tests", "qdp_stub"), "-I", os.path.join(ROOT, "include"),
           "-o", str(exe), os.path.join(ROOT, "tests", "qdp_packer_check.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_staging_pool(tmp_path, sanitizer):
    """The drop-in API's pinned staging (csrc/stager.h) on the CPU through a
    HIP stub: byte-exact round trips across the pipelining threshold and
    restarted pools, under ThreadSanitizer and AddressSanitizer."""
    exe = tmp_path / f"stager_{sanitizer}"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Werror", f"-fsanitize={sanitizer}",
           "-I", os.path.join(ROOT, "tests", "hip_stub"), "-I", os.path.join(ROOT, "sse_wilson_dslash_amd", "csrc"),
           "-o", str(exe), os.path.join(ROOT, "tests", "stager_check.cpp"), "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=180)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0 and out.stdout.startswith("stager ok"), out.stdout + out.stderr[-2000:]


def _ref_unpack_gauge_field(volume, u_tmp, u):
    """lib/packer_nopad.c:47-69 step by step: for each cb, sites in pairs
    (ix, ix+1), the four directions' matrices at [mu][cb][ix] are read into a
    stack buffer v[8] and written back to the same [mu][cb][ix] positions of u."""
    src = u_tmp.reshape(-1, 18)
    dst = u.reshape(-1, 18)
    for cb in range(2):
        for ix in range(0, volume, 2):
            v = [None] * 8
            for mu in range(4):
                v[2 * mu] = src[ix + volume * (cb + 2 * mu)].copy()
                v[2 * mu + 1] = src[ix + 1 + volume * (cb + 2 * mu)].copy()
            for mu in range(4):
                dst[ix + volume * (cb + 2 * mu)] = v[2 * mu]
                dst[ix + 1 + volume * (cb + 2 * mu)] = v[2 * mu + 1]


@pytest.mark.parametrize("prec", [32, 64])
def test_c_gauge_packer(prec):
    """pack_gauge_field reorders [mu][cb][Vh] -> [cb][Vh][mu] without transposing
    (lib/packer_nopad.c:29-45).  unpack_gauge_field is byte-identical to the
    reference's (lib/packer_nopad.c:47-69, restated step by step above, which
    leaves u a copy of u_tmp); unpack_gauge_field_inverse undoes the pack.
    Integer-valued fills, so any misplaced matrix or element shows."""
    from sse_wilson_dslash_amd.legacy import SSEDslash
    lg = SSEDslash(prec)
    dt = np.float32 if prec == 32 else np.float64
    for vh in (12, 8, 128):
        u = np.arange(4 * 2 * vh * 18, dtype=dt).reshape(4, 2, vh, 18)
        out = np.zeros_like(u).reshape(2, vh, 4, 18)
        lg.pack_gauge_field(vh, u, out)
        assert np.array_equal(out, u.transpose(1, 2, 0, 3))
        back = np.zeros_like(u)
        lg.unpack_gauge_field_inverse(vh, out, back)
        assert np.array_equal(back, u)
        # the reference's unpacker: same bytes as the step-by-step restatement
        got = np.full_like(u, -1.0)
        want = np.full_like(u, -1.0)
        lg.unpack_gauge_field(vh, out, got)
        _ref_unpack_gauge_field(vh, out, want)
        assert got.tobytes() == want.tobytes()
        assert np.array_equal(got.reshape(-1), out.reshape(-1))
        # in place (u_tmp == u), as a caller may pass the same array
        same = out.copy()
        lg.unpack_gauge_field(vh, same, same)
        assert same.tobytes() == out.tobytes()


@pytest.mark.parametrize("prec", [32, 64])
def test_config_header_without_define(tmp_path, prec):
    """A caller compiled with the reference's include line and no -DSSE_PRECISION
    (include/sse_config.h:13 takes it from sse_config_internal.h) builds, links
    against the library of that precision and runs; SSEREAL has the library's
    width."""
    src = tmp_path / "cfg.c"
    src.write_text('#include "sse_dslash.h"\n#include "sse_dslash_3d.h"\n#include <stdio.h>\n'
                   "int main(void) {\n"
                   "  SSEREAL u[2 * 8 * 4 * 18], v[2 * 8 * 4 * 18];\n"
                   "  for (int i = 0; i < 2 * 8 * 4 * 18; ++i) u[i] = (SSEREAL)i;\n"
                   "  pack_gauge_field(8, (sse_u_mat_array*)u, (sse_u_mat_array*)v);\n"
                   "  unpack_gauge_field_inverse(8, (sse_u_mat_array*)v, (sse_u_mat_array*)u);\n"
                   "  for (int i = 0; i < 2 * 8 * 4 * 18; ++i) if (u[i] != (SSEREAL)i) return 2;\n"
                   '  printf("SSEREAL %d %d\\n", (int)sizeof(SSEREAL), SSE_PRECISION);\n'
                   "  return sse_su3dslash_last_error();\n}\n")
    exe = tmp_path / "cfg"
    libdir = os.path.join(ROOT, "sse_wilson_dslash_amd", "_lib")
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "include", f"f{prec}"), "-o", str(exe), str(src),
           "-L", libdir, f"-lsse_dslash_f{prec}", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split() == ["SSEREAL", str(prec // 8), str(prec)]


@pytest.mark.parametrize("nd,latt,msg", [(4, "[4,4,3,4]", b"4 dimensions"), (3, "[4,3,4,5]", b"3 dimensions")])
def test_legacy_odd_dimension_exits_1(nd, latt, msg):
    """lib/sse_su3dslash_32bit_scalar.c:103-110: odd global dimension -> stderr + exit(1)."""
    fn = "init_sse_su3dslash" if nd == 4 else "init_sse_su3dslash_3d"
    code = (f"import sys; sys.path.insert(0, {ROOT!r})\n"
            "from sse_wilson_dslash_amd.legacy import SSEDslash\n"
            f"SSEDslash(64).{fn}({latt})\n"
            "print('not reached')\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=120)
    assert r.returncode == 1
    assert msg in r.stderr and b"not reached" not in r.stdout
