"""Batched Dslash: 32 right-hand sides on one gauge, and one persistent launch
over lattices of uneven volumes, checked field by field against the oracle and
against the single-RHS kernel; graph replay and the autotuner."""
import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import F32, F64
from sse_wilson_dslash_amd.batch import (BatchItem, RhsBatch, apply_batch, gen_batch, get_launch, set_launch,
                                        tune_apply, tune_batch)
from sse_wilson_dslash_amd.device import WilsonDslash, gen_gauge_host
from sse_wilson_dslash_amd.layout import volume
from tests import refimpl as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def make_item(latt, nrhs, prec=F32, G=18, isign=1, target_cb=0, seed=1):
    dt = np.float64 if prec == F64 else np.float32
    op = WilsonDslash(latt, 4, prec, G)
    u = gen_gauge_host(latt, 4, seed=seed, dtype=dt)
    op.load_gauge(u)
    psi, host = gen_batch(op, nrhs, 1 - target_cb, seed=seed, field0=10)
    res = RhsBatch(op, nrhs, zero=True)
    return BatchItem(op, psi, res, isign, target_cb), u, host


def check_item(oracle, item, u, host, tol, ks=None):
    op = item.op
    vh = op.vh
    t = slice(item.target_cb * vh, (item.target_cb + 1) * vh)
    dt = u.dtype
    for k in (range(item.psi.nrhs) if ks is None else ks):
        got = op.unpack_cb(item.res.field(k), item.target_cb, dtype=dt)
        want = oracle.dslash(u, host[k], np.zeros_like(host[k]), op.latt, item.isign, 1 - item.target_cb)
        e = R.rel_err(got.reshape(-1, 24)[t], want.reshape(-1, 24)[t])
        assert e < tol, (op.latt, k, e)


@pytest.fixture
def launch_table():
    """Restore the batched launch table after a test that sets it."""
    saved = {(p, g, per): get_launch(p, g, per) for p in (F32, F64) for g in (18, 12) for per in (False, True)}
    yield
    for (p, g, per), c in saved.items():
        set_launch(p, g, per, c["rhs_block"], c["block_threads"], c["grid_blocks"], c["t_block"])


@pytest.mark.parametrize("rhs_block,t_block", [(1, 1), (2, 1), (4, 1), (1, 4), (2, 4)])
@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("latt", [(8, 8, 8, 16), (8, 4, 6, 10)])
def test_mrhs_matches_single_rhs_kernel(oracle, launch_table, rhs_block, t_block, block, latt):
    """Both tile shapes (one site per thread; t-marching over 4 slices, with a
    partial last chunk when T = 10) against the single-RHS kernel."""
    item, u, host = make_item(latt, 7, isign=-1, target_cb=1)  # 7: exercises a partial RHS block
    set_launch(F32, 18, False, rhs_block, block, 0, t_block)
    apply_batch([item], rhs_block=rhs_block, block_threads=block)
    torch.cuda.synchronize()
    op = item.op
    for k in range(7):
        single = op.new_spinor(False)
        op.apply_cb(single, item.psi.field(k), -1, 1)
        a = op.unpack_cb(single, 1, dtype=np.float32)
        b = op.unpack_cb(item.res.field(k), 1, dtype=np.float32)
        assert R.rel_err(b, a) < 1e-6
    check_item(oracle, item, u, host, 1e-5, ks=[0, 6])


@pytest.mark.parametrize("prec,G", [(F32, 18), (F32, 12), (F64, 18)])
def test_32_rhs_24c48(oracle, prec, G):
    """Config: 32 Gaussian right-hand sides on one 24^3 x 48 SU(3) gauge."""
    latt = (24, 24, 24, 48)
    print(f"[mrhs 32 x 24^3x48] prec={prec} G={G}", flush=True)
    item, u, host = make_item(latt, 32, prec, G)
    apply_batch([item])
    torch.cuda.synchronize()
    check_item(oracle, item, u, host, 1e-5 if prec == F32 else 1e-13)


@pytest.mark.parametrize("rhs_block,nrhs", [(1, 7), (2, 8)])
@pytest.mark.parametrize("latt", [(8, 8, 8, 16), (8, 4, 6, 10), (12, 6, 4, 8)])
@pytest.mark.parametrize("isign", [1, -1])
def test_lds_staged_tiles(oracle, launch_table, rhs_block, nrhs, latt, isign):
    """t_block 8: 128-thread blocks walking 8 slices with an LDS slab (T = 10 leaves
    a 2-slice chunk; 12x6x4 has 144 spatial sites, so one block is partial)."""
    item, u, host = make_item(latt, nrhs, isign=isign, target_cb=0)
    set_launch(F32, 18, False, rhs_block, 128, 0, 8)
    apply_batch([item], rhs_block=rhs_block, block_threads=128)
    torch.cuda.synchronize()
    op = item.op
    for k in range(nrhs):
        single = op.new_spinor(False)
        op.apply_cb(single, item.psi.field(k), isign, 0)
        a = op.unpack_cb(single, 0, dtype=np.float32)
        b = op.unpack_cb(item.res.field(k), 0, dtype=np.float32)
        assert R.rel_err(b, a) < 1e-6, (k, R.rel_err(b, a))
    check_item(oracle, item, u, host, 1e-5, ks=[0, nrhs - 1])


def test_lds_tiles_persistent_and_fallback(oracle, launch_table):
    """A persistent LDS-staged launch over two lattices with one isign; with
    mixed isign the tabled LDS shape falls back to the register march."""
    items = [make_item((8, 8, 8, 8), 4, F32, 18, 1, 1, 6), make_item((16, 8, 8, 16), 6, F32, 18, 1, 1, 7)]
    set_launch(F32, 18, True, 2, 128, 64, 8)
    apply_batch([it[0] for it in items], grid_blocks=64)
    torch.cuda.synchronize()
    for item, u, host in items:
        check_item(oracle, item, u, host, 1e-5)
    mixed = uneven_items()
    set_launch(F32, 18, False, 2, 128, 0, 8)
    apply_batch([it[0] for it in mixed])
    torch.cuda.synchronize()
    for item, u, host in mixed:
        check_item(oracle, item, u, host, 1e-5, ks=[0, 3])


@pytest.mark.parametrize("rhs_block,nrhs", [(1, 7), (2, 8)])
@pytest.mark.parametrize("block", [128, 192, 256])
@pytest.mark.parametrize("latt", [(8, 8, 8, 16), (8, 4, 6, 10), (12, 6, 4, 8), (24, 6, 10, 4)])
@pytest.mark.parametrize("isign", [1, -1])
def test_box_tiles(oracle, launch_table, rhs_block, nrhs, block, latt, isign):
    """t_block 16: box tiles of whole x rows x ty x tz (boxes spanning all of y or
    z wrap inside the box; T = 10 leaves a 2-slice chunk; 6, 12 and 18 x rows
    leave idle lanes in the block) against the single-RHS kernel and the oracle."""
    print(f"[box tiles] latt={latt} block={block} R={rhs_block} isign={isign}", flush=True)
    item, u, host = make_item(latt, nrhs, isign=isign, target_cb=1)
    set_launch(F32, 18, False, rhs_block, block, 0, 16)
    apply_batch([item], rhs_block=rhs_block, block_threads=block)
    torch.cuda.synchronize()
    op = item.op
    for k in range(nrhs):
        single = op.new_spinor(False)
        op.apply_cb(single, item.psi.field(k), isign, 1)
        a = op.unpack_cb(single, 1, dtype=np.float32)
        b = op.unpack_cb(item.res.field(k), 1, dtype=np.float32)
        assert R.rel_err(b, a) < 1e-6, (k, R.rel_err(b, a))
    check_item(oracle, item, u, host, 1e-5, ks=[0, nrhs - 1])


@pytest.mark.parametrize("rhs_block,block", [(1, 512), (1, 768), (2, 384)])
@pytest.mark.parametrize("latt,nrhs", [((8, 8, 8, 16), 4), ((24, 24, 24, 16), 6), ((12, 6, 4, 8), 2),
                                       ((24, 6, 10, 4), 4)])
@pytest.mark.parametrize("isign", [1, -1])
def test_wide_box_tiles(oracle, launch_table, rhs_block, block, latt, nrhs, isign):
    """t_block 16 with more than 256 threads (rhs_block x threads <= 768): on
    24^3 x 16 the boxes are 12 x 8 x 8 (768), 12 x 6 x 6 (512) or 12 x 4 x 8 (384)
    sites with hops leaving the box in y and z; on the small lattices one box spans
    the slice (wrap inside the box, idle lanes); T = 4 / 8 / 16 give one or two
    t-chunks.  Against the single-RHS kernel and the oracle."""
    print(f"[wide box tiles] latt={latt} block={block} R={rhs_block} nrhs={nrhs} isign={isign}", flush=True)
    item, u, host = make_item(latt, nrhs, isign=isign, target_cb=0)
    set_launch(F32, 18, False, rhs_block, block, 0, 16)
    apply_batch([item], rhs_block=rhs_block, block_threads=block)
    torch.cuda.synchronize()
    op = item.op
    for k in range(nrhs):
        single = op.new_spinor(False)
        op.apply_cb(single, item.psi.field(k), isign, 0)
        a = op.unpack_cb(single, 0, dtype=np.float32)
        b = op.unpack_cb(item.res.field(k), 0, dtype=np.float32)
        assert R.rel_err(b, a) < 1e-6, (k, R.rel_err(b, a))
    check_item(oracle, item, u, host, 1e-5, ks=[0, nrhs - 1])


def test_wide_box_fallback(oracle, launch_table):
    """A tabled wide shape (768 threads, one field) on a launch that cannot take
    box tiles (mixed isign) runs its fallback with 256 threads; rhs_block 2 with
    more than 384 threads and FP64 above 256 are rejected."""
    from sse_wilson_dslash_amd._lib import DslashError
    set_launch(F32, 18, False, 1, 768, 0, 16)
    mixed = uneven_items()
    apply_batch([it[0] for it in mixed])
    torch.cuda.synchronize()
    for item, u, host in mixed:
        check_item(oracle, item, u, host, 1e-5, ks=[0, 3])
    with pytest.raises(DslashError):
        set_launch(F32, 18, False, 2, 512, 0, 16)
    with pytest.raises(DslashError):
        set_launch(F64, 18, False, 1, 512, 0, 16)


def test_box_tiles_persistent_and_fallback(oracle, launch_table):
    """A persistent box-tile launch over three lattices (4, 8 and 12 sites per x
    row) with one isign; a lattice with Lx/2 > block_threads falls back to the
    LDS tiles, mixed isign to the register march."""
    items = [make_item(latt, 4, F32, 18, -1, 0, seed)
             for latt, seed in (((8, 8, 8, 8), 6), ((16, 16, 16, 8), 7), ((24, 24, 24, 8), 8))]
    set_launch(F32, 18, True, 2, 192, 96, 16)
    apply_batch([it[0] for it in items], grid_blocks=96)
    torch.cuda.synchronize()
    for item, u, host in items:
        check_item(oracle, item, u, host, 1e-5)
    wide = [make_item((264, 2, 2, 4), 2, F32, 18, 1, 0, 9)]
    set_launch(F32, 18, False, 2, 128, 0, 16)
    apply_batch([it[0] for it in wide])
    torch.cuda.synchronize()
    for item, u, host in wide:
        check_item(oracle, item, u, host, 1e-5)
    mixed = uneven_items()
    apply_batch([it[0] for it in mixed])
    torch.cuda.synchronize()
    for item, u, host in mixed:
        check_item(oracle, item, u, host, 1e-5, ks=[0, 3])


def uneven_items():
    out = []
    for latt, isign, tcb, seed in (((8, 8, 8, 8), 1, 0, 3), ((16, 16, 16, 32), -1, 1, 4), ((24, 24, 24, 48), 1, 1, 5)):
        out.append(make_item(latt, 4, F32, 18, isign, tcb, seed))
    return out


@pytest.mark.parametrize("grid", [0, 8, 13, 512, 2048])
@pytest.mark.parametrize("t_block", [1, 4])
def test_persistent_uneven_lattices(oracle, launch_table, grid, t_block):
    """Config: one launch over 8^4, 16^3x32 and 24^3x48 with 4 RHS each."""
    print(f"[persistent] grid={grid} t_block={t_block}", flush=True)
    items = uneven_items()
    set_launch(F32, 18, grid > 0, 2, 128, grid, t_block)
    apply_batch([it[0] for it in items], rhs_block=2, block_threads=128, grid_blocks=grid)
    torch.cuda.synchronize()
    for item, u, host in items:
        check_item(oracle, item, u, host, 1e-5)


def test_persistent_graph_replay(oracle):
    """The persistent launch captured in a graph and replayed with new sources
    written eagerly between replays; every replay matches the eager result."""
    items = uneven_items()
    its = [it[0] for it in items]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        apply_batch(its, rhs_block=2, block_threads=128, grid_blocks=1024)  # warm-up outside capture
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        apply_batch(its, rhs_block=2, block_threads=128, grid_blocks=1024)
    for rep in range(3):
        # new sources: rotate the fields of every batch (eager copies on the current stream)
        for it in its:
            first = it.psi.field(0).clone()
            for k in range(it.psi.nrhs - 1):
                it.psi.field(k).copy_(it.psi.field(k + 1))
            it.psi.field(it.psi.nrhs - 1).copy_(first)
        graph.replay()
        torch.cuda.synchronize()
        for it in its:
            for k in range(it.psi.nrhs):
                ref = it.op.new_spinor(False)
                it.op.apply_cb(ref, it.psi.field(k), it.isign, it.target_cb)
                torch.cuda.synchronize()
                assert torch.equal(ref, it.res.field(k)) or R.rel_err(
                    it.op.unpack_cb(ref, it.target_cb, np.float32),
                    it.op.unpack_cb(it.res.field(k), it.target_cb, np.float32)) < 1e-6
    # the last replay still matches the oracle
    for item, u, host in items:
        n = item.psi.nrhs
        rolled = np.roll(host, -3, axis=0)  # three rotations
        check_item(oracle, item, u, rolled, 1e-5, ks=range(n))


def test_autotuner(oracle):
    items = uneven_items()
    its = [it[0] for it in items]
    for persistent in (False, True):
        r = tune_batch(its, persistent)
        assert r["rhs_block"] in (1, 2, 4) and r["block_threads"] in (64, 128, 192, 256) and r["ms"] > 0
        assert r["t_block"] in (1, 4, 8, 16) and not (r["t_block"] > 1 and r["rhs_block"] == 4)
        assert (r["grid_blocks"] > 0) == persistent
        for it in its:
            it.res.buf.zero_()
        apply_batch(its, grid_blocks=r["grid_blocks"])  # tuned rhs_block / block_threads
        torch.cuda.synchronize()
        for item, u, host in items:
            check_item(oracle, item, u, host, 1e-5, ks=[0, 3])
    op, u = items[2][0].op, items[2][1]
    psi = op.gen_spinor(seed=7)
    res = op.new_spinor(zero=True)
    # the tuner rewrites the process-wide single-RHS entry of (4D, F32, 18 reals);
    # later tests expect the default shape, so it is restored afterwards
    import ctypes as C
    from sse_wilson_dslash_amd import _lib
    lib = _lib.core()
    b0, r0, m0 = C.c_int(), C.c_int(), C.c_int()
    _lib.check(lib.dslash_get_launch(4, F32, 18, C.byref(b0), C.byref(r0)))
    _lib.check(lib.dslash_get_launch_march(4, F32, 18, C.byref(m0)))
    try:
        t = tune_apply(op, res, psi)
        assert t["block_threads"] in (64, 128, 192, 256) and t["xcd_remap"] in (0, 1) and t["ms"] > 0
        assert t["march"] in (1, 4, 8)
        # the tuned shape still computes the same operator
        op.apply(res, psi, 1, 0)
        got = op.unpack_spinor(res, dtype=np.float32)
        h = op.unpack_spinor(psi, dtype=np.float32)
        want = oracle.dslash(u, h, np.zeros_like(h), op.latt, 1, 0)
        assert R.rel_err(got[op.vh:], want[op.vh:]) < 1e-5
    finally:
        _lib.check(lib.dslash_set_launch(4, F32, 18, b0.value, r0.value))
        _lib.check(lib.dslash_set_launch_march(4, F32, 18, m0.value))
