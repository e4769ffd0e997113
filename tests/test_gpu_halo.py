"""GPU tests of the domain-decomposition split of the Dslash (decomp/decomp_hvv
faces, interior, mvv_recons/recons exterior; the stages of
lib/sse_su3dslash_32bit_parscalar.c:108-820), through the C ABI.

* One GPU, every set of partitioned directions, periodic self-exchange: equal
  to the oracle (FP64 1e-13, FP32 1e-5) and to the fused kernel.
* Several emulated nodes on one GPU: the global lattice is cut into local
  lattices, each with its own fields and operator; faces move between them with
  device copies exactly as a multi-node exchange would route them, and the
  assembled result must equal the oracle on the global lattice.
"""
import itertools

import numpy as np
import pytest
import torch

from sse_wilson_dslash_amd import F32, F64
from sse_wilson_dslash_amd.device import WilsonDslash, gen_gauge_host, gen_spinor_host
from sse_wilson_dslash_amd.halo import HaloDslash
from sse_wilson_dslash_amd.layout import coords_of_full, full_of_coords, volume
from tests import refimpl as R

pytestmark = pytest.mark.gpu

TOL = {F64: 1e-13, F32: 1e-5}


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")


def target(a, latt, cb):
    vh = volume(latt) // 2
    return a.reshape(-1, 24)[(1 - cb) * vh:(2 - cb) * vh]


@pytest.mark.parametrize("nd,latt", [(4, (4, 4, 4, 4)), (4, (2, 6, 4, 2)), (4, (6, 2, 2, 10)),
                                     (3, (4, 2, 6, 3)), (3, (4, 4, 2, 4))])
@pytest.mark.parametrize("prec,G", [(F64, 18), (F32, 18), (F32, 12), (F64, 12)])
def test_self_exchange_every_mask(oracle, nd, latt, prec, G):
    dt = np.float64 if prec == F64 else np.float32
    print(f"[halo self-exchange] nd={nd} latt={latt} prec={prec} G={G}", flush=True)
    u = gen_gauge_host(latt, nd, seed=5, dtype=dt)
    psi = gen_spinor_host(latt, nd, seed=5, field=2, dtype=dt)
    op = WilsonDslash(latt, nd=nd, prec=prec, gauge_reals=G)
    op.load_gauge(u)
    d_psi = op.pack_spinor(psi)
    vh = volume(latt) // 2
    for isign in (1, -1):
        for cb in (0, 1):
            want = target(oracle.dslash(u, psi, np.zeros_like(psi), latt, isign, cb, nd=nd), latt, cb)
            fused = op.new_spinor(zero=True)
            op.apply(fused, d_psi, isign, cb)
            f = target(op.unpack_spinor(fused, dtype=dt), latt, cb)
            for mask in range(1 << nd):
                h = HaloDslash(op, mask)
                res = op.pack_spinor(np.full_like(psi, 7.0))
                h.apply(res, d_psi, isign, cb)
                full = op.unpack_spinor(res, dtype=dt)
                got = target(full, latt, cb)
                e2, einf = R.rel_err(got, want), R.max_rel_err(got, want)
                assert e2 < TOL[prec] and einf < TOL[prec], (mask, isign, cb, e2, einf)
                assert R.max_rel_err(got, f) < TOL[prec], (mask, isign, cb)
                assert np.all(full.reshape(-1, 24)[cb * vh:(cb + 1) * vh] == 7.0), "source cb of res modified"


def test_interior_alone_misses_exactly_the_boundary_hops(oracle):
    """interior + exterior(zero faces) leaves out the boundary terms: the
    difference to the full operator is supported on boundary target sites only."""
    latt, nd, prec = (4, 4, 6, 4), 4, F64
    u = gen_gauge_host(latt, nd, seed=6)
    psi = gen_spinor_host(latt, nd, seed=6)
    op = WilsonDslash(latt, nd=nd, prec=prec)
    op.load_gauge(u)
    d_psi = op.pack_spinor(psi)
    mask = 0b0101  # x and z partitioned
    h = HaloDslash(op, mask)
    res = op.new_spinor(zero=True)
    h.interior(res, d_psi, 1, 0)
    got = target(op.unpack_spinor(res), latt, 0)
    want = target(oracle.dslash(u, psi, np.zeros_like(psi), latt, 1, 0), latt, 0)
    vh = volume(latt) // 2
    c = coords_of_full(latt, nd, np.arange(vh, 2 * vh))  # target cb 1
    boundary = np.zeros(vh, bool)
    for mu in (0, 2):
        boundary |= (c[:, mu] == 0) | (c[:, mu] == latt[mu] - 1)
    d = np.abs(got - want).max(axis=1)
    assert d[~boundary].max() < 1e-13
    assert d[boundary].min() > 1e-6


@pytest.mark.parametrize("nd,glatt,split", [(4, (4, 4, 4, 8), (1, 1, 1, 2)), (4, (4, 4, 4, 12), (1, 1, 1, 3)),
                                            (4, (8, 4, 4, 4), (2, 1, 1, 1)), (4, (8, 4, 8, 8), (2, 1, 2, 2)),
                                            (4, (4, 12, 4, 4), (1, 3, 1, 1)), (3, (4, 12, 4, 3), (1, 3, 1, 1)),
                                            (3, (8, 4, 8, 5), (2, 1, 2, 1)),
                                            # degenerate local extents, as the reference's single-host
                                            # multi-rank runs force (tests/scripts/test_all.sh: {2,6,6,2})
                                            (4, (4, 4, 4, 4), (2, 2, 2, 2)), (4, (2, 6, 6, 2), (1, 3, 3, 1)),
                                            (3, (2, 6, 6, 3), (1, 3, 3, 1))])
@pytest.mark.parametrize("prec", [F64, F32])
def test_multi_node_emulation(oracle, nd, glatt, split, prec):
    dt = np.float64 if prec == F64 else np.float32
    print(f"[halo nodes] nd={nd} global={glatt} split={split} prec={prec}", flush=True)
    loc = tuple(g // s for g, s in zip(glatt, split))
    assert all(l % 2 == 0 for l in loc[:nd]), "local extents of hopping directions must stay even"
    mask = sum(1 << mu for mu in range(nd) if split[mu] > 1)
    u = gen_gauge_host(glatt, nd, seed=7, dtype=dt)
    psi = gen_spinor_host(glatt, nd, seed=7, field=3, dtype=dt)
    Vl = volume(loc)
    lc = coords_of_full(loc, nd, np.arange(Vl))
    ranks = list(itertools.product(*[range(s) for s in split]))
    nodes = {}
    for r in ranks:
        gfull = full_of_coords(glatt, nd, lc + np.array(r) * np.array(loc))
        op = WilsonDslash(loc, nd=nd, prec=prec)
        op.load_gauge(np.ascontiguousarray(u.reshape(-1, 72)[gfull]))
        nodes[r] = {"gfull": gfull, "op": op, "halo": HaloDslash(op, mask),
                    "psi": op.pack_spinor(np.ascontiguousarray(psi.reshape(-1, 24)[gfull]))}
    for isign in (1, -1):
        for cb in (0, 1):
            for n in nodes.values():
                n["send"] = n["halo"].new_faces()
                n["halo"].pack(n["send"], n["psi"], isign, cb)
            for r, n in nodes.items():
                h = n["halo"]
                recv = h.new_faces()
                for mu in range(nd):
                    if not (mask >> mu) & 1:
                        continue
                    up = list(r); up[mu] = (r[mu] + 1) % split[mu]
                    dn = list(r); dn[mu] = (r[mu] - 1) % split[mu]
                    h.face(recv, mu, True).copy_(nodes[tuple(up)]["halo"].face(nodes[tuple(up)]["send"], mu, True))
                    h.face(recv, mu, False).copy_(nodes[tuple(dn)]["halo"].face(nodes[tuple(dn)]["send"], mu, False))
                n["res"] = n["op"].new_spinor(zero=True)
                h.interior(n["res"], n["psi"], isign, cb)
                h.exterior(n["res"], recv, isign, cb)
            out = np.zeros((volume(glatt), 24), dtype=dt)
            for n in nodes.values():
                out[n["gfull"]] = n["op"].unpack_spinor(n["res"], dtype=dt).reshape(-1, 24)
            want = oracle.dslash(u, psi, np.zeros_like(psi), glatt, isign, cb, nd=nd)
            got, w = target(out, glatt, cb), target(want, glatt, cb)
            e2, einf = R.rel_err(got, w), R.max_rel_err(got, w)
            print(f"  isign={isign:+d} cb={cb} rel L2 {e2:.2e}", flush=True)
            assert e2 < TOL[prec] and einf < TOL[prec]
