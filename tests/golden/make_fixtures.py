"""Writes the oracle fixtures in this directory.

The reference ships no golden vectors (its tests draw from QDP++'s RNG) and its
scalar path cannot be built here, so these fixtures freeze the output of the
oracle (oracle/dslash_oracle.c, pinned by tests/test_oracle.py) on small seeded
lattices: the packed gauge and full-lattice spinor from the host generators, and
D psi (sse_su3dslash_wilson semantics: the target checkerboard written, the
source checkerboard left as the zero it started as) for isign +1 / -1 and source
checkerboard 0 / 1, in FP64 and in FP32 (the FP64 inputs rounded to FP32 first).

Later changes to the oracle or the generators are caught by comparing against
them, and the GPU tests check the HIP path against them without the oracle.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

# (name, nd, lattice, seed)
FIXTURES = [("oracle_4d_4x4x4x4", 4, (4, 4, 4, 4), 11),
            ("oracle_4d_6x2x4x4", 4, (6, 2, 4, 4), 12),
            ("oracle_3d_4x4x4x3", 3, (4, 4, 4, 3), 13)]


def build(name, nd, latt, seed):
    from oracle import oracle
    from sse_wilson_dslash_amd.device import gen_gauge_host, gen_spinor_host

    u64 = gen_gauge_host(latt, nd, seed=seed, dtype=np.float64)
    psi64 = gen_spinor_host(latt, nd, seed=seed, field=1, dtype=np.float64)
    out = {"nd": np.int32(nd), "latt": np.array(latt, np.int32), "seed": np.int32(seed), "u": u64, "psi": psi64}
    for tag, dt in (("f64", np.float64), ("f32", np.float32)):
        u, psi = u64.astype(dt), psi64.astype(dt)
        for isign in (1, -1):
            for cb in (0, 1):
                res = np.zeros_like(psi)
                oracle.dslash(u, psi, res, latt, isign, cb, nd=nd)
                out[f"res_{tag}_{'p' if isign > 0 else 'm'}{cb}"] = res
    return out


def main():
    for name, nd, latt, seed in FIXTURES:
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **build(name, nd, latt, seed))
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
