"""Host-side site orders of the legacy API (numpy, no GPU).

Canonical orders (full index = cb * Vh + idx):
  * 4D "cb2":  idx = xh + Lxh*(y + Ly*(z + Lz*t)), cb = (x+y+z+t) & 1
    (lib/shift_tables_scalar.c:99-160)
  * 3D "cb3d": idx = t + Lt*(xh + Lxh*(y + Ly*z)), cb = (x+y+z) & 1
    (lib/shift_tables_scalar_3d.c:95-176)
The lexicographic index lex = x + Lx*(y + Ly*(z + Lz*t)) names a site
independently of any order; the synthetic field generators are keyed by it.
"""
from __future__ import annotations

import numpy as np


def volume(latt) -> int:
    return int(latt[0]) * int(latt[1]) * int(latt[2]) * int(latt[3])


def coords_of_full(latt, nd: int, full: np.ndarray) -> np.ndarray:
    """Coordinates (n, 4) of full indices in the canonical order of nd."""
    L = [int(v) for v in latt]
    vh = volume(L) // 2
    full = np.asarray(full, dtype=np.int64)
    cb = full // vh
    i = full % vh
    lxh = L[0] // 2
    c = np.empty(full.shape + (4,), dtype=np.int64)
    if nd == 4:
        xh = i % lxh; i = i // lxh
        c[..., 1] = i % L[1]; i = i // L[1]
        c[..., 2] = i % L[2]; i = i // L[2]
        c[..., 3] = i
        c[..., 0] = 2 * xh + ((cb + c[..., 1] + c[..., 2] + c[..., 3]) & 1)
    else:
        c[..., 3] = i % L[3]; i = i // L[3]
        xh = i % lxh; i = i // lxh
        c[..., 1] = i % L[1]; i = i // L[1]
        c[..., 2] = i
        c[..., 0] = 2 * xh + ((cb + c[..., 1] + c[..., 2]) & 1)
    return c


def full_of_coords(latt, nd: int, c: np.ndarray) -> np.ndarray:
    """Canonical full index of coordinates (..., 4)."""
    L = [int(v) for v in latt]
    vh = volume(L) // 2
    c = np.asarray(c, dtype=np.int64)
    lxh = L[0] // 2
    x, y, z, t = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
    if nd == 4:
        cb = (x + y + z + t) & 1
        idx = x // 2 + lxh * (y + L[1] * (z + L[2] * t))
    else:
        cb = (x + y + z) & 1
        idx = t + L[3] * (x // 2 + lxh * (y + L[1] * z))
    return cb * vh + idx


def lex_of_coords(latt, c: np.ndarray) -> np.ndarray:
    L = [int(v) for v in latt]
    c = np.asarray(c, dtype=np.int64)
    return c[..., 0] + L[0] * (c[..., 1] + L[1] * (c[..., 2] + L[2] * c[..., 3]))


def lex_permutation(latt, nd: int) -> np.ndarray:
    """perm[full] = lexicographic index of the site with canonical index full."""
    return lex_of_coords(latt, coords_of_full(latt, nd, np.arange(volume(latt))))


def to_order(field: np.ndarray, site_table: np.ndarray) -> np.ndarray:
    """Re-order a canonical full-lattice field into a caller order.

    site_table[my] = caller index; returns out with out[site_table[my]] = field[my].
    """
    out = np.empty_like(field)
    out[np.asarray(site_table)] = field
    return out


def from_order(field: np.ndarray, site_table: np.ndarray) -> np.ndarray:
    """Inverse of to_order: canonical field from a caller-ordered one."""
    return field[np.asarray(site_table)]
