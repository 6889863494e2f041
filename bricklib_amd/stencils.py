"""Stencil descriptions and seeded coefficient sets.

Tap layouts follow include/bk_api.h: star coefficients are
``c[0]`` centre then ``(i+r, i-r, j+r, j-r, k+r, k-r)`` for ``r = 1..R``
(stencils/7pt.py:19-26, mpi13pt.py:18-31, mpi25pt.py:20-45); the cube has
one coefficient per offset in ``(dk, dj, di)`` lexicographic order
(the 5x5x5 offsets of stencils/mpi125pt.py:12-162).

Random numbers come from a counter-based SplitMix64 hash so the same seed
gives the same values in C, HIP and Python.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import BK_COND7, BK_CUBE125, BK_STAR7, BK_STAR13, BK_STAR25

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def hash64(seed: int, n: int) -> int:
    """SplitMix64 finaliser of seed + (n+1)*golden (same as the kernels)."""
    z = (seed + (n + 1) * GOLDEN) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def u01(seed: int, n: int) -> float:
    """53-bit uniform in [0, 1)."""
    return (hash64(seed, n) >> 11) * 2.0 ** -53


NAMES = {"7pt": BK_STAR7, "13pt": BK_STAR13, "25pt": BK_STAR25, "125pt": BK_CUBE125,
         "cond": BK_COND7}
RADIUS = {BK_STAR7: 1, BK_STAR13: 2, BK_STAR25: 4, BK_CUBE125: 2, BK_COND7: 1}
NCOEFF = {BK_STAR7: 7, BK_STAR13: 13, BK_STAR25: 25, BK_CUBE125: 125, BK_COND7: 7}
# points per stencil application and flops per point (2*taps - 1; cond is
# counted as the 7-point star, its max/compare/select are not counted)
TAPS = {BK_STAR7: 7, BK_STAR13: 13, BK_STAR25: 25, BK_CUBE125: 125, BK_COND7: 7}


def flops_per_point(stencil: int) -> int:
    return 2 * TAPS[stencil] - 1


def stencil_id(name_or_id) -> int:
    if isinstance(name_or_id, str):
        key = name_or_id.replace(".py", "").replace("mpi", "")
        if key not in NAMES:
            raise ValueError(f"unknown stencil {name_or_id!r}")
        return NAMES[key]
    return int(name_or_id)


def star_offsets(radius: int):
    """(di, dj, dk) of every star tap in coefficient order."""
    offs = [(0, 0, 0)]
    for r in range(1, radius + 1):
        offs += [(r, 0, 0), (-r, 0, 0), (0, r, 0), (0, -r, 0), (0, 0, r), (0, 0, -r)]
    return offs


def cube_offsets(radius: int = 2):
    return [(di, dj, dk) for dk in range(-radius, radius + 1)
            for dj in range(-radius, radius + 1) for di in range(-radius, radius + 1)]


def offsets(stencil: int):
    stencil = stencil_id(stencil)
    if stencil == BK_CUBE125:
        return cube_offsets(2)
    return star_offsets(RADIUS[stencil])


def coeff_plain7(seed: int) -> np.ndarray:
    """Seven coefficients U(0,1), normalised to sum 1 (the plain CPU case;
    single/cpu.cpp:12-17 draws U(0,1) without normalising)."""
    c = np.array([u01(seed, q) for q in range(7)])
    return c / c.sum()


def coeff_star(stencil: int, seed: int, normalise: bool = True) -> np.ndarray:
    """Star coefficients U(0,1) per (axis, distance), symmetric per axis
    (c(axis,+r) == c(axis,-r)), centre U(0,1); normalised to sum 1 so a
    hundred sweeps neither blow up nor vanish."""
    stencil = stencil_id(stencil)
    R = RADIUS[stencil]
    c = [u01(seed, 0)]
    for r in range(1, R + 1):
        ci, cj, ck = (u01(seed, 3 * r - 2), u01(seed, 3 * r - 1), u01(seed, 3 * r))
        c += [ci, ci, cj, cj, ck, ck]
    c = np.array(c)
    return c / c.sum() if normalise else c


def coeff_cube(seed: int) -> np.ndarray:
    """125 coefficients U(-1,1), divided by their L1 norm."""
    c = np.array([2.0 * u01(seed, q) - 1.0 for q in range(125)])
    return c / np.abs(c).sum()


def theta_cond(seed: int) -> float:
    """cond threshold U[-0.5, 0.5]."""
    return u01(seed, 1000) - 0.5


def default_coeff(stencil, seed: int) -> np.ndarray:
    stencil = stencil_id(stencil)
    if stencil == BK_CUBE125:
        return coeff_cube(seed)
    return coeff_star(stencil, seed)


@dataclass
class Stencil:
    """A stencil kind with its coefficients (and cond threshold)."""

    kind: int
    coeff: np.ndarray
    theta: float = 0.0
    cond_mode: int = 0

    @classmethod
    def make(cls, name_or_id, seed: int = 1, cond_mode: int = 1):
        kind = stencil_id(name_or_id)
        c = default_coeff(kind, seed)
        th = theta_cond(seed) if kind == BK_COND7 else 0.0
        return cls(kind, c, th, cond_mode if kind == BK_COND7 else 0)

    @property
    def radius(self) -> int:
        return RADIUS[self.kind]
