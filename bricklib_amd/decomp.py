"""The BrickDecomp layout (bk_decomp_layout) and a periodic grid in it.

BrickDecomp (include/brick-mpi.h:178-713) numbers a subdomain's bricks as
inner bricks, then the skin in a surface order (skin3d_good by default),
then the ghost bricks of each of the 26 neighbours, so that every exchange
moves contiguous runs of bricks.  :class:`DecompLayout` holds that
numbering; :class:`DecompGrid` is a one-subdomain periodic grid in it,
usable wherever a PeriodicGrid is
(sweeps, plans, the fused and deep-ghost forms): its ghost refresh copies
each neighbour's ghost run from the matching skin run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import bk_decomp_region, check, load, longs
from .bricks import (BrickInfo, DeviceBuffer, PeriodicGrid, movBrickInfo)


def _lexicographic():
    """The 27 regions in (k, j, i) lexicographic order, -1 first, i fastest,
    as signed axis numbers (1 = i), the empty centre included."""
    return [tuple(v for v in (i, 2 * j, 3 * k) if v)
            for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)]


# The surface orders of src/brick-mpi.cpp:25-136 in the reference's BitSet
# notation: an entry lists signed axis numbers, (1, -3) = +i and -k.
SKIN3D_GOOD = [(1,), (1, -3), (1, 2, -3), (1, 2), (1, 2, 3), (2, 3), (2,), (2, -3),
               (-1, 2, -3), (-1, 2), (-1, 2, 3), (-1, 3), (-1,), (-3,), (-1, -3),
               (-1, -2, -3), (-1, -2), (-1, -2, 3), (-2, 3), (-2,), (-2, -3), (1, -2, -3),
               (1, -2), (1, -2, 3), (1, 3), (3,)]
SKIN3D_NORMAL = _lexicographic()
SKIN3D_BAD = [e for r in _lexicographic() if r for e in ((), r)][1:]


def encode_skin(skin) -> np.ndarray:
    """Skin entries in BitSet notation -> int32 (i, j, k) triples of -1/0/+1
    (the encoding of bk_decomp_layout_skin)."""
    out = np.zeros((len(skin), 3), np.int32)
    for q, entry in enumerate(skin):
        for v in entry:
            a = abs(int(v)) - 1
            if not 0 <= a < 3 or out[q, a] != 0:
                raise ValueError(f"bad skin entry {entry!r}")
            out[q, a] = 1 if v > 0 else -1
    return out


class DecompLayout:
    """bk_decomp_layout for nb bricks per axis (i, j, k), ghost bricks deep,
    in the surface order `skin` (None: skin3d_good through
    bk_decomp_layout; else a list in BitSet notation, see SKIN3D_GOOD,
    through bk_decomp_layout_skin)."""

    def __init__(self, nb, ghost: int = 1, brick_bytes: int = 4096, page_bytes: int = 4096,
                 skin=None):
        self.nb = tuple(int(x) for x in nb)
        self.g = int(ghost)
        self.padded = tuple(x + 2 * self.g for x in self.nb)
        nbr, nreg = C.c_uint32(), C.c_uint32()
        sep = (C.c_uint32 * 3)()
        args = (longs(self.nb), self.g, C.c_size_t(brick_bytes), C.c_size_t(page_bytes))
        lib = load()
        if skin is None:
            def call(grid, adj, regs, sizes):
                return lib.bk_decomp_layout(*args, grid, adj, C.byref(nbr), sep, regs,
                                            C.byref(nreg))
        else:
            enc = encode_skin(skin)

            def call(grid, adj, regs, sizes):
                return lib.bk_decomp_layout_skin(*args, enc.ctypes.data_as(C.POINTER(C.c_int32)),
                                                 len(enc), grid, adj, C.byref(nbr), sep, regs,
                                                 C.byref(nreg), sizes)
        check(call(None, None, None, None))
        cells = int(np.prod(self.padded))
        grid = np.zeros(cells, np.uint32)
        adj = np.zeros(nbr.value * 27, np.uint32)
        regs = (bk_decomp_region * max(1, nreg.value))()
        sizes = np.zeros(max(1, len(skin or ())), np.uint32)
        check(call(grid.ctypes.data_as(_lib.u32p), adj.ctypes.data_as(_lib.u32p), regs,
                   sizes.ctypes.data_as(_lib.u32p)))
        #: ids per skin-list entry, padding included (given a skin list)
        self.skin_size = sizes[:len(skin)].tolist() if skin is not None else None
        self.nbricks = nbr.value
        self.sep = tuple(int(x) for x in sep)
        self.grid = grid.reshape(self.padded[2], self.padded[1], self.padded[0])
        self.adj = adj.reshape(-1, 27)
        self.regions = [dict(neighbor=tuple(r.neighbor), ghost_pos=r.ghost_pos,
                             skin_pos=r.skin_pos, len=r.len, skin_st=r.skin_st,
                             skin_ed=r.skin_ed, first_pad=r.first_pad, last_pad=r.last_pad)
                        for r in regs[:nreg.value]]

    def interior(self) -> np.ndarray:
        """The interior bricks in grid order (k slowest)."""
        g = self.g
        return np.ascontiguousarray(
            self.grid[g:g + self.nb[2], g:g + self.nb[1], g:g + self.nb[0]].reshape(-1))

    def periodic_pairs(self) -> np.ndarray:
        """(ghost, skin) brick pairs of the one-subdomain periodic exchange:
        every run's ghost bricks from its skin bricks, in order."""
        out = [np.stack([np.arange(r["ghost_pos"], r["ghost_pos"] + r["len"]),
                         np.arange(r["skin_pos"], r["skin_pos"] + r["len"])], 1)
               for r in self.regions if r["len"]]
        return np.ascontiguousarray(np.concatenate(out).astype(np.uint32).reshape(-1))


class DecompGrid(PeriodicGrid):
    """A periodic grid of nb bricks per axis with a one-brick ghost zone in
    the BrickDecomp numbering.  Fields have nbricks (the layout's id range)
    chunks; the refresh copies the 26 neighbours' ghost runs from their
    skin runs.  Sweep order: grid order over the interior.  `skin`: the
    surface order (BitSet notation, None = skin3d_good)."""

    def __init__(self, nb, stream=None, brick_bytes: int = 4096, skin=None):
        self.layout = L = DecompLayout(nb, 1, brick_bytes, skin=skin)
        self.nb = L.nb
        self.ghost = 1
        self.padded = L.padded
        self.grid = L.grid
        self.info = BrickInfo(L.adj)
        movBrickInfo(self.info, stream)
        self.interior = L.interior()
        self.pairs = L.periodic_pairs()
        self.d_interior = DeviceBuffer.from_array(self.interior, stream)
        self.d_pairs = DeviceBuffer.from_array(self.pairs, stream)
        self.d_grid = DeviceBuffer.from_array(self.grid.reshape(-1), stream)
        self.tile = 0
        self.sweep_list = self.interior
        self.d_sweep_list = self.d_interior
        self._plans = {}
        self._tuned = {}
        self.brick_layout = 0  # 8x8x8 bricks in the 8x8 fold
        self.extents = (8, 8, 8)

    @property
    def ghost_images(self):
        """Ghost images of every brick as CSR (for the fused refresh)."""
        if getattr(self, "_images", None) is None:
            pr = self.pairs.reshape(-1, 2)
            order = np.argsort(pr[:, 1], kind="stable")
            counts = np.bincount(pr[:, 1], minlength=self.nbricks)
            off = np.zeros(self.nbricks + 1, np.uint32)
            off[1:] = np.cumsum(counts)
            img = np.ascontiguousarray(pr[order, 0], np.uint32)
            self._images = (DeviceBuffer.from_array(off), DeviceBuffer.from_array(img))
        return self._images

    def fill_uniform(self, brick, seed: int, lo: float = -1.0, hi: float = 1.0,
                     stream=None) -> None:
        """The seeded field of bk_fill_uniform, placed in this numbering
        (generated in a PeriodicGrid of the same size and permuted)."""
        from .bricks import Brick
        ref = PeriodicGrid(self.nb, stream=stream, tile=0)
        tmp = Brick(ref.info, ref.field_storage(brick.dtype), 0)
        ref.fill_uniform(tmp, seed, lo, hi, stream=stream)
        host = tmp.storage.to_host(stream)
        dst = np.zeros((self.nbricks, brick.step), host.dtype)
        dst[self.interior, brick.offset:brick.offset + 512] = host[ref.interior, :512]
        brick.storage.buf.upload(dst.reshape(-1), stream)
