"""Python mirror of bricklib's host interface over libbricks.so.

Names follow the reference (CtopCsUtahEdu/bricklib) so code and tests read
like its drivers: ``init_grid`` (bricksetup.h:73-89), ``BrickInfo``
(brick.h:96-127), ``BrickStorage.allocate`` (brick.h:68-75, here resident in
HBM), ``Brick`` (brick.h:353-395, a field view: storage + offset),
``movBrickInfo`` / ``movBrickStorage`` (brick-gpu.h:42-103),
``copyToBrick`` / ``copyFromBrick`` (bricksetup.h:172-221).  Every GPU
operation goes through the C-ABI; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BK_BF16, BK_F32, BK_F64, bk_sweep_desc, check, load, longs)
from .stencils import Stencil, stencil_id

DTYPE_NP = {BK_F64: np.float64, BK_F32: np.float32, BK_BF16: np.uint16}
DTYPE_SIZE = {BK_F64: 8, BK_F32: 4, BK_BF16: 2}
BRICK_ELEMS = 512  # 8 x 8 x 8


def _vp(x) -> C.c_void_p:
    if x is None:
        return C.c_void_p(None)
    if isinstance(x, int):
        return C.c_void_p(x)
    if isinstance(x, np.ndarray):
        return x.ctypes.data_as(C.c_void_p)
    return C.c_void_p(int(x))


# ---------------------------------------------------------------------------
# device memory, streams, events
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """An HBM allocation owned by Python (freed with the object)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().bk_alloc(C.c_size_t(self.nbytes), C.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, arr: np.ndarray, stream=None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr, stream)
        return buf

    def upload(self, arr: np.ndarray, stream=None, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        check(load().bk_memcpy_h2d(_vp(self.ptr + offset), _vp(arr), C.c_size_t(arr.nbytes),
                                   _vp(stream)))

    def download(self, dtype, count: int | None = None, stream=None, offset: int = 0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download past the end of the buffer")
        check(load().bk_memcpy_d2h(_vp(out), _vp(self.ptr + offset), C.c_size_t(out.nbytes),
                                   _vp(stream)))
        return out

    def zero(self, stream=None) -> None:
        check(load().bk_zero(_vp(self.ptr), C.c_size_t(self.nbytes), _vp(stream)))

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            load().bk_free(_vp(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream (default priority)."""

    def __init__(self):
        h = C.c_void_p()
        check(load().bk_stream_create(C.byref(h)))
        self.handle = h.value

    def sync(self) -> None:
        check(load().bk_stream_sync(_vp(self.handle)))

    def __int__(self):
        return int(self.handle or 0)

    def __del__(self):
        try:
            if self.handle:
                load().bk_stream_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(load().bk_event_create(C.byref(h)))
        self.handle = h.value

    def record(self, stream=None) -> None:
        check(load().bk_event_record(_vp(self.handle), _vp(stream)))

    def sync(self) -> None:
        check(load().bk_event_sync(_vp(self.handle)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(load().bk_event_elapsed_ms(_vp(self.handle), _vp(end.handle), C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                load().bk_event_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int()
    rc = load().bk_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def device_sync() -> None:
    check(load().bk_device_sync())


def _stream_of(stream):
    if stream is None:
        return None
    return int(stream)


# ---------------------------------------------------------------------------
# BrickInfo / init_grid / BrickStorage / Brick
# ---------------------------------------------------------------------------
class BrickInfo:
    """BrickInfo<3>: the adjacency list adj[nbricks][27] (uint32), on the
    host (``adj``) and, after movBrickInfo, in HBM (``dev``)."""

    def __init__(self, adj: np.ndarray):
        self.adj = np.ascontiguousarray(adj, dtype=np.uint32).reshape(-1, 27)
        self.dev: DeviceBuffer | None = None

    @property
    def nbricks(self) -> int:
        return self.adj.shape[0]

    @property
    def dptr(self) -> int:
        if self.dev is None:
            raise RuntimeError("BrickInfo is not on the device; call movBrickInfo first")
        return self.dev.ptr


def init_grid(dims, periodic: bool = False):
    """init_grid<3>(grid_ptr, dims) -> (grid, BrickInfo).  ``dims`` are brick
    counts contiguous-first (i, j, k); ``grid`` has shape (k, j, i)."""
    dims = [int(d) for d in dims]
    n = dims[0] * dims[1] * dims[2]
    grid = np.empty(n, dtype=np.uint32)
    adj = np.empty(n * 27, dtype=np.uint32)
    check(load().bk_init_grid(longs(dims), int(bool(periodic)),
                              grid.ctypes.data_as(_lib.u32p), adj.ctypes.data_as(_lib.u32p)))
    return grid.reshape(dims[2], dims[1], dims[0]), BrickInfo(adj)


def movBrickInfo(info: BrickInfo, stream=None) -> BrickInfo:
    """Copy the adjacency list into HBM (brick-gpu.h:42-55)."""
    info.dev = DeviceBuffer.from_array(info.adj, _stream_of(stream))
    return info


@dataclass
class BrickStorage:
    """Device-resident BrickStorage: ``chunks`` x ``step`` elements."""

    buf: DeviceBuffer
    chunks: int
    step: int
    dtype: int = BK_F64

    @classmethod
    def allocate(cls, chunks: int, step: int, dtype: int = BK_F64) -> "BrickStorage":
        p = C.c_void_p()
        check(load().bk_storage_alloc(int(dtype), C.c_long(int(chunks)), C.c_size_t(int(step)),
                                      C.byref(p)))
        buf = DeviceBuffer.__new__(DeviceBuffer)
        es = DTYPE_SIZE[dtype]
        buf.nbytes = ((int(chunks) * int(step) * es + 2047) // 2048) * 2048
        buf.ptr = p.value or 0
        return cls(buf, int(chunks), int(step), int(dtype))

    @property
    def dat(self) -> int:
        return self.buf.ptr

    def to_host(self, stream=None) -> np.ndarray:
        """movBrickStorage(..., DeviceToHost) (brick-gpu.h:84-103)."""
        n = self.chunks * self.step
        return self.buf.download(DTYPE_NP[self.dtype], n, _stream_of(stream)).reshape(
            self.chunks, self.step)


def movBrickStorage(host: np.ndarray, dtype: int = BK_F64, stream=None) -> BrickStorage:
    """Copy a host storage array (chunks, step) into a new device
    BrickStorage (brick-gpu.h:84-103)."""
    host = np.ascontiguousarray(host, dtype=DTYPE_NP[dtype])
    chunks, step = host.shape
    st = BrickStorage.allocate(chunks, step, dtype)
    st.buf.upload(host.reshape(-1), _stream_of(stream))
    return st


@dataclass
class Brick:
    """Brick<Dim<8,8,8>,Dim<8,8>>(bInfo, bStorage, offset): one field of a
    storage, ``offset`` elements into each chunk (brick.h:390-394)."""

    info: BrickInfo
    storage: BrickStorage
    offset: int = 0

    @property
    def ptr(self) -> int:
        return self.storage.dat + self.offset * DTYPE_SIZE[self.storage.dtype]

    @property
    def step(self) -> int:
        return self.storage.step

    @property
    def dtype(self) -> int:
        return self.storage.dtype

    def bricks_host(self, stream=None) -> np.ndarray:
        """(chunks, 512) view of this field's bricks, copied to the host."""
        return self.storage.to_host(stream)[:, self.offset:self.offset + BRICK_ELEMS]


def copyToBrick(dimlist, padding, ghost, arr: DeviceBuffer, grid: DeviceBuffer, brick: Brick,
                stream=None) -> None:
    """copyToBrick<3> on the device (bricksetup.h:172-200)."""
    check(load().bk_copy_to_brick(brick.dtype, longs(dimlist), longs(padding), longs(ghost),
                                  _vp(arr.ptr), _vp(grid.ptr), _vp(brick.ptr),
                                  C.c_size_t(brick.step), _vp(_stream_of(stream))))


def copyFromBrick(dimlist, padding, ghost, arr: DeviceBuffer, grid: DeviceBuffer, brick: Brick,
                  stream=None) -> None:
    """copyFromBrick<3> on the device (bricksetup.h:213-221)."""
    check(load().bk_copy_from_brick(brick.dtype, longs(dimlist), longs(padding), longs(ghost),
                                    _vp(arr.ptr), _vp(grid.ptr), _vp(brick.ptr),
                                    C.c_size_t(brick.step), _vp(_stream_of(stream))))


# ---------------------------------------------------------------------------
# execution plans
# ---------------------------------------------------------------------------
PLAN_SHAPES = {"column": _lib.BK_PLAN_COLUMN, "quad": _lib.BK_PLAN_QUAD}


def plan_host(info: BrickInfo, blist: np.ndarray, shape: str, kcol: int) -> dict:
    """The plan's arrays built on the host (bk_plan_build_host)."""
    blist = np.ascontiguousarray(blist, np.uint32)
    lib = load()
    nu, nc = C.c_uint32(), C.c_uint32()
    args = (info.adj.ctypes.data_as(_lib.u32p), info.nbricks, blist.ctypes.data_as(_lib.u32p),
            blist.size, PLAN_SHAPES[shape], int(kcol))
    check(lib.bk_plan_build_host(*args, None, None, None, None, C.byref(nu), C.byref(nc)))
    anchor = np.zeros(nu.value, np.uint32)
    ln = np.zeros(nu.value, np.uint32)
    mask = np.zeros(nu.value * kcol, np.uint32)
    flat = np.zeros(nc.value, np.uint32)
    check(lib.bk_plan_build_host(*args, anchor.ctypes.data_as(_lib.u32p),
                                 ln.ctypes.data_as(_lib.u32p), mask.ctypes.data_as(_lib.u32p),
                                 flat.ctypes.data_as(_lib.u32p), C.byref(nu), C.byref(nc)))
    return {"anchor": anchor, "len": ln, "mask": mask.reshape(-1, kcol), "flat": flat}


class Plan:
    """An execution plan in HBM (bk_plan_create): the listed bricks grouped
    into k-columns ("column") or 2x2 column quads ("quad") of up to kcol
    levels."""

    def __init__(self, info: BrickInfo, blist: np.ndarray, shape: str = "column",
                 kcol: int = 16):
        blist = np.ascontiguousarray(blist, np.uint32)
        h = C.c_void_p()
        check(load().bk_plan_create(info.adj.ctypes.data_as(_lib.u32p), info.nbricks,
                                    blist.ctypes.data_as(_lib.u32p), blist.size,
                                    PLAN_SHAPES[shape], int(kcol), C.byref(h)))
        self.handle = h.value
        self.shape = shape
        self.kcol = int(kcol)
        nu, nc = C.c_uint32(), C.c_uint32()
        check(load().bk_plan_query(_vp(self.handle), C.byref(nu), None, None, C.byref(nc)))
        self.nunits = nu.value
        self.ncovered = nc.value

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().bk_plan_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# periodic brick grid with one ghost layer
# ---------------------------------------------------------------------------
# Default schedule per (stencil, dtype): (plan shape, levels), None = one
# brick per wave.  From warmed timings of every schedule at 512^3 on one
# MI355X (ms per step with the refresh fused into the sweep / with a separate
# refresh kernel / kernel alone):
#   stars and cond: 2x2 column quads, whose four waves share one LDS tile
#   of the block, so each wave fetches two of its four in-plane faces --
#   FP64 7-point quad1 0.420 / 0.441 / 0.396 against one brick per wave
#   0.466 / 0.478 / 0.436; 13-point quad1 0.440 / 0.457 against 0.514 /
#   0.516; cond quad1 0.422 / 0.441 against 0.466 / 0.481; 25-point quad2
#   0.505 / 0.510 against k-marching columns of 8 0.598 (separate); FP32
#   7-point quad2 0.235 / 0.241 against 0.257 / 0.269, 13-point and cond
#   quad2, 25-point quad4 0.287 / 0.289 against 0.390 / 0.366.  One level
#   (no k-march) wins for FP64 radius 1-2: four times the units of quad4,
#   and the k-faces a march would save are a quarter of what the shared
#   tile saves.  BF16 (the row-form kernels: 16-byte rows): 7/13-point and
#   cond in 2x2 blocks of 4 levels (kernel alone / with a separate refresh:
#   7-point 0.117 / 0.129 against one brick per wave 0.131 / 0.143 and one
#   level 0.144 / 0.158; 13-point 0.131 / 0.144 against 0.139 / 0.151; cond
#   0.131 / 0.145 against 0.146 / 0.158), the 25-point in blocks of 8
#   levels (0.177-0.180 / 0.183-0.187 against one brick per wave 0.184-0.186
#   / 0.186-0.187).
#   cube: quads of 6 levels (kernel at 512^3: FP64 0.613 ms against 0.628
#   with 4 levels, 0.646 with 2 and 0.635 with 8; FP32 0.514 against 0.520
#   / 0.521 / 0.540).
DEFAULT_PLANS = {
    (0, BK_F64): ("quad", 1), (1, BK_F64): ("quad", 1), (4, BK_F64): ("quad", 1),
    (2, BK_F64): ("quad", 2),
    (0, BK_F32): ("quad", 2), (1, BK_F32): ("quad", 2), (4, BK_F32): ("quad", 2),
    (2, BK_F32): ("quad", 4),
    (0, BK_BF16): ("quad", 4), (1, BK_BF16): ("quad", 4), (4, BK_BF16): ("quad", 4),
    (2, BK_BF16): ("quad", 8),
    (3, BK_F64): ("quad", 6), (3, BK_F32): ("quad", 6), (3, BK_BF16): ("quad", 6),
}
# lane_shift of the quad star kernels per (stencil, dtype); absent = 0 (the
# i-taps from the LDS tile).
DEFAULT_LANE_SHIFT: dict = {}


class PeriodicGrid:
    """n = 8*nb interior cells per axis with periodic boundaries, in one of
    two layouts, all tables in HBM:

    ghost=1 (the reference's): an 8-cell (one-brick) ghost zone -- a
      (nb+2)^3 brick grid numbered by init_grid, its adjacency, the interior
      brick list and the periodic ghost pairs a refresh copies before each
      sweep;
    ghost=0: an nb^3 brick grid whose adjacency wraps (init_grid with
      periodic=True); every brick is interior and nothing needs refreshing.
      A sweep reads the periodic image itself where the ghost layout reads
      its copy, so both give the same values, bit for bit.

    layout: the brick shape and vector fold of the grid's fields
    (BK_LAYOUT_8x8x8, BK_LAYOUT_4x8x16 or BK_LAYOUT_4x8x16_F88); nb counts
    bricks of that shape per axis (i, j, k), so the interior holds
    nb[0]*BI x nb[1]*BJ x nb[2]*BK cells.  Each cell gets the same value in
    every layout."""

    def __init__(self, nb, stream=None, ghost: int = 1, tile=(32, 16), layout: int = 0):
        self.nb = tuple(int(x) for x in nb)
        if ghost not in (0, 1):
            raise ValueError("ghost must be 0 or 1")
        ext = (C.c_int * 3)()
        check(load().bk_layout_extents(int(layout), ext))
        self.brick_layout = int(layout)
        self.extents = tuple(int(x) for x in ext)  # (BK, BJ, BI)
        self.ghost = int(ghost)
        self.padded = tuple(x + 2 * self.ghost for x in self.nb)
        self.grid, self.info = init_grid(self.padded, periodic=not self.ghost)
        movBrickInfo(self.info, stream)
        if self.ghost:
            n = C.c_uint32()
            check(load().bk_interior_list(longs(self.nb), None, C.byref(n)))
            self.interior = np.empty(n.value, dtype=np.uint32)
            check(load().bk_interior_list(longs(self.nb), self.interior.ctypes.data_as(_lib.u32p),
                                          C.byref(n)))
            check(load().bk_ghost_pairs(longs(self.nb), None, C.byref(n)))
            self.pairs = np.empty(2 * n.value, dtype=np.uint32)
            check(load().bk_ghost_pairs(longs(self.nb), self.pairs.ctypes.data_as(_lib.u32p),
                                        C.byref(n)))
        else:
            self.interior = np.ascontiguousarray(self.grid.reshape(-1), dtype=np.uint32)
            self.pairs = np.empty(0, dtype=np.uint32)
        self.d_interior = DeviceBuffer.from_array(self.interior, stream)
        self.d_pairs = DeviceBuffer.from_array(self.pairs, stream) if self.pairs.size else None
        # the order the sweeps visit the interior bricks in: grid order (k
        # slowest) with tile=0, or the (i, j) plane cut into Ti x Tj brick
        # tiles (tile=T: T x T), each swept through all k before the next --
        # neighbours in k are then Ti*Tj list positions apart instead of a
        # whole brick plane (T = 32 measured 4-6% faster for the 7-point star
        # at 512^3 than grid order; for the 2x2 block schedules, whose plans
        # follow this order, 32 x 16 is 0-3% ahead of 32 x 32 and 6-10% ahead
        # of whole planes or 8 x 8; 64 x 8, the same 512 bricks per tile
        # plane, ran the radius-1/2 star kernels 0.5-1.3% faster alone but
        # the fused 7-point step 1.9% slower, and 16 x 16, 16 x 8, 32 x 8 and
        # 64 x 4 were 2-8% slower).  Every brick is computed the same way in
        # any order.
        self.tile = (tuple(int(x) for x in tile) if isinstance(tile, (tuple, list))
                     else ((int(tile), int(tile)) if tile else 0))
        self.sweep_list = self.tiled_order(self.tile) if self.tile else self.interior
        self.d_sweep_list = (DeviceBuffer.from_array(self.sweep_list, stream) if self.tile
                             else self.d_interior)
        self.d_grid = DeviceBuffer.from_array(self.grid.reshape(-1), stream)

        self._plans: dict = {}
        self._tuned: dict = {}

    def tiled_order(self, T) -> np.ndarray:
        """Interior bricks tile by tile (bk_tiled_list): Ti x Tj bricks in
        (i, j) (T an int for square tiles or a pair (Ti, Tj)), all k."""
        Ti, Tj = (int(T[0]), int(T[1])) if isinstance(T, (tuple, list)) else (int(T), int(T))
        if Ti < 1 or Tj < 1:
            raise ValueError("tile sides must be positive")
        n = C.c_uint32()
        check(load().bk_tiled_list(longs(self.nb), self.ghost, Ti, Tj, None, C.byref(n)))
        out = np.empty(n.value, dtype=np.uint32)
        check(load().bk_tiled_list(longs(self.nb), self.ghost, Ti, Tj,
                                   out.ctypes.data_as(_lib.u32p), C.byref(n)))
        return out

    def plan(self, shape: str = "column", kcol: int = 16, every: bool = False) -> Plan:
        """A cached plan over the interior bricks in sweep order (every=True:
        over every brick of the padded grid, ghost bricks included)."""
        key = (shape, int(kcol), bool(every))
        if self.brick_layout != 0:
            raise ValueError("plans are built for the 8x8x8 layout only")
        if key not in self._plans:
            blist = np.arange(self.nbricks, dtype=np.uint32) if every else self.sweep_list
            self._plans[key] = Plan(self.info, blist, shape, kcol)
        return self._plans[key]

    @property
    def ghost_images(self):
        """(d_img_off, d_img): the ghost images of every brick as CSR in HBM
        (bk_ghost_images), for the fused refresh; built on first use."""
        if getattr(self, "_images", None) is None:
            n = C.c_uint32()
            check(load().bk_ghost_images(longs(self.nb), None, None, C.byref(n)))
            off = np.zeros(self.nbricks + 1, np.uint32)
            img = np.zeros(max(1, n.value), np.uint32)
            check(load().bk_ghost_images(longs(self.nb), off.ctypes.data_as(_lib.u32p),
                                         img.ctypes.data_as(_lib.u32p), C.byref(n)))
            self._images = (DeviceBuffer.from_array(off), DeviceBuffer.from_array(img))
        return self._images

    @property
    def d_every(self) -> DeviceBuffer:
        """Every brick of the padded grid in grid order, in HBM (the sweep
        list of brick_stencil(..., skip = 0), weak/main.cpp:26-36)."""
        if getattr(self, "_d_every", None) is None:
            self._d_every = DeviceBuffer.from_array(np.arange(self.nbricks, dtype=np.uint32))
        return self._d_every

    def auto_plan(self, stencil_kind: int, dtype: int = BK_F64):
        """Default schedule per stencil and dtype: the tuned one when tune()
        ran for the stencil on this grid, else DEFAULT_PLANS (measured on
        MI355X at 512^3).  The 4x8x16 layouts always run one brick per
        wave."""
        tuned = self._tuned.get(int(stencil_kind))
        if tuned is not None:
            return None if tuned.plan_shape == 0 else self.plan(
                "quad" if tuned.plan_shape == _lib.BK_PLAN_QUAD else "column", tuned.kcol)
        if self.brick_layout != 0:
            return None
        choice = DEFAULT_PLANS.get((int(stencil_kind), int(dtype)))
        if choice is None:
            return None
        if int(stencil_kind) == _lib.BK_CUBE125 and self.interior.size > (1 << 20):
            # large grids: 8 levels (1024^3 FP64: 4.76 ms per sweep on one
            # host against 5.0-6.2 with 4; on most hosts every schedule lands
            # at 5.8-6.1 ms -- 4, 5, 6, 8 and 12 levels within 1.5 % of each
            # other on one -- and 8 levels were never the slower)
            choice = ("quad", 8)
        return self.plan(*choice)

    def auto_wg_bricks(self, stencil_kind: int) -> int:
        """Bricks per workgroup of the per-brick kernels: the tuned value,
        else the kernels' default (0)."""
        tuned = self._tuned.get(int(stencil_kind))
        return tuned.wg_bricks if tuned is not None else 0

    def auto_lane_shift(self, stencil_kind: int, dtype: int = BK_F64) -> int:
        """lane_shift of the quad star kernels: the tuned value, else
        DEFAULT_LANE_SHIFT."""
        tuned = self._tuned.get(int(stencil_kind))
        if tuned is not None:
            return int(tuned.lane_shift)
        return DEFAULT_LANE_SHIFT.get((int(stencil_kind), int(dtype)), 0)

    @property
    def cell_extents(self) -> tuple:
        """Interior cells per axis (i, j, k)."""
        BK, BJ, BI = self.extents
        return (self.nb[0] * BI, self.nb[1] * BJ, self.nb[2] * BK)

    def tune(self, st, dtype: int = BK_F64, steps: int = 5, stream=None, rounds: int = 3,
             warm_s: float = 0.3) -> dict:
        """Time every schedule of this grid's layout for stencil `st`
        (bk_tune_sweep on a grid of the same cells and ghost layer: per
        brick with 1/2/4/8 bricks per workgroup, k-columns or column quads)
        and make the fastest this grid's choice for plan="auto".  All
        schedules compute the same bits.  Returns {"best": candidate,
        "ms": {candidate: ms per sweep}} with candidates named by
        cand_name()."""
        res = tune(st, self.cell_extents, dtype, ghost=self.ghost, steps=steps, rounds=rounds,
                   warm_s=warm_s, layouts=(self.brick_layout,), stream=stream)
        self._tuned[int(st.kind)] = res[0]
        return {"best": cand_name(res[0]), "ms": {cand_name(c): round(c.ms, 4) for c in res}}

    @property
    def nbricks(self) -> int:
        return self.info.nbricks

    @property
    def npairs(self) -> int:
        return self.pairs.size // 2

    @property
    def cells(self) -> int:
        return 512 * self.interior.size

    def field_storage(self, dtype: int = BK_F64, fields: int = 1) -> BrickStorage:
        """A storage of ``fields`` interleaved fields per chunk (step
        512*fields), like stencils/3axis.cpp:26-28 when fields == 2."""
        return BrickStorage.allocate(self.nbricks, BRICK_ELEMS * fields, dtype)

    def fill_uniform(self, brick: Brick, seed: int, lo: float = -1.0, hi: float = 1.0,
                     stream=None) -> None:
        check(load().bk_fill_uniform_layout(self.brick_layout, brick.dtype, _vp(brick.ptr),
                                            C.c_size_t(brick.step), longs(self.nb), self.ghost,
                                            C.c_uint64(seed), lo, hi, _vp(_stream_of(stream))))

    def ghost_refresh(self, brick: Brick, stream=None) -> None:
        if not self.npairs:
            return
        check(load().bk_ghost_refresh(brick.dtype, _vp(brick.ptr), C.c_size_t(brick.step),
                                      C.c_size_t(BRICK_ELEMS), _vp(self.d_pairs.ptr),
                                      C.c_uint32(self.npairs), _vp(_stream_of(stream))))

    def stats(self, brick: Brick, stream=None) -> np.ndarray:
        out = np.zeros(3)
        check(load().bk_field_stats(brick.dtype, _vp(brick.ptr), C.c_size_t(brick.step),
                                    _vp(self.d_interior.ptr), C.c_uint32(self.interior.size),
                                    C.c_size_t(BRICK_ELEMS), out.ctypes.data_as(_lib.dblp),
                                    _vp(_stream_of(stream))))
        return out

    def compare(self, a: Brick, ref: Brick, floor: float = 1.0, atol: float = 0.0,
                rtol: float = 0.0, stream=None) -> dict:
        out = np.zeros(5)
        check(load().bk_field_compare(a.dtype, _vp(a.ptr), C.c_size_t(a.step), ref.dtype,
                                      _vp(ref.ptr), C.c_size_t(ref.step),
                                      _vp(self.d_interior.ptr), C.c_uint32(self.interior.size),
                                      C.c_size_t(BRICK_ELEMS), floor, atol, rtol,
                                      out.ctypes.data_as(_lib.dblp), _vp(_stream_of(stream))))
        n = self.cells
        return {"max_abs": out[0], "max_rel": out[1],
                "rms_rel": float(np.sqrt(out[2] / out[3])) if out[3] > 0 else 0.0,
                "rms_abs": float(np.sqrt(out[2] / n)), "bad": int(out[4])}


def _resolve_plan(plan, st: Stencil, grid: PeriodicGrid, dtype: int = BK_F64):
    if plan is None:
        return None
    if isinstance(plan, str):
        if plan == "auto":
            return grid.auto_plan(st.kind, dtype)
        return grid.plan(plan)
    return plan


def _desc(st: Stencil, out: Brick, inp: Brick, grid: PeriodicGrid, coeff_keep: list,
          blist: DeviceBuffer | None = None, nlist: int | None = None,
          plan=None, wg_bricks=None, lane_shift=None) -> bk_sweep_desc:
    if out.dtype != inp.dtype:
        raise ValueError("input and output fields must share a dtype")
    coeff = np.ascontiguousarray(st.coeff, dtype=np.float64)
    if coeff.size != _lib.load().bk_stencil_ncoeff(st.kind):
        raise ValueError(f"stencil {st.kind} takes {_lib.load().bk_stencil_ncoeff(st.kind)} "
                         f"coefficients, got {coeff.size}")
    coeff_keep.append(coeff)
    d = bk_sweep_desc()
    d.stencil = st.kind
    d.dtype = inp.dtype
    d.d_adj = grid.info.dptr
    d.d_list = (blist or grid.d_sweep_list).ptr
    d.nlist = grid.interior.size if nlist is None else nlist
    d.d_in = inp.ptr
    d.in_step = inp.step
    d.d_out = out.ptr
    d.out_step = out.step
    d.h_coeff = coeff.ctypes.data_as(_lib.dblp)
    d.theta = float(st.theta)
    d.cond_mode = int(st.cond_mode)
    pl = _resolve_plan(plan, st, grid, inp.dtype)
    d.plan = pl.handle if pl is not None else None
    coeff_keep.append(pl)
    d.layout = getattr(grid, "brick_layout", 0)
    if wg_bricks is None:
        wg_bricks = (grid.auto_wg_bricks(st.kind)
                     if plan == "auto" and hasattr(grid, "auto_wg_bricks") else 0)
    d.wg_bricks = 0 if pl is not None else int(wg_bricks)
    if lane_shift is None:
        # automatic: only where the kernels have it (quad plans of the stars)
        quad_star = pl is not None and pl.shape == "quad" and st.kind != _lib.BK_CUBE125
        lane_shift = (grid.auto_lane_shift(st.kind, inp.dtype)
                      if quad_star and hasattr(grid, "auto_lane_shift") else 0)
    d.lane_shift = int(lane_shift)
    return d


def sweep(st: Stencil, out: Brick, inp: Brick, grid: PeriodicGrid, stream=None,
          plan="auto", wg_bricks=None, lane_shift=None) -> None:
    """One sweep over the interior bricks: out = stencil(in).  ``plan`` is
    "auto", None (plain brick list, one brick per wave), "column", "quad"
    or a Plan; ``wg_bricks`` the bricks per workgroup of the per-brick
    kernels (None: tuned or default); ``lane_shift`` 1 for the i-taps by
    DPP lane shifts in the quad star kernels (None: tuned or default)."""
    keep: list = []
    d = _desc(st, out, inp, grid, keep, plan=plan, wg_bricks=wg_bricks, lane_shift=lane_shift)
    check(load().bk_sweep(C.byref(d), _vp(_stream_of(stream))))


def sweeps(st: Stencil, a: Brick, b: Brick, grid: PeriodicGrid, nsweeps: int,
           use_graph: bool = False, refresh: bool = True, stream=None, plan="auto",
           ghost_iters: int = 1, fused: bool = False, wg_bricks=None,
           lane_shift=None) -> Brick:
    """``nsweeps`` ping-pong sweeps a->b->a..., each preceded by the periodic
    ghost refresh of its input -- or, with ghost_iters > 1, the refresh once
    per ghost_iters sweeps and the sweeps in between over every brick of the
    padded grid (bk_sweeps_deep), or, with fused=True, the refresh folded
    into the sweeps' stores (bk_sweeps_fused; stars and cond, one brick per
    wave or by quad plans).  Same result in every form.  Returns the field
    holding the result."""
    keep: list = []
    if ghost_iters != 1 and lane_shift is None:
        lane_shift = 0  # the sweeps over every brick run without quad plans
    d = _desc(st, b, a, grid, keep, plan=plan, wg_bricks=wg_bricks, lane_shift=lane_shift)
    pairs = grid.d_pairs.ptr if (refresh and grid.d_pairs is not None) else None
    npairs = grid.npairs if refresh else 0
    if fused:
        if ghost_iters != 1:
            raise ValueError("the fused refresh refreshes every sweep (ghost_iters = 1)")
        io, im = grid.ghost_images
        check(load().bk_sweeps_fused(C.byref(d), _vp(io.ptr), _vp(im.ptr), _vp(pairs),
                                     C.c_uint32(npairs), int(nsweeps), int(bool(use_graph)),
                                     _vp(_stream_of(stream))))
    elif ghost_iters == 1:
        check(load().bk_sweeps(C.byref(d), _vp(pairs), C.c_uint32(npairs), int(nsweeps),
                               int(bool(use_graph)), _vp(_stream_of(stream))))
    else:
        # sweeps over every brick: column plans as for the interior; the
        # cube per brick (quad plans cannot take the outer layer of a
        # non-periodic grid, bk_plan_create rejects them)
        ip = _resolve_plan(plan, st, grid, a.dtype)
        ap = (grid.plan(ip.shape, ip.kcol, every=True)
              if ip is not None and ip.shape != "quad" else None)
        keep.append(ap)
        check(load().bk_sweeps_deep(C.byref(d), _vp(grid.d_every.ptr), C.c_uint32(grid.nbricks),
                                    _vp(ap.handle if ap is not None else None), _vp(pairs),
                                    C.c_uint32(npairs), int(nsweeps), int(ghost_iters),
                                    int(bool(use_graph)), _vp(_stream_of(stream))))
    return b if nsweeps % 2 == 1 else a


def cand_name(c) -> str:
    """A tuner candidate as text: layout, then the schedule."""
    lay = {0: "8x8x8", 1: "4x8x16", 2: "4x8x16/f88"}[c.layout]
    if c.plan_shape == 0:
        return f"{lay} brick wg{c.wg_bricks}"
    shape = "quad" if c.plan_shape == _lib.BK_PLAN_QUAD else "column"
    return f"{lay} {shape}{c.kcol}" + ("+dpp" if c.lane_shift else "")


def tune(st: Stencil, cells, dtype: int = BK_F64, ghost: int = 1, steps: int = 5,
         rounds: int = 3, warm_s: float = 0.3, layouts=(0, 1, 2), batch_only: bool = False,
         stream=None) -> list:
    """bk_tune_sweep: time every (layout, schedule) candidate for stencil
    `st` on a periodic grid of `cells` (i, j, k) and return them fastest
    first (bk_tune_cand records: layout, plan_shape, kcol, wg_bricks,
    lane_shift, ms)."""
    flags = 0
    for lay in layouts:
        flags |= 1 << int(lay)
    if batch_only:
        flags |= _lib.BK_TUNE_BATCH_SCHEDULES
    coeff = np.ascontiguousarray(st.coeff, dtype=np.float64)
    n = C.c_uint32()
    args = (int(st.kind), int(dtype), coeff.ctypes.data_as(_lib.dblp), float(st.theta),
            int(st.cond_mode), longs(cells), int(ghost), int(steps), int(rounds), float(warm_s),
            C.c_uint(flags))
    check(load().bk_tune_sweep(*args, None, 0, C.byref(n), None))
    out = (_lib.bk_tune_cand * n.value)()
    check(load().bk_tune_sweep(*args, out, n.value, C.byref(n), _vp(_stream_of(stream))))
    return list(out[:n.value])


__all__ = [
    "tune", "cand_name", "DeviceBuffer", "Stream", "Event", "BrickInfo", "init_grid", "movBrickInfo",
    "BrickStorage", "movBrickStorage", "Brick", "copyToBrick", "copyFromBrick",
    "PeriodicGrid", "Plan", "plan_host", "sweep", "sweeps", "device_count", "device_sync", "DTYPE_NP",
    "DTYPE_SIZE", "BK_F64", "BK_F32", "BK_BF16",
]
