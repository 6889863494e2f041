"""Batched sweeps over many independent grids (bk_batch_* of the C-ABI).

A :class:`Shape` holds what grids of one brick-grid shape share, in HBM:
the adjacency of the padded (nb+2)^3 grid, the interior brick list, the
periodic ghost pairs and a 2x2 column-quad plan.  A :class:`Request` is one
grid with its own two fields, stencil, dtype, coefficients and sweep count.
A :class:`Batch` runs any set of requests together: per tick one persistent
refresh kernel per dtype and one persistent sweep kernel per (stencil,
dtype) over the concatenated bricks of all its requests, captured once into
a hipGraph; between runs the host only rewrites the device tables.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import BK_CUBE125, bk_batch_req, check, load
from .bricks import DTYPE_SIZE, DeviceBuffer, PeriodicGrid, _stream_of, _vp
from .stencils import Stencil


def _record_dtype():
    """A numpy view of bk_batch_req (field offsets from the ctypes layout;
    pointers as uint64), so the per-run fields of many requests can be
    written at once."""
    names, formats, offsets = [], [], []
    for fname, ftype in bk_batch_req._fields_:
        names.append(fname)
        offsets.append(getattr(bk_batch_req, fname).offset)
        if ftype in (C.c_int, C.c_int32):
            formats.append(np.int32)
        elif ftype in (C.c_uint32,):
            formats.append(np.uint32)
        elif ftype in (C.c_double,):
            formats.append(np.float64)
        elif ftype in (C.c_size_t, C.c_uint64):
            formats.append(np.uint64)
        else:  # pointers and opaque handles
            formats.append(np.uint64)
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": C.sizeof(bk_batch_req)})


REQ_DTYPE = _record_dtype()


# (stencil, dtype) pairs whose batched sweeps march k-columns (the column
# schedule) instead of one brick per wave.  None by default.  A batch of 60
# FP64 25-point grids of 32..256 cells alone runs faster by columns (209.7
# GStencil/s against 175.4; FP32 362 against 384), but in the serving mix
# the column units -- up to eight bricks swept in turn by one wave -- make
# ticks with little 25-point work last longer: the saturating run took
# 0.409 s against 0.365 and the Poisson stream's p50 rose 10.5 -> 13.3 ms.
BATCH_COLUMN_STENCILS: frozenset = frozenset()


class Shape:
    """Shared per-shape device tables (adjacency, interior list, ghost
    pairs, quad plan, column plan).  ghost=0: the wrapped-adjacency layout
    (no ghost bricks, no refresh); ghost=1: the reference's one-brick ghost
    zone with a periodic refresh before every sweep."""

    def __init__(self, nb, stream=None, kcol: int = 8, ghost: int = 0):
        self.grid = PeriodicGrid(nb, stream=stream, ghost=ghost)
        self.nb = self.grid.nb
        self.ghost = self.grid.ghost
        self.kcol = int(kcol)
        self.quad = self.grid.plan("quad", 4)

    @property
    def columns(self):
        """k-marching column plan of the interior bricks (built on first use)."""
        return self.grid.plan("column", self.kcol)

    def plan_for(self, stencil_kind: int, dtype: int = _lib.BK_F64):
        """The plan a batched request of this stencil and dtype sweeps by
        (None: its bricks one per wave).  Cube: quads of 4 levels -- a batch
        of small grids has few units per persistent wave, and a tick lasts
        as long as the waves with one unit more: 60 grids of 32..256 cells
        ran 20% (FP64) and 11% (FP32) faster than with 8 levels, and a
        lightly loaded stream's latency is lower too (a grid alone in a tick
        is split into twice as many units).  A device work queue handing
        units out dynamically was tried: 8% more saturated FP64 throughput
        with 8-level quads, but a third more latency at light load (its
        atomics cost every short launch several microseconds)."""
        if stencil_kind == BK_CUBE125:
            return self.grid.plan("quad", 4)
        if (stencil_kind, dtype) in BATCH_COLUMN_STENCILS:
            return self.columns
        return None

    @property
    def nbricks(self) -> int:
        return self.grid.nbricks

    @property
    def cells(self) -> int:
        return self.grid.cells


class ShapeCache:
    def __init__(self, stream=None):
        self.stream = stream
        self.shapes: dict = {}

    def get(self, nb, ghost: int = 0) -> Shape:
        key = (tuple(int(x) for x in nb), int(ghost))
        if key not in self.shapes:
            self.shapes[key] = Shape(key[0], self.stream, ghost=key[1])
        return self.shapes[key]


class AsyncBuffer:
    """Device memory from the stream-ordered allocator."""

    def __init__(self, nbytes: int, stream=None):
        p = C.c_void_p()
        self.stream = stream
        check(load().bk_alloc_async(C.c_size_t(int(nbytes)), C.byref(p), _vp(_stream_of(stream))))
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            load().bk_free_async(_vp(self.ptr), _vp(_stream_of(self.stream)))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Arena:
    """Field arena (bk_arena_*): device blocks carved best-fit out of
    hipMalloc'd slabs and recycled without a driver call.  Blocks must be
    used on one stream (a freed block can be handed out again while queued
    work still reads it)."""

    def __init__(self, slab_bytes: int = 1 << 30, capacity: int = 0):
        h = C.c_void_p()
        check(load().bk_arena_create(C.c_size_t(int(slab_bytes)), C.c_size_t(int(capacity)),
                                     C.byref(h)))
        self.handle = h.value

    def alloc(self, nbytes: int) -> "ArenaBlock":
        return ArenaBlock(self, nbytes)

    def stats(self) -> dict:
        u, r, lf = C.c_size_t(), C.c_size_t(), C.c_size_t()
        n = C.c_uint32()
        check(load().bk_arena_stats(_vp(self.handle), C.byref(u), C.byref(r), C.byref(lf),
                                    C.byref(n)))
        return {"used": u.value, "reserved": r.value, "largest_free": lf.value,
                "live_blocks": n.value}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().bk_arena_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass


class ArenaBlock:
    """One block of an Arena (keeps the arena alive)."""

    def __init__(self, arena: Arena, nbytes: int):
        p = C.c_void_p()
        check(load().bk_arena_alloc(_vp(arena.handle), C.c_size_t(int(nbytes)), C.byref(p)))
        self.arena = arena
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)

    def free(self):
        if self.ptr:
            load().bk_arena_free(_vp(self.arena.handle), _vp(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Request:
    """One independent stencil solve: nb interior bricks per axis, `steps`
    sweeps from the seeded U[-1,1] input."""

    rid: int
    nb: tuple
    stencil: Stencil
    dtype: int
    steps: int
    seed: int
    arrival: float = 0.0
    ghost: int = 0  # field layout: 0 wrapped adjacency, 1 ghost zone + refresh
    # runtime state
    shape: Shape | None = None
    f0: object = None  # AsyncBuffer or ArenaBlock
    f1: object = None
    parity: int = 0
    done: int = 0
    admitted: float = 0.0
    finished: float = 0.0
    schedule: tuple = (0, 0)  # (plan shape, kcol) a server ran it by
    _coeff: np.ndarray = field(default=None, repr=False)
    _rec: bytes = field(default=None, repr=False)
    _server: object = field(default=None, repr=False)  # serve.Server that ran it
    _sid: int = field(default=-1, repr=False)

    @property
    def cells(self) -> int:
        return int(np.prod(self.nb)) * 512

    @property
    def field_bytes(self) -> int:
        return int(np.prod([x + 2 * self.ghost for x in self.nb])) * 512 * DTYPE_SIZE[self.dtype]

    @property
    def hbm_bytes(self) -> int:
        return 2 * self.field_bytes

    @property
    def remaining(self) -> int:
        return self.steps - self.done

    def allocate(self, shapes: ShapeCache, stream=None, arena: Arena | None = None) -> None:
        """Fields in HBM -- from `arena` when given (then all of this
        request's work must run on `stream`), else stream-ordered -- and the
        input filled from the seed."""
        self.shape = shapes.get(self.nb, self.ghost)
        if arena is not None:
            self.f0 = arena.alloc(self.field_bytes)
            self.f1 = arena.alloc(self.field_bytes)
        else:
            self.f0 = AsyncBuffer(self.field_bytes, stream)
            self.f1 = AsyncBuffer(self.field_bytes, stream)
        check(load().bk_fill_uniform_grid(self.dtype, _vp(self.f0.ptr), C.c_size_t(512),
                                          _lib.longs(self.nb), self.ghost, C.c_uint64(self.seed),
                                          -1.0, 1.0, _vp(_stream_of(stream))))
        self.parity = 0
        self.done = 0
        self._coeff = np.ascontiguousarray(self.stencil.coeff, np.float64)
        self._rec = bytes(self.req_struct(0))

    def release(self) -> None:
        if self._server is not None:
            check(load().bk_server_release(_vp(self._server.handle), self._sid))
            return
        for f in (self.f0, self.f1):
            if f is not None:
                f.free()
        self.f0 = self.f1 = None
        self._rec = None

    @property
    def result_ptr(self) -> int:
        """The field holding the latest result."""
        if self._server is not None:
            info = _lib.bk_serve_info()
            check(load().bk_server_info(_vp(self._server.handle), self._sid, C.byref(info)))
            if not info.d_result:
                raise ValueError(f"request {self.rid}: no kept result on the server")
            return info.d_result
        return (self.f1 if self.parity else self.f0).ptr

    def result_host(self, stream=None) -> np.ndarray:
        from .bricks import DTYPE_NP
        out = np.empty(self.field_bytes // DTYPE_SIZE[self.dtype], DTYPE_NP[self.dtype])
        if self._server is not None:
            check(load().bk_server_result(_vp(self._server.handle), self._sid,
                                          out.ctypes.data_as(C.c_void_p),
                                          C.c_size_t(out.nbytes)))
        else:
            check(load().bk_memcpy_d2h(out.ctypes.data_as(C.c_void_p), _vp(self.result_ptr),
                                       C.c_size_t(out.nbytes), _vp(_stream_of(stream))))
        return out.reshape(-1, 512)

    def interior_index(self) -> np.ndarray:
        """Ids of the interior bricks in grid order (k slowest) for this
        request's layout (ghost 1: inside the ghost layer; 0: every brick)."""
        if self.shape is not None:
            return self.shape.grid.interior
        if not self.ghost:
            return np.arange(int(np.prod(self.nb)), dtype=np.uint32)
        n = C.c_uint32()
        check(load().bk_interior_list(_lib.longs(self.nb), None, C.byref(n)))
        out = np.empty(n.value, np.uint32)
        check(load().bk_interior_list(_lib.longs(self.nb), out.ctypes.data_as(_lib.u32p),
                                      C.byref(n)))
        return out

    def interior_host(self, stream=None) -> np.ndarray:
        """The latest result's interior bricks in grid order (k slowest), the
        same array for either layout."""
        return self.result_host(stream)[self.interior_index()]

    def req_struct(self, ticks: int) -> bk_batch_req:
        g = self.shape.grid
        r = bk_batch_req()
        r.stencil = self.stencil.kind
        r.dtype = self.dtype
        r.d_adj = g.info.dptr
        r.d_list = g.d_interior.ptr
        r.nlist = g.interior.size
        pl = self.shape.plan_for(self.stencil.kind, self.dtype)
        r.plan = pl.handle if pl is not None else None
        r.d_pairs = g.d_pairs.ptr if g.d_pairs is not None else None
        r.npairs = g.npairs
        r.d_f0 = self.f0.ptr
        r.d_f1 = self.f1.ptr
        r.step = 512
        r.h_coeff = self._coeff.ctypes.data_as(_lib.dblp)
        r.theta = float(self.stencil.theta)
        r.cond_mode = int(self.stencil.cond_mode)
        r.parity = self.parity
        r.ticks = int(min(ticks, self.remaining))
        return r


class Batch:
    """Runs sets of requests tick by tick (bk_batch_*)."""

    def __init__(self, max_reqs: int = 2048, stream=None):
        h = C.c_void_p()
        check(load().bk_batch_create(int(max_reqs), C.byref(h)))
        self.handle = h.value
        self.max_reqs = int(max_reqs)
        self.stream = stream

    def run(self, reqs: list, ticks: int, use_graph: bool = True) -> None:
        """`ticks` ticks: each request sweeps min(ticks, remaining) times.
        The requests' records are built once at allocation; per run only
        their tick counts and parities are written (vectorised)."""
        n = len(reqs)
        if n > self.max_reqs:
            raise ValueError("more requests than the batch holds")
        st = _vp(_stream_of(self.stream))
        if n == 0:
            check(load().bk_batch_set(_vp(self.handle), None, 0, st))
            return
        recs = np.frombuffer(bytearray(b"".join([r._rec for r in reqs])), dtype=REQ_DTYPE)
        rem = np.fromiter((r.steps - r.done for r in reqs), np.int64, n)
        take = np.minimum(rem, int(ticks)).astype(np.uint32)
        recs["ticks"] = take
        recs["parity"] = np.fromiter((r.parity for r in reqs), np.uint32, n)
        check(load().bk_batch_set(_vp(self.handle), recs.ctypes.data_as(C.POINTER(bk_batch_req)),
                                  n, st))
        check(load().bk_batch_run(_vp(self.handle), int(ticks), int(bool(use_graph)), st))
        for r, t in zip(reqs, take.tolist()):
            r.done += t
            r.parity ^= t & 1

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().bk_batch_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass
