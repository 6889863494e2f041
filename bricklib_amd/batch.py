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


class Shape:
    """Shared per-shape device tables (adjacency, interior list, ghost
    pairs, quad plan)."""

    def __init__(self, nb, stream=None, kcol: int = 8):
        self.grid = PeriodicGrid(nb, stream=stream)
        self.nb = self.grid.nb
        self.quad = self.grid.plan("quad", kcol)

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

    def get(self, nb) -> Shape:
        key = tuple(int(x) for x in nb)
        if key not in self.shapes:
            self.shapes[key] = Shape(key, self.stream)
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
    # runtime state
    shape: Shape | None = None
    f0: AsyncBuffer | None = None
    f1: AsyncBuffer | None = None
    parity: int = 0
    done: int = 0
    admitted: float = 0.0
    finished: float = 0.0
    _coeff: np.ndarray = field(default=None, repr=False)

    @property
    def cells(self) -> int:
        return int(np.prod(self.nb)) * 512

    @property
    def field_bytes(self) -> int:
        return int(np.prod([x + 2 for x in self.nb])) * 512 * DTYPE_SIZE[self.dtype]

    @property
    def hbm_bytes(self) -> int:
        return 2 * self.field_bytes

    @property
    def remaining(self) -> int:
        return self.steps - self.done

    def allocate(self, shapes: ShapeCache, stream=None) -> None:
        """Fields in HBM (stream-ordered), input filled from the seed."""
        self.shape = shapes.get(self.nb)
        self.f0 = AsyncBuffer(self.field_bytes, stream)
        self.f1 = AsyncBuffer(self.field_bytes, stream)
        check(load().bk_fill_uniform(self.dtype, _vp(self.f0.ptr), C.c_size_t(512),
                                     _lib.longs(self.nb), C.c_uint64(self.seed), -1.0, 1.0,
                                     _vp(_stream_of(stream))))
        self.parity = 0
        self.done = 0
        self._coeff = np.ascontiguousarray(self.stencil.coeff, np.float64)

    def release(self) -> None:
        for f in (self.f0, self.f1):
            if f is not None:
                f.free()
        self.f0 = self.f1 = None

    @property
    def result_ptr(self) -> int:
        """The field holding the latest result."""
        return (self.f1 if self.parity else self.f0).ptr

    def result_host(self, stream=None) -> np.ndarray:
        from .bricks import DTYPE_NP
        out = np.empty(self.field_bytes // DTYPE_SIZE[self.dtype], DTYPE_NP[self.dtype])
        check(load().bk_memcpy_d2h(out.ctypes.data_as(C.c_void_p), _vp(self.result_ptr),
                                   C.c_size_t(out.nbytes), _vp(_stream_of(stream))))
        return out.reshape(-1, 512)

    def req_struct(self, ticks: int) -> bk_batch_req:
        g = self.shape.grid
        r = bk_batch_req()
        r.stencil = self.stencil.kind
        r.dtype = self.dtype
        r.d_adj = g.info.dptr
        r.d_list = g.d_interior.ptr
        r.nlist = g.interior.size
        r.plan = self.shape.quad.handle if self.stencil.kind == BK_CUBE125 else None
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
        """`ticks` ticks: each request sweeps min(ticks, remaining) times."""
        if len(reqs) > self.max_reqs:
            raise ValueError("more requests than the batch holds")
        arr = (bk_batch_req * max(1, len(reqs)))()
        takes = []
        for q, r in enumerate(reqs):
            arr[q] = r.req_struct(ticks)
            takes.append(arr[q].ticks)
        st = _vp(_stream_of(self.stream))
        check(load().bk_batch_set(_vp(self.handle), arr, len(reqs), st))
        if reqs:
            check(load().bk_batch_run(_vp(self.handle), int(ticks), int(bool(use_graph)), st))
        for r, t in zip(reqs, takes):
            r.done += t
            r.parity ^= t & 1

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().bk_batch_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass
