"""Serving loop: continuous batching of stencil solves on one GPU.

The loop itself is C++ (bk_server_* of libbricks.so); this module is its
Python handle plus the workload generator of the benchmark.

Requests (grid size, stencil, dtype, sweep count, seed) arrive over time.
The loop admits every arrived request whose fields fit the HBM budget
(fields stay resident until the request completes), runs all admitted
requests together for a chunk of ticks in one batched run, retires the
completed ones (their memory returns to the budget) and admits new arrivals
between chunks -- requests that arrive while others are on the GPU join the
next chunk, so their work overlaps on the device.  Latency of a request =
completion time - arrival time; throughput = interior cells x sweeps of all
completed requests / elapsed time.

The admitted requests are split over `lanes` batches, each on its own
stream, joined once per chunk: lane 0 takes the FP64 cubes and the other
dtypes' stars, lane 1 the other cubes and the FP64 stars, so a lane's
compute-bound cube ticks run beside the other lane's memory-bound star
ticks and fill the tails of each other's persistent kernels (the
saturating mix of the benchmark: 385 -> 357 ms on one MI355X; both cubes
on one lane and the stars on the other gained nothing).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import BK_CUBE125, BK_F32, BK_F64, check, load
from .batch import Request
from .bricks import _vp
from .stencils import NAMES, Stencil, u01


def workload(n: int, rate: float, seed: int = 2024, edges=(32, 256), max_steps: int = 200,
             stencils=("7pt", "13pt", "25pt", "125pt"), dtypes=(BK_F64, BK_F32)) -> list:
    """Poisson arrivals at `rate` per second; each edge uniform over the
    multiples of 8 in [edges[0], edges[1]] (non-cubic), stencil uniform over
    `stencils`, dtype uniform over `dtypes`, steps uniform in [1, max_steps],
    per-request seed.  Deterministic from `seed`."""
    out = []
    t = 0.0
    choices = np.arange(edges[0], edges[1] + 1, 8)
    for r in range(n):
        h = lambda q: u01(seed, 16 * r + q)  # noqa: E731
        t += -np.log(1.0 - h(0)) / rate if rate > 0 else 0.0
        nb = tuple(int(choices[min(int(h(1 + a) * len(choices)), len(choices) - 1)]) // 8
                   for a in range(3))
        sname = stencils[min(int(h(4) * len(stencils)), len(stencils) - 1)]
        dt = dtypes[min(int(h(5) * len(dtypes)), len(dtypes) - 1)]
        steps = 1 + min(int(h(6) * max_steps), max_steps - 1)
        rseed = int(h(7) * (1 << 62))
        out.append(Request(rid=r, nb=nb, stencil=Stencil.make(sname, seed=rseed), dtype=dt,
                           steps=steps, seed=rseed, arrival=t))
    return out


@dataclass
class ServeStats:
    """What a serve() call completed and the server's statistics for it
    (bk_server_stats)."""

    completed: list = field(default_factory=list)
    elapsed: float = 0.0
    peak_bytes: float = 0.0
    reserved_bytes: float = 0.0  # the arena's slabs at the end
    chunks: int = 0
    native: dict = field(default_factory=dict)

    def latencies(self) -> np.ndarray:
        return np.array([r.finished - r.arrival for r in self.completed])

    def summary(self) -> dict:
        n = self.native
        return {"requests": len(self.completed),
                "p50_latency_s": n.get("p50_s") if self.completed else None,
                "p99_latency_s": n.get("p99_s") if self.completed else None,
                "mean_latency_s": n.get("mean_s") if self.completed else None,
                "elapsed_s": self.elapsed,
                "gstencil_s": n.get("gstencil_s", 0.0),
                "stencil_updates": int(n.get("stencil_updates", 0)),
                "peak_hbm_bytes": int(self.peak_bytes),
                "arena_reserved_bytes": int(self.reserved_bytes),
                "chunks": self.chunks, "tuned_classes": n.get("tuned_classes", 0)}


def lane_of(kind: int, dtype: int, lanes: int) -> int:
    """The lane of a request with `lanes` lanes: 0 for FP64 cubes and the
    stars (and cond) of the other dtypes, 1 for the other cubes and the
    FP64 stars; 0 with one lane (the rule of the C++ loop)."""
    if lanes == 1:
        return 0
    return 0 if (kind == BK_CUBE125) == (dtype == BK_F64) else 1


class Server:
    """Python handle of the C++ serving loop (bk_server_*): continuous
    batching over `lanes` batches (one stream each), requests admitted
    against an HBM budget, chunks of `chunk` ticks, the next chunk queued
    before the current one is waited for (`pipeline`), and with `tune` each
    request's schedule chosen by the autotuner per size class; otherwise
    stars sweep by `star_plan` ("quad": 2x2 brick blocks, "column": k-columns
    of 8, None: one brick per wave) and the cube by 2x2 column quads."""

    def __init__(self, hbm_budget: float = 200e9, chunk: int = 16, max_active: int = 2048,
                 use_graph: bool = True, ghost: int = 0, slab_bytes: int = 1 << 30,
                 lanes: int = 2, pipeline: bool = True, tune: bool = False,
                 star_plan: str | None = "quad"):
        if lanes not in (1, 2):
            raise ValueError("lanes must be 1 or 2")
        cfg = _lib.bk_serve_config()
        cfg.hbm_budget = float(hbm_budget)
        cfg.chunk = int(chunk)
        cfg.max_active = int(max_active)
        cfg.lanes = int(lanes)
        cfg.use_graph = int(bool(use_graph))
        cfg.ghost = int(ghost)
        cfg.pipeline = int(bool(pipeline))
        cfg.tune = int(bool(tune))
        cfg.slab_bytes = int(slab_bytes)
        cfg.star_plan = {None: 0, "quad": _lib.BK_PLAN_QUAD,
                         "column": _lib.BK_PLAN_COLUMN}[star_plan]
        h = C.c_void_p()
        check(load().bk_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h.value
        self.budget = float(hbm_budget)
        self.ghost = int(ghost)
        self.lanes = int(lanes)
        self.chunk = int(chunk)

    def lane(self, r: Request) -> int:
        return lane_of(r.stencil.kind, r.dtype, self.lanes)

    def _spec(self, r: Request, keep: bool) -> "_lib.bk_serve_req":
        q = _lib.bk_serve_req()
        for a in range(3):
            q.nb[a] = int(r.nb[a])
        q.stencil = int(r.stencil.kind)
        q.dtype = int(r.dtype)
        q.steps = int(r.steps)
        q.seed = int(r.seed)
        r._coeff = np.ascontiguousarray(r.stencil.coeff, np.float64)
        q.h_coeff = r._coeff.ctypes.data_as(_lib.dblp)
        q.theta = float(r.stencil.theta)
        q.cond_mode = int(r.stencil.cond_mode)
        q.arrival = float(r.arrival)
        q.keep = int(bool(keep))
        return q

    def prewarm(self, requests: list) -> None:
        """Build the shared per-shape tables (and, with tuning, time the
        size classes) of these requests ahead of serving."""
        for r in requests:
            r.ghost = self.ghost
        specs = (_lib.bk_serve_req * len(requests))(*[self._spec(r, False) for r in requests])
        check(load().bk_server_prewarm(_vp(self.handle), specs, len(requests)))

    def serve(self, requests: list, realtime: bool = True, keep_results=False) -> ServeStats:
        """Serve `requests` to completion (bk_server_run).  realtime: each is
        admitted once the clock from the start passes its arrival; otherwise
        all are present at t=0 (a saturating, closed-loop run).
        keep_results: True, or a set of request ids whose result fields stay
        allocated (Request.interior_host / release)."""
        for r in requests:
            r.ghost = self.ghost
            if r.hbm_bytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.hbm_bytes} B, above the budget")
        lib = load()
        check(lib.bk_server_reset_stats(_vp(self.handle)))
        for r in sorted(requests, key=lambda r: r.arrival):
            keep = keep_results if isinstance(keep_results, bool) else r.rid in keep_results
            q = self._spec(r, keep)
            sid = C.c_uint32()
            check(lib.bk_server_submit(_vp(self.handle), C.byref(q), C.byref(sid)))
            r._server, r._sid = self, sid.value
        check(lib.bk_server_run(_vp(self.handle), int(bool(realtime))))
        info = _lib.bk_serve_info()
        for r in requests:
            check(lib.bk_server_info(_vp(self.handle), r._sid, C.byref(info)))
            r.arrival, r.admitted, r.finished = info.arrival, info.admitted, info.finished
            r.done = info.sweeps_done
            r.schedule = (info.plan_shape, info.kcol)
        c = _lib.bk_serve_stats()
        check(lib.bk_server_stats(_vp(self.handle), C.byref(c)))
        native = {f: getattr(c, f) for f, _ in _lib.bk_serve_stats._fields_}
        return ServeStats(completed=sorted(requests, key=lambda r: r.finished),
                          elapsed=c.elapsed_s, peak_bytes=c.peak_hbm_bytes,
                          reserved_bytes=c.arena_reserved_bytes, chunks=int(c.chunks),
                          native=native)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().bk_server_destroy(_vp(self.handle))
                self.handle = None
        except Exception:
            pass


def stencil_name(kind: int) -> str:
    for k, v in NAMES.items():
        if v == kind:
            return k
    return str(kind)
