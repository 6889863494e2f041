"""Serving loop: continuous batching of stencil solves on one GPU.

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

import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import BK_CUBE125, BK_F32, BK_F64
from .batch import Arena, Batch, Request, ShapeCache
from .bricks import Event, Stream
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
    completed: list = field(default_factory=list)
    elapsed: float = 0.0
    peak_bytes: int = 0
    reserved_bytes: int = 0  # the arena's slabs at the end
    chunks: int = 0

    def latencies(self) -> np.ndarray:
        return np.array([r.finished - r.arrival for r in self.completed])

    def summary(self) -> dict:
        lat = self.latencies()
        work = sum(r.cells * r.steps for r in self.completed)
        return {"requests": len(self.completed),
                "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
                "mean_latency_s": float(lat.mean()) if lat.size else None,
                "elapsed_s": self.elapsed,
                "gstencil_s": work / self.elapsed / 1e9 if self.elapsed > 0 else 0.0,
                "stencil_updates": work, "peak_hbm_bytes": self.peak_bytes,
                "arena_reserved_bytes": self.reserved_bytes,
                "chunks": self.chunks}


def lane_of(kind: int, dtype: int, lanes: int) -> int:
    """The lane of a request with `lanes` lanes: 0 for FP64 cubes and the
    stars (and cond) of the other dtypes, 1 for the other cubes and the
    FP64 stars; 0 with one lane."""
    if lanes == 1:
        return 0
    return 0 if (kind == BK_CUBE125) == (dtype == BK_F64) else 1


class Server:
    """Continuous-batching loop over `lanes` batches (one stream each)."""

    def __init__(self, hbm_budget: float = 200e9, chunk: int = 16, max_active: int = 2048,
                 use_graph: bool = True, ghost: int = 0, slab_bytes: int = 1 << 30,
                 lanes: int = 2, pipeline: bool = True):
        if lanes not in (1, 2):
            raise ValueError("lanes must be 1 or 2")
        self.streams = [Stream() for _ in range(lanes)]
        self.stream = self.streams[0]
        # request fields come from an arena; a retired request's memory is
        # reused at once, without a driver call: blocks are released only
        # after every lane's stream has finished the chunk
        self.arena = Arena(slab_bytes)
        self.batches = [Batch(max_active, stream=s) for s in self.streams]
        self.batch = self.batches[0]
        self.shapes = ShapeCache()
        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.max_active = int(max_active)
        self.use_graph = use_graph
        # pipeline: queue the next chunk before waiting for the current one
        # (the host's retire/admit/table work then overlaps the GPU; requests
        # that arrive meanwhile join one chunk later).  At 200 req/s: p50
        # 8.75 -> 8.6 ms, p99 47.9 -> 46 ms, saturating 0.362 -> 0.359 s.
        self.pipeline = bool(pipeline)
        # field layout of admitted requests: 0 = wrapped adjacency (no ghost
        # bricks, no refresh), 1 = one-brick ghost zone refreshed every tick
        self.ghost = int(ghost)

    def lane(self, r: Request) -> int:
        return lane_of(r.stencil.kind, r.dtype, len(self.batches))

    def admissible(self, r: Request, used: int, nactive: int) -> bool:
        return nactive < self.max_active and used + r.hbm_bytes <= self.budget

    def prewarm(self, requests: list) -> None:
        """Build the shared per-shape tables of these requests' grid shapes
        ahead of time (a shape catalogue known to the server)."""
        for r in requests:
            self.shapes.get(r.nb, self.ghost)
        self.stream.sync()

    def serve(self, requests: list, realtime: bool = True, keep_results=False,
              on_complete=None) -> ServeStats:
        """Serve `requests` (sorted by arrival).  realtime: arrivals follow
        the wall clock from the start; otherwise all are present at t=0
        (a saturating, closed-loop run).  keep_results: True, or a set of
        request ids whose fields stay allocated after completion."""
        pending = sorted(requests, key=lambda r: r.arrival)
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        for r in pending:
            r.ghost = self.ghost
            if r.hbm_bytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.hbm_bytes} B, above the budget")
        active: list = []
        used = 0
        st = ServeStats()
        t0 = time.perf_counter()
        qi = 0
        inflight = None  # pipelined: (events, finishing requests) of the queued chunk

        def retire(done: list) -> None:
            nonlocal used
            now = time.perf_counter() - t0
            for r in done:
                r.finished = now
                if on_complete is not None:
                    on_complete(r)
                keep = keep_results if isinstance(keep_results, bool) else (
                    r.rid in keep_results)
                if not keep:
                    r.release()
                used -= r.hbm_bytes
                st.completed.append(r)

        while qi < len(pending) or active or inflight is not None:
            now = time.perf_counter() - t0
            # admit arrived requests that fit
            while qi < len(pending) and pending[qi].arrival <= now:
                r = pending[qi]
                if not self.admissible(r, used, len(active)):
                    break
                # filled on its lane's stream, which runs all its ticks
                r.allocate(self.shapes, self.streams[self.lane(r)], self.arena)
                r.admitted = now
                used += r.hbm_bytes
                active.append(r)
                qi += 1
            st.peak_bytes = max(st.peak_bytes, used)
            if not active and inflight is None:
                if qi < len(pending):
                    time.sleep(max(0.0, pending[qi].arrival - (time.perf_counter() - t0)))
                continue
            queued = None
            if active:
                ticks = min(self.chunk, max(r.remaining for r in active))
                lanes = [[] for _ in self.batches]
                for r in active:
                    lanes[self.lane(r)].append(r)
                for b, reqs in zip(self.batches, lanes):
                    if reqs:
                        b.run(reqs, ticks, self.use_graph)
                st.chunks += 1
                done = [r for r in active if r.remaining == 0]
                active = [r for r in active if r.remaining > 0]
                if self.pipeline:
                    events = []
                    for stream in self.streams:
                        e = Event()
                        e.record(stream)
                        events.append(e)
                    queued = (events, done)
                else:
                    for stream in self.streams:
                        stream.sync()
                    retire(done)
            if inflight is not None:
                # the previous chunk, queued before this one: its requests
                # are complete once its events are
                for e in inflight[0]:
                    e.sync()
                retire(inflight[1])
            inflight = queued
        st.elapsed = time.perf_counter() - t0
        st.reserved_bytes = self.arena.stats()["reserved"]
        return st


def stencil_name(kind: int) -> str:
    for k, v in NAMES.items():
        if v == kind:
            return k
    return str(kind)
