"""Batched sweeps and the serving loop on the GPU.

Each request's result is the same alone and in the mix: bit for bit against
the single-grid path (and, for FP64, the oracle), whatever the mix, chunk
size or admission order; requests that do not fit the HBM budget wait and
still complete; malformed requests are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_BF16, BK_CUBE125, BK_F32, BK_F64, serve
from bricklib_amd.batch import Batch, Request, ShapeCache
from bricklib_amd.stencils import Stencil

pytestmark = pytest.mark.gpu

MIX = [((2, 3, 4), "7pt", BK_F64, 5), ((4, 4, 4), "125pt", BK_F64, 3),
       ((3, 2, 2), "13pt", BK_F32, 7), ((2, 2, 5), "25pt", BK_F64, 2),
       ((5, 3, 1), "cond", BK_F64, 4), ((2, 4, 2), "125pt", BK_F32, 6),
       ((1, 1, 1), "7pt", BK_BF16, 9), ((4, 2, 3), "25pt", BK_F32, 1)]


def make_mix():
    reqs = []
    for q, (nb, name, dt, steps) in enumerate(MIX):
        reqs.append(Request(rid=q, nb=nb, stencil=Stencil.make(name, seed=100 + q), dtype=dt,
                            steps=steps, seed=1000 + q))
    return reqs


def alone(gpu, r: Request, ghost: int = 1) -> np.ndarray:
    """The same solve through the single-grid path (ghost zone + refresh by
    default): its interior bricks in grid order."""
    g = gpu.PeriodicGrid(r.nb, ghost=ghost)
    a = gpu.Brick(g.info, g.field_storage(r.dtype), 0)
    b = gpu.Brick(g.info, g.field_storage(r.dtype), 0)
    g.fill_uniform(a, r.seed)
    res = gpu.sweeps(r.stencil, a, b, g, r.steps)
    return res.storage.to_host()[g.interior]


@pytest.mark.parametrize("chunk,ghost,lanes,pipeline", [
    (1, 0, 2, True), (3, 0, 2, True), (8, 0, 2, True), (3, 1, 2, True), (8, 1, 2, True),
    (3, 0, 1, True), (8, 1, 1, True), (3, 0, 2, False), (8, 1, 1, False)])
def test_mix_matches_alone(gpu, chunk, ghost, lanes, pipeline):
    """Every request of a mixed batch -- on one stream, or split over two
    lanes that run concurrently, with the next chunk queued before the
    current one is waited for or not -- gives its single-grid result bit
    for bit."""
    srv = serve.Server(chunk=chunk, ghost=ghost, lanes=lanes, pipeline=pipeline)
    reqs = make_mix()
    if lanes == 2:
        assert {srv.lane(r) for r in reqs} == {0, 1}
    stats = srv.serve(reqs, realtime=False, keep_results=True)
    assert len(stats.completed) == len(reqs)
    for r in reqs:
        print(f"[case] rid={r.rid} {r.nb} kind={r.stencil.kind} dtype={r.dtype} "
              f"steps={r.steps} ghost={ghost} lanes={lanes} pipeline={pipeline}", flush=True)
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "125pt", "cond"])
def test_wrapped_single_grid_matches_ghost(gpu, name):
    """The wrapped-adjacency layout (no ghost bricks, no refresh) gives the
    ghost-zone layout's values bit for bit, with every schedule."""
    nb = (3, 2, 5)
    r = Request(rid=0, nb=nb, stencil=Stencil.make(name, seed=9), dtype=BK_F64, steps=4, seed=11)
    ref = alone(gpu, r, ghost=1)
    np.testing.assert_array_equal(alone(gpu, r, ghost=0), ref)
    g = gpu.PeriodicGrid(nb, ghost=0)
    assert g.npairs == 0 and g.nbricks == 30
    for shape, kcol in (("column", 2), ("column", 8), ("quad", 2), ("quad", 8)):
        if name == "125pt" and shape != "quad":
            continue
        a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        g.fill_uniform(a, r.seed)
        res = gpu.sweeps(r.stencil, a, b, g, r.steps, plan=g.plan(shape, kcol))
        np.testing.assert_array_equal(res.storage.to_host()[g.interior], ref)


def test_mix_fp64_matches_oracle(gpu):
    srv = serve.Server(chunk=4)
    reqs = [r for r in make_mix() if r.dtype == BK_F64]
    srv.serve(reqs, realtime=False, keep_results=True)
    for r in reqs:
        og = O.Grid(r.nb)
        oa, ob = og.storage(), og.storage()
        og.fill_uniform(oa, r.seed)
        ores = og.sweeps(r.stencil.kind, r.stencil.coeff, oa, ob, r.steps,
                         theta=r.stencil.theta, cond_mode=r.stencil.cond_mode)
        np.testing.assert_array_equal(r.interior_host(), ores[og.interior])


def test_tight_budget_queues_and_completes(gpu):
    reqs = make_mix()
    budget = max(r.hbm_bytes for r in reqs) * 2  # about two requests at a time
    srv = serve.Server(hbm_budget=budget, chunk=2)
    stats = srv.serve(reqs, realtime=False, keep_results=True)
    assert len(stats.completed) == len(reqs)
    assert stats.peak_bytes <= budget
    for r in reqs[:4]:
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))


def test_poisson_stream_small(gpu):
    reqs = serve.workload(24, rate=400.0, edges=(16, 48), max_steps=12, seed=5)
    srv = serve.Server(chunk=4)
    stats = srv.serve(reqs, realtime=True)
    s = stats.summary()
    print(s)
    assert s["requests"] == 24 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


@pytest.mark.parametrize("ghost", [0, 1])
def test_star_schedules_in_one_batch(gpu, ghost):
    """Star requests by all three batched schedules in one batch -- one
    brick per wave, k-marching column units (a BK_PLAN_COLUMN plan in the
    request) and 2x2 brick blocks (a BK_PLAN_QUAD plan, one to three levels)
    -- each give the single-grid result bit for bit, for every star."""
    shapes = ShapeCache()
    reqs, q = [], 0
    for name in ("7pt", "13pt", "25pt", "cond"):
        for dt in (BK_F64, BK_F32, BK_BF16):
            for sched in ("brick", "columns", "quad1", "quad3"):
                nb = ((2, 3, 4), (3, 1, 5), (1, 2, 2), (4, 2, 3))[q % 4]
                r = Request(rid=q, nb=nb, stencil=Stencil.make(name, seed=200 + q), dtype=dt,
                            steps=3 + q % 2, seed=300 + q, ghost=ghost)
                r.allocate(shapes)
                rq = r.req_struct(0)
                rq.plan = {"brick": None, "columns": r.shape.columns.handle,
                           "quad1": r.shape.grid.plan("quad", 1).handle,
                           "quad3": r.shape.grid.plan("quad", 3).handle}[sched]
                r._rec = bytes(rq)
                reqs.append(r)
                q += 1
    b = Batch(len(reqs))
    b.run(reqs, 2)
    b.run(reqs, 5)
    for r in reqs:
        assert r.done == r.steps
        print(f"[case] rid={r.rid} {r.nb} kind={r.stencil.kind} dtype={r.dtype} "
              f"steps={r.steps} ghost={ghost}", flush=True)
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))


def test_malformed_requests_rejected(gpu):
    from bricklib_amd import BrickError
    shapes = ShapeCache()
    r = Request(rid=0, nb=(2, 2, 2), stencil=Stencil.make("125pt"), dtype=BK_F64, steps=1,
                seed=1)
    r.allocate(shapes)
    b = Batch(4)
    rq = r.req_struct(1)
    rq.plan = None  # a cube request without its quad plan
    from bricklib_amd._lib import bk_batch_req, load
    arr = (bk_batch_req * 1)(rq)
    assert load().bk_batch_set(C.c_void_p(b.handle), arr, 1, None) == 1
    rq = r.req_struct(1)
    rq.d_f1 = rq.d_f0  # aliased fields
    arr = (bk_batch_req * 1)(rq)
    assert load().bk_batch_set(C.c_void_p(b.handle), arr, 1, None) == 1
    # a star request takes a column or a quad plan
    s = Request(rid=1, nb=(2, 2, 2), stencil=Stencil.make("13pt"), dtype=BK_F64, steps=1,
                seed=2)
    s.allocate(shapes)
    rq = s.req_struct(1)
    for pl in (s.shape.quad, s.shape.columns):
        rq.plan = pl.handle
        arr = (bk_batch_req * 1)(rq)
        assert load().bk_batch_set(C.c_void_p(b.handle), arr, 1, None) == 0
    # a cube request needs a quad plan, not a column plan
    rq = r.req_struct(1)
    rq.plan = s.shape.columns.handle
    arr = (bk_batch_req * 1)(rq)
    assert load().bk_batch_set(C.c_void_p(b.handle), arr, 1, None) == 1
    with pytest.raises(BrickError):
        Batch(0)
    with pytest.raises(ValueError):
        serve.Server(hbm_budget=1000).serve([r], realtime=False)


def test_serving_stream_at_workload_parameters(gpu, orc):
    """The serving workload's own parameters: Poisson arrivals at 200 req/s,
    each grid edge uniform over the multiples of 8 in [32, 256], stencil
    uniform over {7, 13, 25, 125}-point, FP64 or FP32 with probability 1/2,
    1..200 sweeps; two lanes, pipelined chunks.  A sample of the requests
    (the first ones, the largest, both dtypes, every stencil) gives its
    single-grid result bit for bit, and the small FP64 ones the oracle's."""
    reqs = serve.workload(64, rate=200.0, seed=11)
    assert {r.stencil.kind for r in reqs} == {0, 1, 2, 3}
    assert {r.dtype for r in reqs} == {BK_F64, BK_F32}
    assert min(min(r.nb) for r in reqs) >= 4 and max(max(r.nb) for r in reqs) <= 32
    by_work = sorted(reqs, key=lambda r: r.cells * r.steps)
    sample = {r.rid for r in reqs[:4]} | {r.rid for r in by_work[-3:]}
    for kind in range(4):
        for dt in (BK_F64, BK_F32):
            hit = [r for r in reqs if r.stencil.kind == kind and r.dtype == dt]
            if hit:
                sample.add(hit[0].rid)
    small64 = [r for r in by_work if r.dtype == BK_F64][:3]
    sample |= {r.rid for r in small64}
    srv = serve.Server()
    stats = srv.serve(reqs, realtime=True, keep_results=sample)
    s = stats.summary()
    print(s)
    assert s["requests"] == len(reqs) and s["p99_latency_s"] >= s["p50_latency_s"] > 0
    for r in reqs:
        if r.rid not in sample:
            continue
        print(f"[case] rid={r.rid} {r.nb} kind={r.stencil.kind} dtype={r.dtype} "
              f"steps={r.steps}", flush=True)
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))
    for r in small64:
        og = O.Grid(r.nb)
        oa, ob = og.storage(), og.storage()
        og.fill_uniform(oa, r.seed)
        ores = og.sweeps(r.stencil.kind, r.stencil.coeff, oa, ob, r.steps,
                         theta=r.stencil.theta, cond_mode=r.stencil.cond_mode)
        np.testing.assert_array_equal(r.interior_host(), ores[og.interior])
    for r in reqs:
        r.release()
