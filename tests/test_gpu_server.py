"""The C++ serving loop (bk_server_*) through its C-ABI: served on its own
thread with requests submitted while it runs, with the autotuner choosing
each request's schedule per size class, and its bookkeeping (request
states, latencies, statistics, kept results).  Every result is checked bit
for bit against the same solve through the single-grid path."""
import ctypes as C

import numpy as np
import pytest

from bricklib_amd import BK_F32, BK_F64, BrickError, _lib, serve
from bricklib_amd.batch import Request
from bricklib_amd.stencils import Stencil

pytestmark = pytest.mark.gpu

MIX = [((2, 3, 4), "7pt", BK_F64, 5), ((4, 4, 4), "125pt", BK_F64, 3),
       ((3, 2, 2), "13pt", BK_F32, 7), ((2, 2, 5), "25pt", BK_F64, 2),
       ((5, 3, 1), "cond", BK_F64, 4), ((2, 4, 2), "125pt", BK_F32, 6)]


def make_mix(base=0):
    return [Request(rid=base + q, nb=nb, stencil=Stencil.make(name, seed=400 + q), dtype=dt,
                    steps=steps, seed=500 + q) for q, (nb, name, dt, steps) in enumerate(MIX)]


def alone(gpu, r):
    g = gpu.PeriodicGrid(r.nb)
    a = gpu.Brick(g.info, g.field_storage(r.dtype), 0)
    b = gpu.Brick(g.info, g.field_storage(r.dtype), 0)
    g.fill_uniform(a, r.seed)
    return gpu.sweeps(r.stencil, a, b, g, r.steps).storage.to_host()[g.interior]


def test_threaded_server_submit_while_running(gpu):
    """bk_server_start, then requests submitted from this thread while the
    server's thread sweeps earlier ones; bk_server_wait on each; results as
    alone; stop returns once all are done."""
    lib = gpu.load()
    srv = serve.Server(chunk=2)
    h = C.c_void_p(srv.handle)
    assert lib.bk_server_start(h) == 0
    reqs = make_mix()
    ids = []
    for r in reqs:
        r.arrival = -1.0  # stamped on submission
        q = srv._spec(r, keep=True)
        sid = C.c_uint32()
        assert lib.bk_server_submit(h, C.byref(q), C.byref(sid)) == 0
        r._server, r._sid, r.ghost = srv, sid.value, srv.ghost
        ids.append(sid.value)
    print("[case] threaded server: submitted", ids, flush=True)
    # another thread polls every request's state and the statistics while
    # the server's thread sweeps (ctypes calls release the GIL): the states
    # and sweep counts it sees only move forward
    import threading
    seen, stop_poll, errors = {}, threading.Event(), []

    def poll():
        inf, st = _lib.bk_serve_info(), _lib.bk_serve_stats()
        while not stop_poll.is_set():
            for sid in ids:
                if lib.bk_server_info(h, sid, C.byref(inf)) != 0:
                    errors.append(sid)
                    continue
                prev = seen.get(sid, (0, 0))
                if inf.state < prev[0] or inf.sweeps_done < prev[1]:
                    errors.append((sid, prev, (inf.state, inf.sweeps_done)))
                seen[sid] = (inf.state, inf.sweeps_done)
            if lib.bk_server_stats(h, C.byref(st)) != 0:
                errors.append("stats")

    poller = threading.Thread(target=poll)
    poller.start()
    try:
        for sid in ids:
            rc = lib.bk_server_wait(h, sid, 60.0)
            assert rc == 0, lib.bk_last_error()
    finally:
        stop_poll.set()
        poller.join()
    assert not errors, errors
    assert lib.bk_server_stop(h) == 0
    info = _lib.bk_serve_info()
    for r in reqs:
        assert lib.bk_server_info(h, r._sid, C.byref(info)) == 0
        assert info.state == 2 and info.sweeps_done == r.steps
        assert info.finished >= info.admitted >= info.arrival >= 0.0
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))
        r.release()
        assert lib.bk_server_info(h, r._sid, C.byref(info)) == 0 and info.state == 3
    st = _lib.bk_serve_stats()
    assert lib.bk_server_stats(h, C.byref(st)) == 0
    assert st.requests == len(reqs) and st.p99_s >= st.p50_s > 0.0
    # a released or unknown request has no result
    out = np.empty(reqs[0].field_bytes // 8)
    assert lib.bk_server_result(h, reqs[0]._sid, out.ctypes.data_as(C.c_void_p),
                                C.c_size_t(out.nbytes)) != 0
    assert lib.bk_server_wait(h, 999, 0.1) != 0


@pytest.mark.parametrize("ghost", [0, 1])
def test_tuned_schedules_give_the_same_bits(gpu, ghost):
    """With tuning on, each (size class, stencil, dtype) sweeps by the
    schedule the tuner timed fastest among the batched kernels' schedules;
    results are those of the single-grid path whatever it chose."""
    srv = serve.Server(chunk=3, ghost=ghost, tune=True)
    reqs = make_mix()
    srv.prewarm(reqs)
    stats = srv.serve(reqs, realtime=False, keep_results=True)
    s = stats.summary()
    print(s, [(r.rid, r.schedule) for r in reqs])
    assert s["tuned_classes"] >= 4 and s["requests"] == len(reqs)
    for r in reqs:
        if r.stencil.kind == 3:
            assert r.schedule[0] == _lib.BK_PLAN_QUAD and r.schedule[1] in (2, 4, 8)
        else:
            assert r.schedule[0] in (0, _lib.BK_PLAN_COLUMN, _lib.BK_PLAN_QUAD)
        print(f"[case] tuned rid={r.rid} schedule={r.schedule} ghost={ghost}", flush=True)
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))


def test_server_rejects_malformed(gpu):
    lib = gpu.load()
    cfg = _lib.bk_serve_config()
    cfg.hbm_budget, cfg.chunk, cfg.max_active, cfg.lanes = 1e9, 16, 16, 3
    h = C.c_void_p()
    assert lib.bk_server_create(C.byref(cfg), C.byref(h)) != 0  # three lanes
    srv = serve.Server(hbm_budget=1e9)
    r = make_mix()[0]
    q = srv._spec(r, keep=False)
    q.stencil = 9
    sid = C.c_uint32()
    assert lib.bk_server_submit(C.c_void_p(srv.handle), C.byref(q), C.byref(sid)) != 0
    q = srv._spec(r, keep=False)
    q.steps = 0
    assert lib.bk_server_submit(C.c_void_p(srv.handle), C.byref(q), C.byref(sid)) != 0
    q = srv._spec(r, keep=False)
    q.nb[0] = q.nb[1] = 4000  # beyond the budget
    assert lib.bk_server_submit(C.c_void_p(srv.handle), C.byref(q), C.byref(sid)) != 0
    with pytest.raises(BrickError):
        serve.Server(lanes=2, chunk=0)


@pytest.mark.parametrize("star_plan", [None, "column", "quad"])
@pytest.mark.parametrize("ghost", [0, 1])
def test_star_plan_choices_give_the_same_bits(gpu, star_plan, ghost):
    """Every fixed star schedule of the serving loop (one brick per wave,
    k-columns of 8, 2x2 brick blocks) gives the single-grid results."""
    srv = serve.Server(chunk=2, ghost=ghost, star_plan=star_plan)
    reqs = make_mix(base=100)
    stats = srv.serve(reqs, realtime=False, keep_results=True)
    assert len(stats.completed) == len(reqs)
    for r in reqs:
        want = {None: 0, "column": _lib.BK_PLAN_COLUMN, "quad": _lib.BK_PLAN_QUAD}[star_plan]
        assert r.schedule[0] == (_lib.BK_PLAN_QUAD if r.stencil.kind == 3 else want)
        if r.stencil.kind == 3:
            # a light load: the cubes ran one-level quads
            assert r.schedule[1] == 1
        print(f"[case] star_plan={star_plan} ghost={ghost} rid={r.rid}", flush=True)
        np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))


def test_heavy_cube_load_runs_two_level_quads(gpu):
    """A cube load far above what the batched kernel holds at once (a 512^3
    FP64 grid: 32768 two-level units, at least four times the resident
    waves of any occupancy) runs two-level quads, with the single-grid
    bits."""
    r = Request(rid=0, nb=(64, 64, 64), stencil=Stencil.make("125pt", seed=47), dtype=BK_F64,
                steps=2, seed=48)
    srv = serve.Server(chunk=2)
    stats = srv.serve([r], realtime=False, keep_results=True)
    assert len(stats.completed) == 1
    print(f"[case] heavy cube schedule={r.schedule}", flush=True)
    assert r.schedule == (_lib.BK_PLAN_QUAD, 2)
    np.testing.assert_array_equal(r.interior_host(), alone(gpu, r))
