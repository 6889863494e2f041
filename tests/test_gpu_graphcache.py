"""Graph caches under pressure: more distinct captured runs than a cache
holds (64), launched back to back without waiting, so the caches must evict
graphs whose launches are still queued or running; and a freed plan whose
host handle may be reused by a new plan.  Every result is checked against
the eager (uncaptured) run of the same work, bit for bit, and the single-grid
case against the oracle."""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_F32, BK_F64, stencils as S
from bricklib_amd.batch import Arena, Batch, Request, ShapeCache

pytestmark = pytest.mark.gpu


def _requests(shapes, stream, arena, base_rid=0):
    specs = [((2, 2, 2), "7pt", BK_F64, 11), ((2, 3, 2), "125pt", BK_F32, 12),
             ((3, 2, 2), "13pt", BK_F32, 13), ((2, 2, 3), "125pt", BK_F64, 14)]
    reqs = []
    for q, (nb, name, dt, seed) in enumerate(specs):
        r = Request(rid=base_rid + q, nb=nb, stencil=S.Stencil.make(name, seed=seed), dtype=dt,
                    steps=10 ** 6, seed=seed)
        r.allocate(shapes, stream, arena)
        reqs.append(r)
    return reqs


def test_batch_graph_eviction_while_queued(gpu):
    """90 runs of 1..90 ticks each (90 distinct graph keys), queued on one
    stream with no host wait in between: the batch's cache of 64 must evict
    graphs that were launched moments before.  Same bits as the eager run
    of the same tick sequence."""
    print("[case] batch graph cache: 90 distinct keys, pipelined", flush=True)
    stream = gpu.Stream()
    shapes = ShapeCache()
    arena = Arena(1 << 26)
    ticks = list(range(1, 91))
    results = []
    for use_graph in (True, False):
        batch = Batch(16, stream=stream)
        reqs = _requests(shapes, stream, arena)
        for t in ticks:
            batch.run(reqs, t, use_graph)
        stream.sync()
        results.append([r.interior_host(stream) for r in reqs])
        for r in reqs:
            assert r.done == sum(ticks)
            r.release()
        del batch
    for g, e in zip(*results):
        np.testing.assert_array_equal(g, e)


def test_sweeps_graph_eviction_while_queued(gpu, orc):
    """bk_sweeps with graphs for 70 distinct sweep counts, queued without
    waiting (the process-wide cache holds 64), then checked against the
    oracle's ping-pong of the same total."""
    print("[case] sweeps graph cache: 70 distinct keys, pipelined", flush=True)
    nb = (2, 2, 2)
    st = S.Stencil.make("7pt", seed=5)
    g = gpu.PeriodicGrid(nb)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    g.fill_uniform(a, 9)
    s = gpu.Stream()
    total = 0
    cur, other = a, b
    for n in range(1, 71):
        res = gpu.sweeps(st, cur, other, g, n, use_graph=True, stream=s)
        if res is other:
            cur, other = other, cur
        total += n
    s.sync()
    og = O.Grid(nb)
    oa, ob = og.storage(), og.storage()
    og.fill_uniform(oa, 9)
    ores = og.sweeps(st.kind, st.coeff, oa, ob, total)
    np.testing.assert_array_equal(cur.storage.to_host()[og.interior], ores[og.interior])


@pytest.mark.parametrize("shape", ["column", "quad"])
def test_freed_plan_never_replayed(gpu, orc, shape):
    """A graph captured with a plan is dropped when the plan is freed, and a
    new plan (which may get the old host handle) is captured afresh: each
    run gives the oracle's bits."""
    print(f"[case] freed plan then new plan ({shape})", flush=True)
    name = "125pt" if shape == "quad" else "25pt"
    st = S.Stencil.make(name, seed=3)
    nb = (4, 4, 4)
    g = gpu.PeriodicGrid(nb)
    og = O.Grid(nb)
    for kcol in (2, 4, 8, 3):
        pl = gpu.Plan(g.info, g.sweep_list, shape, kcol)
        a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
        g.fill_uniform(a, 20 + kcol)
        res = gpu.sweeps(st, a, b, g, 3, use_graph=True, plan=pl)
        got = res.storage.to_host()
        del pl
        oa, ob = og.storage(), og.storage()
        og.fill_uniform(oa, 20 + kcol)
        ores = og.sweeps(st.kind, st.coeff, oa, ob, 3)
        np.testing.assert_array_equal(got[og.interior], ores[og.interior])
