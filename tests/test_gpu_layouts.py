"""Brick shapes, vector folds, bricks per workgroup and the autotuner on the
GPU, through the C-ABI, against the oracle:

  - 4x8x16 bricks in the 4x16 fold (row-major 128k + 16j + i) and in the
    8x8 fold (two 8-wide halves per row), every stencil, FP64/FP32/BF16,
    1/2/4 bricks per workgroup, ghost and wrapped layouts: bit for bit the
    oracle's sweep over the same layout, and the same dense field as the
    8x8x8 sweep of the same cells;
  - the 8x8x8 star kernel with 1, 2 and 8 bricks per workgroup;
  - bk_tune_sweep: every candidate it reports is one the library can run,
    the grid's tuned choice gives the same bits as the default schedule.
"""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_BF16, BK_F32, BK_F64, stencils as S

pytestmark = pytest.mark.gpu

NAMES = ["7pt", "13pt", "25pt", "125pt", "cond"]


def gpu_sweeps(gpu, nb, layout, ghost, st, dtype, n, wg, seed=5):
    g = gpu.PeriodicGrid(nb, ghost=ghost, layout=layout)
    a = gpu.Brick(g.info, g.field_storage(dtype), 0)
    b = gpu.Brick(g.info, g.field_storage(dtype), 0)
    a.storage.buf.zero()
    b.storage.buf.zero()
    g.fill_uniform(a, seed)
    res = gpu.sweeps(st, a, b, g, n, wg_bricks=wg, use_graph=(wg == 2))
    return g, res.storage.to_host()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("layout", [1, 2], ids=["4x8x16", "4x8x16f88"])
def test_layout_sweeps_match_oracle(gpu, orc, layout, name):
    st = S.Stencil.make(name, seed=12)
    cells = (48, 24, 16)  # three 16-wide bricks along i: neighbours differ
    BK, BJ, BI = O.layout_extents(layout)
    nb = (cells[0] // BI, cells[1] // BJ, cells[2] // BK)
    dtypes = (BK_F64, BK_F32, BK_BF16) if name in ("7pt", "125pt") else (BK_F64, BK_F32)
    for dt in dtypes:
        og8 = O.Grid((cells[0] // 8, cells[1] // 8, cells[2] // 8))
        a8, b8 = og8.storage(dt), og8.storage(dt)
        og8.fill_uniform(a8, 5, dtype=dt)
        r8 = og8.sweeps(st.kind, st.coeff, a8, b8, 3, dtype=dt, theta=st.theta,
                        cond_mode=st.cond_mode)
        dense8 = O.LayoutGrid(og8.nb, 0).to_dense(r8, dt)
        for ghost in (1, 0):
            lg = O.LayoutGrid(nb, layout, ghost)
            oa, ob = lg.storage(dt), lg.storage(dt)
            lg.fill_uniform(oa, 5, dtype=dt)
            want = lg.sweeps(st.kind, st.coeff, oa, ob, 3, dtype=dt, theta=st.theta,
                             cond_mode=st.cond_mode, refresh=bool(ghost))
            for wg in (1, 2, 4):
                print(f"[case] layout {layout} {name} dt={dt} ghost={ghost} wg={wg}", flush=True)
                g, got = gpu_sweeps(gpu, nb, layout, ghost, st, dt, 3, wg)
                np.testing.assert_array_equal(got[lg.interior], want[lg.interior])
                np.testing.assert_array_equal(lg.to_dense(got, dt), dense8)


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "cond"])
@pytest.mark.parametrize("wg", [1, 2, 8])
def test_888_star_bricks_per_workgroup(gpu, orc, name, wg):
    st = S.Stencil.make(name, seed=13)
    nb = (5, 3, 4)  # 60 interior bricks: not a multiple of 8
    og = O.Grid(nb)
    for dt in (BK_F64, BK_F32, BK_BF16):
        print(f"[case] 8x8x8 {name} wg={wg} dt={dt}", flush=True)
        g, got = gpu_sweeps(gpu, nb, 0, 1, st, dt, 3, wg)
        oa, ob = og.storage(dt), og.storage(dt)
        og.fill_uniform(oa, 5, dtype=dt)
        want = og.sweeps(st.kind, st.coeff, oa, ob, 3, dtype=dt, theta=st.theta,
                         cond_mode=st.cond_mode)
        np.testing.assert_array_equal(got[og.interior], want[og.interior])


def test_layout_rejects_unsupported_forms(gpu):
    from bricklib_amd import BrickError
    g = gpu.PeriodicGrid((2, 3, 4), layout=1)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    st = S.Stencil.make("7pt", seed=1)
    with pytest.raises(BrickError):  # the fused refresh is an 8x8x8 kernel
        gpu.sweeps(st, a, b, g, 2, fused=True)
    with pytest.raises(ValueError):  # plans are 8x8x8 only
        g.plan("column", 4)
    with pytest.raises(BrickError):  # 8 bricks per workgroup: 8x8x8 stars only
        gpu.sweeps(st, a, b, g, 1, wg_bricks=8)
    with pytest.raises(BrickError):  # radius 2 x 3 sweeps exceed a 4-deep ghost zone
        gpu.sweeps(S.Stencil.make("13pt", seed=1), a, b, g, 3, ghost_iters=3)


def test_deep_ghost_on_4x8x16(gpu, orc):
    """Deep-ghost sweeps on 4-deep bricks (radius * ghost_iters <= 4): the
    interior equals the refresh-every-sweep oracle."""
    st = S.Stencil.make("7pt", seed=3)
    nb = (3, 3, 5)
    g = gpu.PeriodicGrid(nb, layout=1)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    a.storage.buf.zero()
    b.storage.buf.zero()
    g.fill_uniform(a, 7)
    got = gpu.sweeps(st, a, b, g, 7, ghost_iters=4).storage.to_host()
    lg = O.LayoutGrid(nb, 1)
    oa, ob = lg.storage(), lg.storage()
    lg.fill_uniform(oa, 7)
    want = lg.sweeps(st.kind, st.coeff, oa, ob, 7)
    np.testing.assert_array_equal(got[lg.interior], want[lg.interior])


@pytest.mark.parametrize("name", ["7pt", "125pt"])
def test_tuner_candidates_and_choice(gpu, orc, name):
    st = S.Stencil.make(name, seed=8)
    cells = (64, 32, 32)
    print(f"[case] tune {name} {cells}", flush=True)
    res = gpu.tune(st, cells, BK_F64, ghost=1, steps=2, rounds=1, warm_s=0.02)
    names = [gpu.cand_name(c) for c in res]
    print("[tune]", [(n, round(c.ms, 4)) for n, c in zip(names, res)])
    assert len(set(names)) == len(names)
    assert {c.layout for c in res} == {0, 1, 2}
    assert all(res[q].ms <= res[q + 1].ms for q in range(len(res) - 1))
    assert all(0.0 < c.ms < 1e3 for c in res)
    if name == "125pt":
        assert {"8x8x8 quad2", "8x8x8 quad4", "8x8x8 quad6", "8x8x8 quad8"} <= set(names)
    else:
        assert {"8x8x8 brick wg1", "8x8x8 brick wg8", "8x8x8 column16"} <= set(names)
    # restricted to the batched kernels' schedules
    rb = gpu.tune(st, cells, BK_F64, ghost=0, steps=2, rounds=1, warm_s=0.0, batch_only=True)
    assert {c.layout for c in rb} == {0}
    assert all(c.plan_shape or c.wg_bricks == 4 for c in rb)
    # a grid's tuned choice sweeps to the oracle's bits
    nb = (cells[0] // 8, cells[1] // 8, cells[2] // 8)
    g = gpu.PeriodicGrid(nb)
    out = g.tune(st, BK_F64, steps=2, rounds=1, warm_s=0.02)
    assert out["best"] in out["ms"]
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    g.fill_uniform(a, 4)
    got = gpu.sweeps(st, a, b, g, 3).storage.to_host()
    og = O.Grid(nb)
    oa, ob = og.storage(), og.storage()
    og.fill_uniform(oa, 4)
    want = og.sweeps(st.kind, st.coeff, oa, ob, 3)
    np.testing.assert_array_equal(got[og.interior], want[og.interior])
