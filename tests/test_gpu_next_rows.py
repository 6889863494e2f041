"""The rows around the sweep, anchored on the oracle on the GPU:

  deep-ghost sweeps   bk_sweeps_deep against the oracle running the same
                      communication-avoiding loop (weak/main.cpp:26-36,
                      203-211: refresh once per g sweeps, sweeps over every
                      brick -- ghost bricks included, with the reference's
                      non-periodic init_grid adjacency at the outer layer --
                      and the last sweep of each group over the interior),
                      bit for bit over the whole field, ghost bricks too;
                      its interior also equals the refresh-every-sweep loop;
  BrickDecomp layout  sweeps on the BrickDecomp numbering
                      (brick-mpi.h:322-460) against the oracle's brick loop
                      on that numbering's own adjacency and exchange pairs;
  ingested stencils   DSL text (the codegen expression classes of the
                      reference's stencils/*.py, composed here) parsed to a
                      kernel + coefficients and swept on the GPU, against
                      the oracle and against a numpy brute force of the
                      text's own taps.
"""
import numpy as np
import pytest

import oracle as O
from bricklib_amd import BK_F32, BK_F64, ingest, stencils as S
from bricklib_amd.decomp import SKIN3D_BAD, SKIN3D_NORMAL, DecompGrid

from _dsl import star, text_of

pytestmark = pytest.mark.gpu

DEEP = [("7pt", 8), ("7pt", 3), ("13pt", 4), ("25pt", 2), ("125pt", 4), ("125pt", 2),
        ("cond", 8)]


@pytest.mark.parametrize("name,gi", DEEP)
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_deep_ghost_matches_oracle_loop(gpu, orc, name, gi, use_graph):
    nb = (3, 2, 4)
    st = S.Stencil.make(name, seed=23)
    n = 2 * gi + 3
    for dt in ((BK_F64, BK_F32) if name in ("7pt", "125pt") else (BK_F64,)):
        print(f"[case] deep {name} gi={gi} dt={dt} graph={use_graph} n={n}", flush=True)
        g = gpu.PeriodicGrid(nb)
        a = gpu.Brick(g.info, g.field_storage(dt), 0)
        b = gpu.Brick(g.info, g.field_storage(dt), 0)
        a.storage.buf.zero()
        b.storage.buf.zero()
        g.fill_uniform(a, 5)
        res = gpu.sweeps(st, a, b, g, n, use_graph=use_graph, ghost_iters=gi)
        got = res.storage.to_host()
        og = O.Grid(nb)
        oa, ob = og.storage(dt), og.storage(dt)
        og.fill_uniform(oa, 5, dtype=dt)
        ores = og.deep_sweeps(st.kind, st.coeff, oa, ob, n, gi, dtype=dt, theta=st.theta,
                              cond_mode=st.cond_mode)
        np.testing.assert_array_equal(got, ores)
        # the interior is the refresh-every-sweep result (radius * gi <= 8)
        oa, ob = og.storage(dt), og.storage(dt)
        og.fill_uniform(oa, 5, dtype=dt)
        every = og.sweeps(st.kind, st.coeff, oa, ob, n, dtype=dt, theta=st.theta,
                          cond_mode=st.cond_mode)
        np.testing.assert_array_equal(got[og.interior], every[og.interior])


def _decomp_oracle(dg, og, oa_init, dt):
    """The oracle on the decomposition's numbering, its input the init_grid
    field `oa_init` moved brick by brick to the decomposition's ids."""
    lay = O.Layout(dg.info.adj, dg.interior, dg.pairs)
    src = lay.storage(dt)
    src[dg.grid.reshape(-1)] = oa_init[og.grid]
    return lay, src


@pytest.mark.parametrize("skin", ["good", "normal", "bad"])
@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "125pt", "cond"])
def test_decomp_layout_matches_oracle(gpu, orc, name, skin):
    nb = (4, 3, 5)
    st = S.Stencil.make(name, seed=43)
    dg = DecompGrid(nb, skin={"good": None, "normal": SKIN3D_NORMAL, "bad": SKIN3D_BAD}[skin])
    og = O.Grid(nb)
    for dt in ((BK_F64, BK_F32) if name in ("7pt", "125pt") else (BK_F64,)):
        print(f"[case] decomp {name} skin={skin} dt={dt}", flush=True)
        oa = og.storage(dt)
        og.fill_uniform(oa, 19, dtype=dt)
        lay, src = _decomp_oracle(dg, og, oa, dt)
        a = gpu.Brick(dg.info, dg.field_storage(dt), 0)
        b = gpu.Brick(dg.info, dg.field_storage(dt), 0)
        a.storage.buf.upload(src.reshape(-1))
        res = gpu.sweeps(st, a, b, dg, 5)
        got = res.storage.to_host()
        want = lay.sweeps(st.kind, st.coeff, src, lay.storage(dt), 5, dtype=dt, theta=st.theta,
                          cond_mode=st.cond_mode)
        np.testing.assert_array_equal(got[dg.interior], want[dg.interior])
        # the same values as the init_grid numbering, brick for brick
        ob = og.storage(dt)
        ores = og.sweeps(st.kind, st.coeff, oa, ob, 5, dtype=dt, theta=st.theta,
                         cond_mode=st.cond_mode)
        pk, pj, pi = og.padded[2], og.padded[1], og.padded[0]
        inner = og.grid.reshape(pk, pj, pi)[1:-1, 1:-1, 1:-1].reshape(-1)
        dinner = dg.grid[1:-1, 1:-1, 1:-1].reshape(-1)
        np.testing.assert_array_equal(got[dinner], ores[inner])


def _spec_cases():
    rng = np.random.default_rng(5)
    out = []
    for R, name in ((1, "7pt"), (2, "13pt"), (4, "25pt")):
        offs = star(R)
        perm = rng.permutation(len(offs))
        taps = [(offs[p], f"coeff[{p}]") for p in perm]
        out.append((name, text_of(taps), rng.uniform(0.05, 1.0, len(offs)), False))
    cube = [(di, dj, dk) for dk in range(-2, 3) for dj in range(-2, 3) for di in range(-2, 3)]
    perm = rng.permutation(125)
    out.append(("125pt", text_of([(cube[p], f"coeff[{p}]") for p in perm]),
                rng.uniform(-1, 1, 125) / 62.5, False))
    taps = [(d, f"coeff[{q}]") for q, d in enumerate(star(1))][::-1]
    out.append(("cond", text_of(taps, relu=True, select=True), rng.uniform(-1, 1, 7), True))
    return out


@pytest.mark.parametrize("case", _spec_cases(), ids=lambda c: c[0])
def test_ingested_stencil_on_gpu(gpu, orc, case):
    name, text, coeff, cond = case
    spec = ingest.parse_text(text)
    st = spec.stencil(coeff)
    assert st.kind == S.stencil_id(name)
    nb = (3, 4, 2)
    print(f"[case] ingested {name}", flush=True)
    g = gpu.PeriodicGrid(nb)
    a = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    b = gpu.Brick(g.info, g.field_storage(BK_F64), 0)
    g.fill_uniform(a, 61)
    res = gpu.sweeps(st, a, b, g, 2)
    got = res.storage.to_host()
    og = O.Grid(nb)
    oa, ob = og.storage(), og.storage()
    og.fill_uniform(oa, 61)
    inp = og.to_dense(oa)
    ores = og.sweeps(st.kind, st.coeff, oa, ob, 2, theta=st.theta, cond_mode=st.cond_mode)
    np.testing.assert_array_equal(got[og.interior], ores[og.interior])
    # against the text's own taps, applied twice by numpy on the periodic field
    want = inp
    for _ in range(2):
        nxt = np.zeros_like(want)
        for t in spec.taps:
            di, dj, dk = t.offset
            x = np.roll(want, shift=(-dk, -dj, -di), axis=(0, 1, 2))
            if t.relu:
                x = np.where(x < 0, 0.0, x)
            nxt += ingest._resolve(t.coeff, coeff) * x
        if spec.select_abs:
            nxt = np.where(nxt > 0, nxt, -nxt)
        want = nxt
    np.testing.assert_allclose(og.to_dense(got), want, rtol=0, atol=1e-13 * np.abs(want).max())
