"""Execution plans built on the host (bk_plan_build_host): every listed brick
is stored exactly once, units follow k-chains of the adjacency, quad members
are the documented neighbours, and the k-march order agrees with the grid."""
import numpy as np
import pytest

import oracle as O


def members(adj, a):
    m1 = adj[a][14]
    return [a, m1, adj[a][16], adj[m1][16]]


def check_plan(bk, nb, shape, kcol, order="grid", seed=0):
    g = O.Grid(nb)
    info = bk.BrickInfo(g.adj)
    blist = g.interior.copy()
    if order == "shuffled":
        np.random.default_rng(seed).shuffle(blist)
    p = bk.plan_host(info, blist, shape, kcol)
    # stored bricks == listed bricks, each once
    assert sorted(p["flat"].tolist()) == sorted(blist.tolist())
    assert len(set(p["flat"].tolist())) == p["flat"].size
    stored = []
    for u, a in enumerate(p["anchor"]):
        L = int(p["len"][u])
        assert 1 <= L <= kcol
        assert np.all(p["mask"][u, L:] == 0)
        for l in range(L):
            m = int(p["mask"][u, l])
            mem = [a] if shape == "column" else members(g.adj, a)
            for q, b in enumerate(mem):
                if m >> q & 1:
                    stored.append(int(b))
            a = g.adj[a][22]
    assert stored == p["flat"].tolist()
    return p, g


@pytest.mark.parametrize("nb", [(4, 4, 4), (3, 5, 7), (1, 1, 1), (2, 1, 9)])
@pytest.mark.parametrize("shape", ["column", "quad"])
@pytest.mark.parametrize("kcol", [1, 3, 16])
def test_plan_covers_each_brick_once(bk, nb, shape, kcol):
    check_plan(bk, nb, shape, kcol)


@pytest.mark.parametrize("shape", ["column", "quad"])
def test_plan_shuffled_list(bk, shape):
    check_plan(bk, (4, 3, 5), shape, 4, order="shuffled", seed=3)


def test_column_plan_of_grid_order_is_full_columns(bk):
    nb = (4, 4, 8)
    p, g = check_plan(bk, nb, "column", 8)
    assert p["anchor"].size == nb[0] * nb[1]
    assert np.all(p["len"] == 8)


def test_quad_plan_even_grid_is_full_quads(bk):
    nb = (4, 6, 4)
    p, g = check_plan(bk, nb, "quad", 4)
    assert p["anchor"].size == (nb[0] // 2) * (nb[1] // 2)
    assert np.all(p["mask"] == 0xF)


def test_quad_plan_odd_grid_has_partial_quads(bk):
    p, g = check_plan(bk, (3, 3, 2), "quad", 2)
    masks = set(p["mask"].reshape(-1).tolist())
    assert 0xF in masks and (0x5 in masks or 0x3 in masks or 0x1 in masks)


def test_plan_rejects_bad_input(bk):
    from bricklib_amd import BrickError
    g = O.Grid((2, 2, 2))
    info = bk.BrickInfo(g.adj)
    with pytest.raises(BrickError):
        bk.plan_host(info, np.array([g.nbricks + 5], np.uint32), "column", 4)
    with pytest.raises((BrickError, KeyError)):
        bk.plan_host(info, g.interior, "cube", 4)


def _padded(nb):
    return bk_init_grid([x + 2 for x in nb])


def bk_init_grid(dims, periodic=False):
    from bricklib_amd.bricks import init_grid
    return init_grid(dims, periodic)


@pytest.mark.parametrize("nb", [(3, 2, 4), (5, 7, 3), (1, 1, 1), (4, 4, 9)])
def test_quad_plans_need_consistent_neighbourhoods(bk, nb):
    """Interior quads of a padded grid and every-brick quads of a periodic
    grid build; every-brick quads of a padded (non-periodic) grid, whose
    outer layer has out-of-range neighbours, are rejected."""
    from bricklib_amd import BrickError
    from bricklib_amd.bricks import plan_host
    g, info = _padded(nb)
    S0, S1 = nb[0] + 2, nb[1] + 2
    interior = np.array([(i + 1) + S0 * ((j + 1) + S1 * (k + 1)) for k in range(nb[2])
                         for j in range(nb[1]) for i in range(nb[0])], np.uint32)
    for kcol in (1, 4, 8):
        plan_host(info, interior, "quad", kcol)
        with pytest.raises(BrickError, match="consistently"):
            plan_host(info, np.arange(info.nbricks, dtype=np.uint32), "quad", kcol)
    gp, infop = bk_init_grid(list(nb), periodic=True)
    for kcol in (1, 8):
        p = plan_host(infop, np.arange(infop.nbricks, dtype=np.uint32), "quad", kcol)
        assert np.sort(p["flat"]).tolist() == list(range(infop.nbricks))


@pytest.mark.parametrize("nb", [(3, 2, 4), (2, 2, 9)])
def test_column_links_agree_both_ways(bk, nb):
    """Column plans over every brick of a padded grid: consecutive bricks
    of a unit are k-neighbours in both directions, every brick once."""
    from bricklib_amd.bricks import plan_host
    g, info = _padded(nb)
    adj = info.adj
    for kcol in (2, 8, 64):
        p = plan_host(info, np.arange(info.nbricks, dtype=np.uint32), "column", kcol)
        assert np.sort(p["flat"]).tolist() == list(range(info.nbricks))
        pos = 0
        for a, ln in zip(p["anchor"], p["len"]):
            chain = p["flat"][pos:pos + ln]
            assert chain[0] == a
            for lo, hi in zip(chain, chain[1:]):
                assert adj[lo, 22] == hi and adj[hi, 4] == lo
            pos += ln
