"""The BrickDecomp layout (bk_decomp_layout), pinned without the reference
build (its BrickDecomp needs MPI): a hand-derived small case of the
skin3d_good order, the section boundaries, the reference's own adjacency
check (weak/main.cpp:87-94), and the exchange property -- every ghost run
is the periodic image of its skin run, brick by brick."""
import numpy as np
import pytest

from bricklib_amd.decomp import SKIN3D_BAD, SKIN3D_GOOD, SKIN3D_NORMAL, DecompLayout


def coords(L):
    """id -> padded (i, j, k) for every grid position."""
    out = {}
    for k in range(L.padded[2]):
        for j in range(L.padded[1]):
            for i in range(L.padded[0]):
                out[int(L.grid[k, j, i])] = (i, j, k)
    return out


def test_two_cubed_follows_skin_order():
    # nb = 2: no inner bricks; the eight interior bricks are all corners of
    # the skin, numbered in skin3d_good order: {1,2,-3}, {1,2,3}, {-1,2,-3},
    # {-1,2,3}, {-1,-2,-3}, {-1,-2,3}, {1,-2,-3}, {1,-2,3}
    L = DecompLayout((2, 2, 2))
    want = {(2, 2, 1): 1, (2, 2, 2): 2, (1, 2, 1): 3, (1, 2, 2): 4,
            (1, 1, 1): 5, (1, 1, 2): 6, (2, 1, 1): 7, (2, 1, 2): 8}
    for (i, j, k), b in want.items():
        assert L.grid[k, j, i] == b
    assert L.sep == (1, 9, 65) and L.nbricks == 65


def test_skin3d_good_list_equals_default():
    a = DecompLayout((4, 3, 5), brick_bytes=2048)
    b = DecompLayout((4, 3, 5), brick_bytes=2048, skin=SKIN3D_GOOD)
    np.testing.assert_array_equal(a.grid, b.grid)
    np.testing.assert_array_equal(a.adj, b.adj)
    assert a.regions == b.regions and a.sep == b.sep


def test_skin3d_normal_three_cubed():
    # nb = 3, one brick per region: the inner brick first (id 1), then the
    # 26 regions in lexicographic order (k slowest, -1 first), the empty
    # centre entry taking no id
    L = DecompLayout((3, 3, 3), skin=SKIN3D_NORMAL)
    for k in range(1, 4):
        for j in range(1, 4):
            for i in range(1, 4):
                r = 9 * (k - 1) + 3 * (j - 1) + (i - 1)
                want = 1 if r == 13 else (2 + r if r < 13 else 1 + r)
                assert L.grid[k, j, i] == want, (i, j, k)
    assert L.sep == (2, 28, 126)
    assert L.skin_size == [1] * 13 + [0] + [1] * 13


def test_skin3d_bad_groups_pages_per_region():
    # nb = 2, 2048-byte bricks (two to a 4096-byte page): only corners have
    # bricks; each is alone between empty entries, so each takes a page:
    # interior ids 2, 4, .., 16 in lexicographic order, and each of the
    # 8 x 7 ghost bricks (a corner seen from each neighbour that sees it)
    # its own page: 18 + 2 * 56 = 130
    L = DecompLayout((2, 2, 2), brick_bytes=2048, skin=SKIN3D_BAD)
    ids = L.grid[1:3, 1:3, 1:3].reshape(-1)
    np.testing.assert_array_equal(ids, np.arange(2, 18, 2))
    assert L.sep == (2, 18, 130) and L.nbricks == 130
    assert sum(L.skin_size) == 16 and L.skin_size[1::2] == [0] * 25
    # every region is its own run: 26 neighbours x the regions containing
    # the opposite direction (9 per face, 3 per edge, 1 per corner)
    assert len(L.regions) == 6 * 9 + 12 * 3 + 8 * 1


def test_rejects_bad_skin_lists():
    from bricklib_amd import BrickError
    with pytest.raises(BrickError):
        DecompLayout((4, 4, 4), skin=SKIN3D_GOOD[:-1])  # a region missing
    with pytest.raises(BrickError):
        DecompLayout((4, 4, 4), skin=SKIN3D_GOOD[:-1] + [(1,)])  # a region twice
    with pytest.raises(ValueError):
        DecompLayout((4, 4, 4), skin=SKIN3D_GOOD[:-1] + [(4,)])  # no axis 4


SKINS = {"good": None, "normal": SKIN3D_NORMAL, "bad": SKIN3D_BAD,
         "shuffled": [SKIN3D_GOOD[q] for q in np.random.default_rng(5).permutation(26)]}


@pytest.mark.parametrize("skin", sorted(SKINS))
@pytest.mark.parametrize("nb,bb", [((2, 2, 2), 4096), ((4, 3, 5), 4096), ((6, 6, 6), 4096),
                                   ((4, 3, 5), 2048), ((5, 4, 2), 1024)])
def test_sections_adjacency_and_exchange(nb, bb, skin):
    L = DecompLayout(nb, 1, brick_bytes=bb, skin=SKINS[skin])
    factor = max(1, 4096 // bb)
    C = coords(L)
    ids = np.array(sorted(C))
    assert len(ids) == np.prod(L.padded) and ids.min() >= factor and ids.max() < L.nbricks
    assert len(set(ids.tolist())) == len(ids)
    g = 1

    def kind(c):
        inside = [g <= c[a] < g + nb[a] for a in range(3)]
        if not all(inside):
            return "ghost"
        inner = all(2 * g <= c[a] < nb[a] for a in range(3))
        return "inner" if inner else "skin"
    for b, c in C.items():
        k = kind(c)
        lo, hi = {"inner": (factor, L.sep[0]), "skin": (L.sep[0], L.sep[1]),
                  "ghost": (L.sep[1], L.sep[2])}[k]
        assert lo <= b < hi, (b, c, k)
    # adjacency: coordinates, symmetric for interior bricks, self at 13
    for b, (i, j, k) in C.items():
        interior = kind((i, j, k)) != "ghost"
        for s in range(27):
            di, dj, dk = s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1
            n = (i + di, j + dj, k + dk)
            inr = all(0 <= n[a] < L.padded[a] for a in range(3))
            assert L.adj[b, s] == (L.grid[n[2], n[1], n[0]] if inr else 0)
            if interior:
                assert L.adj[L.adj[b, s], 26 - s] == b
        assert L.adj[b, 13] == b
    # every neighbour has its runs; a ghost run mirrors its skin run
    seen = set()
    for r in L.regions:
        seen.add(r["neighbor"])
        for q in range(r["len"]):
            gb, sb = r["ghost_pos"] + q, r["skin_pos"] + q
            assert (gb in C) == (sb in C)  # padding at the same offsets
            if gb not in C:
                continue
            gc, sc = C[gb], C[sb]
            for a in range(3):
                d = r["neighbor"][a]
                if d > 0:
                    assert gc[a] >= g + nb[a]
                elif d < 0:
                    assert gc[a] < g
                else:
                    assert g <= gc[a] < g + nb[a]
                assert (gc[a] - g) % nb[a] == (sc[a] - g) % nb[a]
            assert kind(sc) != "ghost"
    assert len(seen) == 26
    # the runs cover every ghost brick exactly once
    covered = [r["ghost_pos"] + q for r in L.regions for q in range(r["len"])
               if r["ghost_pos"] + q in C]
    assert sorted(covered) == sorted(b for b, c in C.items() if kind(c) == "ghost")


def test_rejects_bad_arguments():
    from bricklib_amd import BrickError
    with pytest.raises(BrickError):
        DecompLayout((1, 4, 4))  # fewer than 2 * ghost bricks on an axis
    with pytest.raises(BrickError):
        DecompLayout((4, 4, 4), brick_bytes=3000)  # not a factor of the page


@pytest.mark.parametrize("skin", ["good", "normal", "bad"])
@pytest.mark.parametrize("name", ["7pt", "25pt", "125pt", "cond"])
def test_oracle_on_decomp_numbering_matches_init_grid(name, skin):
    """The oracle's brick loop on the BrickDecomp numbering (its adjacency
    and its exchange pairs as the refresh) gives the init_grid numbering's
    values brick for brick: the layout is a valid periodic grid."""
    import oracle as O
    from bricklib_amd import stencils as S
    from bricklib_amd.decomp import SKIN3D_BAD, SKIN3D_NORMAL, DecompLayout
    nb = (4, 3, 5)
    L = DecompLayout(nb, skin={"good": None, "normal": SKIN3D_NORMAL, "bad": SKIN3D_BAD}[skin])
    st = S.Stencil.make(name, seed=43)
    og = O.Grid(nb)
    oa = og.storage()
    og.fill_uniform(oa, 19)
    lay = O.Layout(L.adj, L.interior(), L.periodic_pairs())
    src = lay.storage()
    src[L.grid.reshape(-1)] = oa[og.grid]
    got = lay.sweeps(st.kind, st.coeff, src, lay.storage(), 5, theta=st.theta,
                     cond_mode=st.cond_mode)
    want = og.sweeps(st.kind, st.coeff, oa, og.storage(), 5, theta=st.theta,
                     cond_mode=st.cond_mode)
    pk, pj, pi = og.padded[2], og.padded[1], og.padded[0]
    inner = og.grid.reshape(pk, pj, pi)[1:-1, 1:-1, 1:-1].reshape(-1)
    dinner = L.grid[1:-1, 1:-1, 1:-1].reshape(-1)
    np.testing.assert_array_equal(got[dinner], want[inner])
