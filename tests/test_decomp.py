"""The BrickDecomp layout (bk_decomp_layout), pinned without the reference
build (its BrickDecomp needs MPI): a hand-derived small case of the
skin3d_good order, the section boundaries, the reference's own adjacency
check (weak/main.cpp:87-94), and the exchange property -- every ghost run
is the periodic image of its skin run, brick by brick."""
import numpy as np
import pytest

from bricklib_amd.decomp import DecompLayout


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


@pytest.mark.parametrize("nb,bb", [((2, 2, 2), 4096), ((4, 3, 5), 4096), ((6, 6, 6), 4096),
                                   ((4, 3, 5), 2048), ((5, 4, 2), 1024)])
def test_sections_adjacency_and_exchange(nb, bb):
    L = DecompLayout(nb, 1, brick_bytes=bb)
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
