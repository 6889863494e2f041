"""Brick shapes and vector folds without a GPU: the oracle's addressing of
the 4x8x16 brick in the 4x16 and 8x8 folds against the reference's own
accessor (oracle/_ref, brick.h:235-246, 276-284) -- every element of a brick
and every cell of its 3x3x3 neighbourhood through the adjacency -- and the
oracle's sweeps over those layouts against the 8x8x8 sweeps (a cell's value
cannot depend on how the cells are packed into bricks), plus the library's
host helpers for layouts (extents, tiled sweep order)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from bricklib_amd import stencils as S

LAYOUTS = [O.LAYOUT_4816, O.LAYOUT_4816_F88]


@pytest.mark.parametrize("lay,fold", [(O.LAYOUT_888, 0), (O.LAYOUT_4816, 5),
                                      (O.LAYOUT_4816_F88, 6)])
def test_element_offsets_match_reference_accessor(orc, lay, fold):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    grid, adj = O.init_grid((3, 3, 3))
    exp = np.empty(512, np.int64)
    assert O.ref().ref_fold_offsets(fold, O.P(adj), adj.shape[0], 13, C.c_size_t(512),
                                    O.P(exp)) == 0
    BK, BJ, BI = O.layout_extents(lay)
    got = np.array([13 * 512 + O.lib().or_layout_offset(lay, k, j, i)
                    for k in range(BK) for j in range(BJ) for i in range(BI)])
    np.testing.assert_array_equal(got, exp)
    assert sorted(got - 13 * 512) == list(range(512))  # a permutation of the brick


@pytest.mark.parametrize("lay", LAYOUTS)
def test_neighbour_addressing_matches_reference_accessor(orc, lay):
    """(k, j, i) anywhere in the 3x3x3 neighbourhood of brick 13: the
    neighbour digit (x + B) / B per axis with each axis's own extent."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    R = O.ref()
    R.ref_neighbor_offset_layout.restype = C.c_long
    R.ref_neighbor_offset_layout.argtypes = [C.c_int, C.c_void_p, C.c_uint, C.c_uint,
                                             C.c_size_t, C.c_int, C.c_int, C.c_int]
    grid, adj = O.init_grid((3, 3, 3))
    BK, BJ, BI = O.layout_extents(lay)
    for k in range(-BK, 2 * BK, 3):
        for j in range(-BJ, 2 * BJ, 3):
            for i in range(-BI, 2 * BI, 5):
                kd, jd, id_ = (k + BK) // BK, (j + BJ) // BJ, (i + BI) // BI
                e = O.lib().or_layout_offset(lay, (k + BK) % BK, (j + BJ) % BJ, (i + BI) % BI)
                mine = int(adj[13][kd * 9 + jd * 3 + id_]) * 512 + e
                assert mine == R.ref_neighbor_offset_layout(lay, O.P(adj), adj.shape[0], 13,
                                                            512, k, j, i), (k, j, i)


def _dense_888(cells, st, n, dtype, seed=5):
    g = O.Grid((cells[0] // 8, cells[1] // 8, cells[2] // 8))
    a, b = g.storage(dtype), g.storage(dtype)
    g.fill_uniform(a, seed, dtype=dtype)
    r = g.sweeps(st.kind, st.coeff, a, b, n, dtype=dtype, theta=st.theta,
                 cond_mode=st.cond_mode)
    lg = O.LayoutGrid(g.nb, O.LAYOUT_888)
    return lg.to_dense(r, dtype)


@pytest.mark.parametrize("name", ["7pt", "13pt", "25pt", "125pt", "cond"])
@pytest.mark.parametrize("lay", LAYOUTS)
def test_layout_sweeps_equal_888_sweeps(orc, name, lay):
    cells = (32, 32, 16)
    st = S.Stencil.make(name, seed=4)
    for dtype in (O.F64, O.F32):
        want = _dense_888(cells, st, 3, dtype)
        BK, BJ, BI = O.layout_extents(lay)
        for ghost in (1, 0):
            lg = O.LayoutGrid((cells[0] // BI, cells[1] // BJ, cells[2] // BK), lay, ghost)
            a, b = lg.storage(dtype), lg.storage(dtype)
            lg.fill_uniform(a, 5, dtype=dtype)
            r = lg.sweeps(st.kind, st.coeff, a, b, 3, dtype=dtype, theta=st.theta,
                          cond_mode=st.cond_mode, refresh=bool(ghost))
            np.testing.assert_array_equal(lg.to_dense(r, dtype), want)


def test_layout_sweep_catches_planted_mistakes(orc):
    """The layout comparison is not vacuous: an i-tap and a j-tap
    coefficient swapped (the stars' coefficients are symmetric per axis, so
    the swap crosses axes) or a mis-addressed neighbour brick shows up."""
    cells = (64, 16, 16)  # four bricks along i: the i-1 and i+1 bricks differ
    st = S.Stencil.make("13pt", seed=6)
    want = _dense_888(cells, st, 1, O.F64)
    lg = O.LayoutGrid((4, 2, 4), O.LAYOUT_4816)
    c = np.array(st.coeff)
    assert c[1] != c[3]
    c[[1, 3]] = c[[3, 1]]  # i+1 <-> j+1
    a, b = lg.storage(), lg.storage()
    lg.fill_uniform(a, 5)
    r = lg.sweeps(st.kind, c, a, b, 1)
    assert not np.array_equal(lg.to_dense(r), want)
    a, b = lg.storage(), lg.storage()
    lg.fill_uniform(a, 5)
    lg.adj = lg.adj.copy()
    lg.adj[:, [12, 14]] = lg.adj[:, [14, 12]]  # i-1 / i+1 neighbour swapped
    r = lg.sweeps(st.kind, st.coeff, a, b, 1)
    assert not np.array_equal(lg.to_dense(r), want)


def test_fill_values_do_not_depend_on_layout(orc):
    cells = (32, 24, 8)
    ref = _dense_888(cells, S.Stencil.make("7pt", seed=1), 0, O.F64, seed=9)
    for lay in LAYOUTS:
        BK, BJ, BI = O.layout_extents(lay)
        for ghost in (0, 1):
            lg = O.LayoutGrid((cells[0] // BI, cells[1] // BJ, cells[2] // BK), lay, ghost)
            a = lg.storage()
            lg.fill_uniform(a, 9)
            np.testing.assert_array_equal(lg.to_dense(a), ref)


def test_library_layout_extents(bk):
    lib = bk.load()
    for lay in (0, 1, 2):
        ext = (C.c_int * 3)()
        assert lib.bk_layout_extents(lay, ext) == 0
        assert tuple(ext) == O.layout_extents(lay)
    assert lib.bk_layout_extents(3, (C.c_int * 3)()) != 0


@pytest.mark.parametrize("nb,ghost", [((3, 5, 7), 0), ((3, 5, 7), 1), ((40, 33, 2), 1)])
def test_tiled_list_is_a_tiled_permutation(bk, nb, ghost):
    """bk_tiled_list: every interior brick once, tile by tile, each tile
    through all k before the next."""
    from bricklib_amd._lib import longs, u32p
    lib = bk.load()
    n = C.c_uint32()
    assert lib.bk_tiled_list(longs(nb), ghost, 32, 2, None, C.byref(n)) == 0
    out = np.empty(n.value, np.uint32)
    assert lib.bk_tiled_list(longs(nb), ghost, 32, 2, out.ctypes.data_as(u32p), C.byref(n)) == 0
    padded = [x + 2 * ghost for x in nb]
    grid, _ = O.init_grid(padded, periodic=not ghost)
    g3 = grid.reshape(padded[2], padded[1], padded[0])
    inner = g3[ghost:ghost + nb[2], ghost:ghost + nb[1], ghost:ghost + nb[0]]
    assert sorted(out) == sorted(inner.reshape(-1))
    want = []
    for j0 in range(0, nb[1], 2):
        for i0 in range(0, nb[0], 32):
            want.append(inner[:, j0:j0 + 2, i0:i0 + 32].reshape(-1))
    np.testing.assert_array_equal(out, np.concatenate(want))
