"""The oracle against the reference's own headers compiled here (oracle/_ref).

Pins: brick numbering and adjacency (init_grid), element addressing through
the accessor, copyToBrick / copyFromBrick, compareBrick's rule, and the
brick-loop 7pt and cond sweeps (bit for bit in the reference order).
Skipped where oracle/_ref was not built (it needs the reference tree).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def ref_grid(dims):
    R = O.ref()
    n = int(np.prod(dims))
    grid = np.empty(n, np.uint32)
    adj = np.empty(n * 27, np.uint32)
    R.ref_init_grid.restype = C.c_int
    nb = R.ref_init_grid(O.L3(dims), O.P(grid), O.P(adj))
    assert nb == n
    return grid, adj.reshape(n, 27)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (4, 4, 4), (5, 2, 7), (18, 18, 18)])
def test_init_grid_matches_reference(dims):
    g_ref, a_ref = ref_grid(dims)
    g, a = O.init_grid(dims, periodic=False)
    np.testing.assert_array_equal(g, g_ref)
    np.testing.assert_array_equal(a, a_ref)


def test_library_init_grid_matches_reference(bk):
    dims = (6, 5, 4)
    g_ref, a_ref = ref_grid(dims)
    g, info = bk.init_grid(dims)
    np.testing.assert_array_equal(g.reshape(-1), g_ref)
    np.testing.assert_array_equal(info.adj, a_ref)


def test_accessor_neighbor_addressing():
    """adj[b][kd*9+jd*3+id]*step + ((k+8)&7)*64 + ((j+8)&7)*8 + ((i+8)&7)
    equals the reference accessor's address for k, j, i in [-8, 16)."""
    R = O.ref()
    R.ref_neighbor_offset.restype = C.c_long
    R.ref_neighbor_offset.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_size_t, C.c_int,
                                      C.c_int, C.c_int]
    dims = (5, 4, 3)
    grid, adj = O.init_grid(dims)
    step = 1024
    rng = np.random.default_rng(3)
    b = int(grid[1 + 5 * (1 + 4 * 1)])  # an interior brick
    for _ in range(400):
        k, j, i = (int(x) for x in rng.integers(-8, 16, 3))
        exp = int(adj[b][((k + 8) >> 3) * 9 + ((j + 8) >> 3) * 3 + ((i + 8) >> 3)]) * step + \
            ((k + 8) & 7) * 64 + ((j + 8) & 7) * 8 + ((i + 8) & 7)
        got = R.ref_neighbor_offset(O.P(adj), adj.shape[0], b, step, k, j, i)
        assert got == exp, (k, j, i)


FOLD_CASES = {0: (8, 8), 1: (8,), 2: (2, 4), 3: (4, 8), 4: (2, 2)}


def fold_offsets_py(fold, b, step):
    """Closed form of the accessor offset for an 8x8x8 brick in `fold`:
    the unfolded outer dims and the D/F parts of the folded inner dims form
    the vector index n; the F parts form the lane w; offset = n*VECLEN + w."""
    dims = [8, 8, 8]
    nf = len(fold)
    veclen = int(np.prod(fold))
    out = np.empty(512, np.int64)
    q = 0
    for k in range(8):
        for j in range(8):
            for i in range(8):
                idx = [k, j, i]
                n, w = 0, 0
                for d in range(3):
                    D = dims[d]
                    if d < 3 - nf:
                        n = n * D + idx[d]
                    else:
                        F = fold[d - (3 - nf)]
                        n = n * (D // F) + idx[d] // F
                        w = w * F + idx[d] % F
                out[q] = b * step + n * veclen + w
                q += 1
    return out


@pytest.mark.parametrize("fold", sorted(FOLD_CASES))
def test_fold_offsets_match_reference(fold):
    R = O.ref()
    dims = (3, 3, 3)
    grid, adj = O.init_grid(dims)
    b, step = 13, 512
    out = np.empty(512, np.int64)
    assert R.ref_fold_offsets(fold, O.P(adj), adj.shape[0], b, C.c_size_t(step), O.P(out)) == 0
    np.testing.assert_array_equal(out, fold_offsets_py(FOLD_CASES[fold], b, step))
    if FOLD_CASES[fold] in ((8, 8), (8,)):  # the HIP/AVX-512 folds are row-major
        np.testing.assert_array_equal(out - b * step, np.arange(512))


@pytest.mark.parametrize("pad,ghost", [(0, 0), (8, 0), (8, 8), (3, 8)])
def test_copy_to_from_brick_match_reference(pad, ghost):
    R = O.ref()
    dims = [16, 24, 8]
    bdims = [(d + 2 * ghost) // 8 for d in dims]
    grid, adj = O.init_grid(bdims)
    nb = adj.shape[0]
    ext = [d + 2 * (pad + ghost) for d in dims]
    rng = np.random.default_rng(7)
    arr = rng.standard_normal(int(np.prod(ext)))
    dat_ref = np.zeros((nb, 512))
    dat_or = np.zeros((nb, 512))
    L = O.L3
    R.ref_copy_to_brick(L(dims), L([pad] * 3), L([ghost] * 3), O.P(arr), O.P(grid), O.P(adj),
                        nb, O.P(dat_ref), C.c_size_t(512))
    O.lib().or_copy_to_brick(O.F64, L(dims), L([pad] * 3), L([ghost] * 3), O.P(arr),
                             O.P(grid, O.u32p), O.P(dat_or), 512)
    np.testing.assert_array_equal(dat_or, dat_ref)
    back_ref = np.zeros_like(arr)
    back_or = np.zeros_like(arr)
    R.ref_copy_from_brick(L(dims), L([pad] * 3), L([ghost] * 3), O.P(back_ref), O.P(grid),
                          O.P(adj), nb, O.P(dat_ref), C.c_size_t(512))
    O.lib().or_copy_from_brick(O.F64, L(dims), L([pad] * 3), L([ghost] * 3), O.P(back_or),
                               O.P(grid, O.u32p), O.P(dat_or), 512)
    np.testing.assert_array_equal(back_or, back_ref)


def test_compare_rule_matches_reference():
    R = O.ref()
    dims = [8, 8, 8]
    grid, adj = O.init_grid([1, 1, 1])
    arr = np.linspace(-2, 2, 512)
    L = O.L3
    for delta, tol in [(0.0, 1e-6), (5e-7, 1e-6), (2e-6, 1e-6), (1e-3, 1e-6)]:
        dat = (arr + delta).reshape(1, 512).copy()
        r = R.ref_compare_brick(L(dims), L([0] * 3), L([0] * 3), O.P(arr), O.P(grid), O.P(adj),
                                1, O.P(dat), C.c_size_t(512))
        o = O.lib().or_compare_brick(L(dims), L([0] * 3), L([0] * 3), O.P(arr, O.dblp),
                                     O.P(grid, O.u32p), O.P(dat, O.dblp), 512, tol)
        assert bool(r) == bool(o), delta


@pytest.mark.parametrize("which", ["7pt", "cond"])
def test_brick_sweep_matches_reference_bit_for_bit(which):
    """Oracle (reference order, no FMA) == the reference accessor loop,
    stencils/3axis.cpp:51-66 and :123-140, over a periodic grid."""
    R = O.ref()
    g = O.Grid((3, 4, 2))
    a = g.storage()
    g.fill_uniform(a, seed=11)
    g.refresh(a)
    coeff = np.random.default_rng(5).uniform(0, 1, 7)
    out_ref = np.zeros_like(a)
    fn = R.ref_sweep_7pt if which == "7pt" else R.ref_sweep_cond
    fn(O.P(g.adj), g.nbricks, O.P(g.interior), g.interior.size, O.P(a), C.c_size_t(512),
       O.P(out_ref), C.c_size_t(512), O.P(coeff))
    st = O.STAR7 if which == "7pt" else O.COND7
    out_or = np.zeros_like(a)
    g.sweep(st, coeff, a, out_or, mode=O.MODE_REF, cond_mode=0)
    np.testing.assert_array_equal(out_or, out_ref)
    if which == "7pt":
        out_acc = np.zeros_like(a)
        O.lib().or_brick_sweep_7pt_accessor(O.P(g.adj, O.u32p), O.P(g.interior, O.u32p),
                                            g.interior.size, O.P(a, O.dblp), 512,
                                            O.P(out_acc, O.dblp), 512, O.P(coeff, O.dblp))
        np.testing.assert_array_equal(out_acc, out_ref)


def test_adjacency_symmetry_invariant():
    """adj[adj[l][s]][26-s] == l and adj[l][13] == l for interior bricks
    (weak/main.cpp:87-94, stencils/fake.h:376-377)."""
    g = O.Grid((4, 3, 5))
    for l in g.interior:
        assert g.adj[l][13] == l
        for s in range(27):
            assert g.adj[g.adj[l][s]][26 - s] == l
