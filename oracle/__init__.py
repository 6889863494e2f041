"""ctypes wrapper of the oracle (oracle/oracle.c) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py, always as the reference the GPU results are
compared with, never as the thing measured on the GPU or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

F64, F32, BF16 = 0, 1, 2
STAR7, STAR13, STAR25, CUBE125, COND7 = 0, 1, 2, 3, 4
MODE_FMA, MODE_REF = 0, 1
NP = {F64: np.float64, F32: np.float32, BF16: np.uint16}

u32p = C.POINTER(C.c_uint32)
dblp = C.POINTER(C.c_double)
longp = C.POINTER(C.c_long)
vp = C.c_void_p

_SIG = {
    "or_hash64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "or_u01": (C.c_double, [C.c_uint64, C.c_uint64]),
    "or_f32_to_bf16": (C.c_uint16, [C.c_float]),
    "or_bf16_to_f32": (C.c_float, [C.c_uint16]),
    "or_init_grid": (None, [longp, C.c_int, u32p, u32p]),
    "or_interior_list": (C.c_uint32, [longp, u32p]),
    "or_ghost_pairs": (C.c_uint32, [longp, u32p]),
    "or_fill_uniform": (None, [C.c_int, vp, C.c_size_t, longp, C.c_uint64, C.c_double,
                               C.c_double]),
    "or_fill_uniform_array": (None, [dblp, longp, C.c_uint64, C.c_double, C.c_double]),
    "or_ghost_refresh": (None, [C.c_int, vp, C.c_size_t, C.c_size_t, u32p, C.c_uint32]),
    "or_copy_to_brick": (None, [C.c_int, longp, longp, longp, vp, u32p, vp, C.c_size_t]),
    "or_copy_from_brick": (None, [C.c_int, longp, longp, longp, vp, u32p, vp, C.c_size_t]),
    "or_stencil_radius": (C.c_int, [C.c_int]),
    "or_stencil_ncoeff": (C.c_int, [C.c_int]),
    "or_brick_sweep": (None, [C.c_int, C.c_int, C.c_int, u32p, u32p, C.c_uint32, vp,
                              C.c_size_t, vp, C.c_size_t, dblp, C.c_double, C.c_int]),
    "or_brick_sweep_7pt_accessor": (None, [u32p, u32p, C.c_uint32, dblp, C.c_size_t, dblp,
                                           C.c_size_t, dblp]),
    "or_array_sweep": (None, [C.c_int, C.c_int, longp, C.c_long, dblp, dblp, dblp,
                              C.c_double, C.c_int]),
    "or_array_periodic_fill": (None, [dblp, longp, C.c_long]),
    "or_compare_brick": (C.c_int, [longp, longp, longp, dblp, u32p, dblp, C.c_size_t,
                                   C.c_double]),
    "or_field_compare": (None, [C.c_int, vp, C.c_size_t, C.c_int, vp, C.c_size_t, u32p,
                                C.c_uint32, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                dblp]),
    "or_layout_extent": (C.c_int, [C.c_int, C.c_int]),
    "or_layout_offset": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "or_fill_uniform_layout": (None, [C.c_int, C.c_int, vp, C.c_size_t, longp, C.c_int,
                                      C.c_uint64, C.c_double, C.c_double]),
    "or_to_dense_layout": (None, [C.c_int, C.c_int, vp, C.c_size_t, longp, C.c_int, dblp]),
    "or_brick_sweep_layout": (None, [C.c_int, C.c_int, C.c_int, C.c_int, u32p, u32p,
                                     C.c_uint32, vp, C.c_size_t, vp, C.c_size_t, dblp,
                                     C.c_double, C.c_int]),
    "or_num_threads": (C.c_int, []),
    "or_set_num_threads": (None, [C.c_int]),
    "or_wtime": (C.c_double, []),
}

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing; run `make -C oracle`")
        l = C.CDLL(LIB)
        for n, (r, a) in _SIG.items():
            f = getattr(l, n)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref() -> C.CDLL:
    """oracle/_ref/libref.so: the reference's own headers compiled here."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_LIB} missing; run `make -C oracle ref`")
        _ref = C.CDLL(REF_LIB)
    return _ref


def L3(v):
    return (C.c_long * 3)(*[int(x) for x in v])


def P(a: np.ndarray, t=None):
    return a.ctypes.data_as(t or vp)


# ---------------------------------------------------------------------------
def init_grid(dims, periodic=False):
    n = int(np.prod(dims))
    grid = np.empty(n, np.uint32)
    adj = np.empty(n * 27, np.uint32)
    lib().or_init_grid(L3(dims), int(periodic), P(grid, u32p), P(adj, u32p))
    return grid, adj.reshape(n, 27)


def interior_list(nb):
    n = lib().or_interior_list(L3(nb), None)
    out = np.empty(n, np.uint32)
    lib().or_interior_list(L3(nb), P(out, u32p))
    return out


def ghost_pairs(nb):
    n = lib().or_ghost_pairs(L3(nb), None)
    out = np.empty(2 * n, np.uint32)
    lib().or_ghost_pairs(L3(nb), P(out, u32p))
    return out


class Layout:
    """The oracle over any brick numbering: an adjacency adj[nbricks][27],
    the interior list a sweep covers and the (dst, src) ghost pairs a
    refresh copies -- e.g. the BrickDecomp layout."""

    def __init__(self, adj, interior, pairs):
        self.adj = np.ascontiguousarray(adj, np.uint32).reshape(-1, 27)
        self.interior = np.ascontiguousarray(interior, np.uint32)
        self.pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1)
        self.nbricks = self.adj.shape[0]

    def deep_sweeps(self, stencil, coeff, a, b, n, ghost_iters, dtype=F64, mode=MODE_FMA,
                    theta=0.0, cond_mode=0):
        """The communication-avoiding loop of weak/main.cpp:26-36,203-211:
        the ghost bricks of the input are refreshed before sweeps
        0, g, 2g, ...; the last sweep of each group of g (and of the
        sequence) covers the interior (skip = 1), the others every brick of
        the layout, ghost bricks included (skip = 0).  Returns the array
        holding the result."""
        every = np.arange(self.nbricks, dtype=np.uint32)
        src, dst = a, b
        for t in range(n):
            if t % ghost_iters == 0:
                self.refresh(src, dtype)
            last = t % ghost_iters == ghost_iters - 1 or t == n - 1
            self.sweep(stencil, coeff, src, dst, dtype, mode, theta, cond_mode,
                       blist=None if last else every)
            src, dst = dst, src
        return src

    def storage(self, dtype=F64, fields=1):
        return np.zeros((self.nbricks, 512 * fields), NP[dtype])

    def refresh(self, dat, dtype=F64, offset=0):
        base = dat.reshape(-1)[offset:]
        lib().or_ghost_refresh(dtype, P(base), dat.shape[1], 512, P(self.pairs, u32p),
                               self.pairs.size // 2)

    def sweep(self, stencil, coeff, inp, out, dtype=F64, mode=MODE_FMA, theta=0.0,
              cond_mode=0, in_off=0, out_off=0, blist=None):
        coeff = np.ascontiguousarray(coeff, np.float64)
        blist = self.interior if blist is None else np.ascontiguousarray(blist, np.uint32)
        lib().or_brick_sweep(stencil, dtype, mode, P(self.adj, u32p), P(blist, u32p),
                             blist.size, P(inp.reshape(-1)[in_off:]), inp.shape[1],
                             P(out.reshape(-1)[out_off:]), out.shape[1], P(coeff, dblp),
                             theta, cond_mode)

    def sweeps(self, stencil, coeff, a, b, n, dtype=F64, mode=MODE_FMA, theta=0.0,
               cond_mode=0, refresh=True):
        """ping-pong a->b->a..., refreshing ghosts of the input first;
        returns the array holding the result"""
        src, dst = a, b
        for _ in range(n):
            if refresh:
                self.refresh(src, dtype)
            self.sweep(stencil, coeff, src, dst, dtype, mode, theta, cond_mode)
            src, dst = dst, src
        return src


class Grid(Layout):
    """Host twin of bricklib_amd.PeriodicGrid: (nb+2)^3 bricks, one ghost
    layer, numbering/adjacency of init_grid."""

    def __init__(self, nb):
        self.nb = tuple(int(x) for x in nb)
        self.padded = tuple(x + 2 for x in self.nb)
        self.grid, adj = init_grid(self.padded)
        super().__init__(adj, interior_list(self.nb), ghost_pairs(self.nb))

    def fill_uniform(self, dat, seed, lo=-1.0, hi=1.0, dtype=F64, step=None, offset=0):
        step = step or dat.shape[1]
        base = dat.reshape(-1)[offset:]
        lib().or_fill_uniform(dtype, P(base), step, L3(self.nb), seed, lo, hi)

    def to_dense(self, dat, offset=0):
        """interior field as a dense (n2, n1, n0) array (x fastest)."""
        nb0, nb1, nb2 = self.nb
        S0, S1 = self.padded[0], self.padded[1]
        out = np.empty((nb2 * 8, nb1 * 8, nb0 * 8), dat.dtype)
        f = dat[:, offset:offset + 512].reshape(-1, 8, 8, 8)
        for tk in range(nb2):
            for tj in range(nb1):
                idx = [(ti + 1) + S0 * ((tj + 1) + S1 * (tk + 1)) for ti in range(nb0)]
                blk = f[idx]  # (nb0, 8, 8, 8) [ti][k][j][i]
                out[tk * 8:(tk + 1) * 8, tj * 8:(tj + 1) * 8, :] = \
                    blk.transpose(1, 2, 0, 3).reshape(8, 8, nb0 * 8)
        return out

    def from_dense(self, arr, dat, offset=0):
        nb0, nb1, nb2 = self.nb
        S0, S1 = self.padded[0], self.padded[1]
        for tk in range(nb2):
            for tj in range(nb1):
                for ti in range(nb0):
                    b = (ti + 1) + S0 * ((tj + 1) + S1 * (tk + 1))
                    dat[b, offset:offset + 512] = arr[tk * 8:(tk + 1) * 8, tj * 8:(tj + 1) * 8,
                                                      ti * 8:(ti + 1) * 8].reshape(-1)


LAYOUT_888, LAYOUT_4816, LAYOUT_4816_F88 = 0, 1, 2


def layout_extents(layout):
    """(k, j, i) extents of a layout's bricks."""
    return tuple(lib().or_layout_extent(layout, a) for a in range(3))


class LayoutGrid(Layout):
    """A periodic grid of `nb` interior bricks per axis (contiguous-first) in
    a brick layout (0: 8x8x8 fold 8x8, 1: 4x8x16 fold 4x16, 2: 4x8x16 fold
    8x8), numbered by init_grid over (nb+2) bricks per axis with one ghost
    layer, or over nb with the adjacency wrapped (ghost=0)."""

    def __init__(self, nb, layout, ghost=1):
        self.nb = tuple(int(x) for x in nb)
        self.layout = int(layout)
        self.ghost = int(ghost)
        self.ext = layout_extents(self.layout)  # (BK, BJ, BI)
        self.cells = (self.nb[0] * self.ext[2], self.nb[1] * self.ext[1],
                      self.nb[2] * self.ext[0])  # contiguous-first
        if ghost:
            self.padded = tuple(x + 2 for x in self.nb)
            grid, adj = init_grid(self.padded)
            interior, pairs = interior_list(self.nb), ghost_pairs(self.nb)
        else:
            self.padded = self.nb
            grid, adj = init_grid(self.nb, periodic=True)
            interior, pairs = np.arange(grid.size, dtype=np.uint32), np.zeros(0, np.uint32)
        self.grid = grid
        super().__init__(adj, interior, pairs)

    def fill_uniform(self, dat, seed, lo=-1.0, hi=1.0, dtype=F64):
        lib().or_fill_uniform_layout(self.layout, dtype, P(dat), dat.shape[1], L3(self.nb),
                                     self.ghost, seed, lo, hi)

    def to_dense(self, dat, dtype=F64):
        """the interior as a dense FP64 (n2, n1, n0) array, x fastest"""
        out = np.empty((self.cells[2], self.cells[1], self.cells[0]), np.float64)
        lib().or_to_dense_layout(self.layout, dtype, P(dat), dat.shape[1], L3(self.nb),
                                 self.ghost, P(out, dblp))
        return out

    def sweep(self, stencil, coeff, inp, out, dtype=F64, mode=MODE_FMA, theta=0.0,
              cond_mode=0, in_off=0, out_off=0, blist=None):
        coeff = np.ascontiguousarray(coeff, np.float64)
        blist = self.interior if blist is None else np.ascontiguousarray(blist, np.uint32)
        lib().or_brick_sweep_layout(self.layout, stencil, dtype, mode, P(self.adj, u32p),
                                    P(blist, u32p), blist.size, P(inp.reshape(-1)[in_off:]),
                                    inp.shape[1], P(out.reshape(-1)[out_off:]), out.shape[1],
                                    P(coeff, dblp), theta, cond_mode)


def field_compare(a, r, interior, dta=F64, dtr=F64, floor=1.0, atol=0.0, rtol=0.0):
    out = np.zeros(5)
    lib().or_field_compare(dta, P(a), a.shape[1], dtr, P(r), r.shape[1], P(interior, u32p),
                           interior.size, 512, floor, atol, rtol, P(out, dblp))
    return {"max_abs": out[0], "max_rel": out[1], "sum_sq_diff": out[2], "sum_sq_ref": out[3],
            "bad": int(out[4])}


def bf16_to_f32(a: np.ndarray) -> np.ndarray:
    return (a.astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r
