"""The C-ABI library without a GPU: it loads, exports every symbol that
include/bk_api.h declares, its host-side helpers agree with the oracle, and
argument errors come back as status codes (checked before any HIP call)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        if 'extern "C"' not in txt:
            continue
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bk_[a-z_0-9]+)\s*\(", txt,
                            flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol(bk):
    lib = bk.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header(bk):
    from bricklib_amd._lib import SIGNATURES
    assert set(declared_functions()) <= set(SIGNATURES)


def test_stencil_tables(bk):
    lib = bk.load()
    for st in range(5):
        assert lib.bk_stencil_ncoeff(st) == O.lib().or_stencil_ncoeff(st)
        assert lib.bk_stencil_radius(st) == O.lib().or_stencil_radius(st)
    assert lib.bk_stencil_ncoeff(99) == 0
    assert [lib.bk_dtype_size(d) for d in (0, 1, 2, 7)] == [8, 4, 2, 0]


@pytest.mark.parametrize("nb", [(1, 1, 1), (2, 3, 4), (7, 1, 2)])
def test_grid_helpers_match_oracle(bk, nb):
    from bricklib_amd._lib import longs, u32p
    lib = bk.load()
    padded = [x + 2 for x in nb]
    g_or, a_or = O.init_grid(padded)
    g, info = bk.init_grid(padded)
    np.testing.assert_array_equal(g.reshape(-1), g_or)
    np.testing.assert_array_equal(info.adj, a_or)
    gp, ap = O.init_grid(padded, periodic=True)
    g2, info2 = bk.init_grid(padded, periodic=True)
    np.testing.assert_array_equal(info2.adj, ap)
    n = C.c_uint32()
    assert lib.bk_interior_list(longs(nb), None, C.byref(n)) == 0
    lst = np.empty(n.value, np.uint32)
    lib.bk_interior_list(longs(nb), lst.ctypes.data_as(u32p), C.byref(n))
    np.testing.assert_array_equal(lst, O.interior_list(nb))
    assert lib.bk_ghost_pairs(longs(nb), None, C.byref(n)) == 0
    pairs = np.empty(2 * n.value, np.uint32)
    lib.bk_ghost_pairs(longs(nb), pairs.ctypes.data_as(u32p), C.byref(n))
    np.testing.assert_array_equal(pairs, O.ghost_pairs(nb))
    total = int(np.prod(padded))
    assert n.value == total - int(np.prod(nb))


def test_argument_errors_without_gpu(bk):
    from bricklib_amd._lib import BK_ERR_ARG, bk_sweep_desc, longs
    lib = bk.load()
    assert lib.bk_sweep(None, None) == BK_ERR_ARG
    assert b"null descriptor" in lib.bk_last_error()
    d = bk_sweep_desc()
    d.stencil = 42
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_ARG
    d.stencil = 0
    d.dtype = 9
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_ARG
    d.dtype = 0
    coeff = np.ones(7)
    d.h_coeff = coeff.ctypes.data_as(C.POINTER(C.c_double))
    d.nlist = 4
    d.d_adj = d.d_list = d.d_in = d.d_out = 4096
    d.in_step = d.out_step = 512
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_ARG  # in == out aliasing
    assert b"alias" in lib.bk_last_error()
    d.d_out = 8192
    d.in_step = 100
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_ARG
    p = C.c_void_p()
    assert lib.bk_storage_alloc(5, 10, 512, C.byref(p)) == BK_ERR_ARG
    assert lib.bk_storage_alloc(0, -1, 512, C.byref(p)) == BK_ERR_ARG
    assert lib.bk_storage_alloc(0, 1 << 62, 1 << 20, C.byref(p)) == BK_ERR_ARG
    assert lib.bk_init_grid(longs([0, 1, 1]), 0, None, None) == BK_ERR_ARG
    assert lib.bk_sweeps(C.byref(d), None, 0, -1, 0, None) == BK_ERR_ARG
    d.in_step = 512
    every = 12288
    assert lib.bk_sweeps_deep(C.byref(d), every, 8, None, None, 0, 4, 0, 0, None) == BK_ERR_ARG
    assert lib.bk_sweeps_deep(C.byref(d), every, 8, None, None, 0, 4, 9, 0, None) == BK_ERR_ARG
    assert b"ghost zone" in lib.bk_last_error()
    d.stencil = 2  # 25-point star, radius 4: at most 2 sweeps per refresh
    coeff25 = np.ones(25)
    d.h_coeff = coeff25.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.bk_sweeps_deep(C.byref(d), every, 8, None, None, 0, 4, 3, 0, None) == BK_ERR_ARG
    assert lib.bk_sweeps_deep(C.byref(d), every, 8, None, None, 0, -1, 2, 0, None) == BK_ERR_ARG
    assert lib.bk_sweeps_deep(C.byref(d), None, 8, None, None, 0, 4, 2, 0, None) == BK_ERR_ARG


def test_python_mirror_raises(bk):
    from bricklib_amd import BrickError
    with pytest.raises(BrickError):
        bk.BrickStorage.allocate(-1, 512)


@pytest.mark.parametrize("nb", [(1, 1, 1), (3, 2, 4), (5, 5, 5)])
def test_ghost_images_invert_pairs(bk, nb):
    from bricklib_amd import _lib
    from bricklib_amd._lib import longs
    lib = bk.load()
    n = C.c_uint32()
    assert lib.bk_ghost_images(longs(nb), None, None, C.byref(n)) == 0
    S = [x + 2 for x in nb]
    N = S[0] * S[1] * S[2]
    off = np.zeros(N + 1, np.uint32)
    img = np.zeros(n.value, np.uint32)
    assert lib.bk_ghost_images(longs(nb), off.ctypes.data_as(_lib.u32p),
                               img.ctypes.data_as(_lib.u32p), C.byref(n)) == 0
    pairs = np.zeros(2 * n.value, np.uint32)
    lib.bk_ghost_pairs(longs(nb), pairs.ctypes.data_as(_lib.u32p), C.byref(n))
    pairs = pairs.reshape(-1, 2)
    for b in range(N):
        assert sorted(img[off[b]:off[b + 1]].tolist()) == sorted(pairs[pairs[:, 1] == b, 0].tolist())
    assert off[-1] == n.value


def test_fused_sweeps_argument_errors(bk):
    from bricklib_amd._lib import BK_ERR_ARG, BK_ERR_UNSUPPORTED, bk_sweep_desc
    lib = bk.load()
    d = bk_sweep_desc()
    coeff = np.ones(125)
    d.h_coeff = coeff.ctypes.data_as(C.POINTER(C.c_double))
    d.nlist = 4
    d.d_adj = d.d_list = d.d_in = 4096
    d.d_out = 8192
    d.in_step = d.out_step = 512
    assert lib.bk_sweeps_fused(C.byref(d), None, None, None, 0, 2, 0, None) == BK_ERR_ARG
    d.stencil = 3  # the cube has no fused refresh
    assert lib.bk_sweeps_fused(C.byref(d), 64, 128, None, 0, 2, 0, None) == BK_ERR_UNSUPPORTED


def test_lane_shift_argument_errors(bk):
    """lane_shift takes 0 or 1 and is built into the quad star kernels
    only: without a plan it is rejected before any HIP call."""
    from bricklib_amd._lib import BK_ERR_ARG, BK_ERR_UNSUPPORTED, bk_sweep_desc
    lib = bk.load()
    d = bk_sweep_desc()
    coeff = np.ones(7)
    d.h_coeff = coeff.ctypes.data_as(C.POINTER(C.c_double))
    d.nlist = 4
    d.d_adj = d.d_list = d.d_in = 4096
    d.d_out = 8192
    d.in_step = d.out_step = 512
    d.lane_shift = 2
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_ARG
    assert b"lane_shift" in lib.bk_last_error()
    d.lane_shift = 1
    assert lib.bk_sweep(C.byref(d), None) == BK_ERR_UNSUPPORTED
    assert b"quad plans" in lib.bk_last_error()


def test_struct_layouts_match_the_header(bk, tmp_path):
    """The ctypes structures are the C structures of include/bk_api.h:
    sizes and field offsets from a C program compiled against the header,
    and the library's ABI version is the binding's."""
    import subprocess
    from bricklib_amd import _lib
    structs = {"bk_sweep_desc": _lib.bk_sweep_desc, "bk_tune_cand": _lib.bk_tune_cand,
               "bk_serve_config": _lib.bk_serve_config, "bk_batch_req": _lib.bk_batch_req}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bk_api.h"',
             'int main(void) {', 'printf("version %d\\n", BK_API_VERSION);']
    for s, cls in structs.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            *key, val = line.split()
            got[tuple(key)] = int(val)
    assert got[("version",)] == _lib.API_VERSION == bk.load().bk_version()
    for s, cls in structs.items():
        assert got[(s, "size")] == C.sizeof(cls), s
        for f, _ in cls._fields_:
            assert got[(s, f)] == getattr(cls, f).offset, (s, f)
