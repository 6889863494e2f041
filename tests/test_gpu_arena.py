"""The field arena (bk_arena_*) on the GPU: blocks do not overlap, freed
neighbours merge, capacity and foreign pointers are rejected, and data
written through one block survives the churn of others."""
import ctypes as C

import numpy as np
import pytest

from bricklib_amd import BrickError
from bricklib_amd._lib import load
from bricklib_amd.batch import Arena

pytestmark = pytest.mark.gpu

MB = 1 << 20


def upload(ptr, arr):
    from bricklib_amd.bricks import _vp
    assert load().bk_memcpy_h2d(_vp(ptr), arr.ctypes.data_as(C.c_void_p),
                                C.c_size_t(arr.nbytes), None) == 0


def download(ptr, n):
    from bricklib_amd.bricks import _vp
    out = np.empty(n, np.uint32)
    assert load().bk_memcpy_d2h(out.ctypes.data_as(C.c_void_p), _vp(ptr),
                                C.c_size_t(out.nbytes), None) == 0
    return out


def test_blocks_disjoint_and_merge(gpu):
    a = Arena(slab_bytes=64 * MB)
    blocks = [a.alloc(s) for s in (3 * MB, 5 * MB + 7, 1, 8 * MB)]
    spans = sorted((b.ptr, b.ptr + b.nbytes) for b in blocks)
    for (s0, e0), (s1, _) in zip(spans, spans[1:]):
        assert e0 <= s1
    assert all(b.ptr % 2048 == 0 for b in blocks)
    st = a.stats()
    assert st["reserved"] == 64 * MB and st["live_blocks"] == 4
    # free the middle two, then the first: one free run at the front merges
    for b in (blocks[1], blocks[2], blocks[0]):
        b.free()
    big = a.alloc(8 * MB + 2048)  # fits only in the merged run (or the tail)
    assert a.stats()["live_blocks"] == 2
    big.free()
    blocks[3].free()
    st = a.stats()
    assert st["used"] == 0 and st["largest_free"] == 64 * MB and st["live_blocks"] == 0


def test_growth_capacity_and_foreign_pointer(gpu):
    a = Arena(slab_bytes=16 * MB, capacity=40 * MB)
    x = a.alloc(10 * MB)
    y = a.alloc(10 * MB)  # second slab
    assert a.stats()["reserved"] == 32 * MB
    with pytest.raises(BrickError):
        a.alloc(10 * MB)  # a third slab would pass the capacity
    from bricklib_amd.bricks import _vp
    assert load().bk_arena_free(_vp(a.handle), _vp(x.ptr + 2048)) == 1
    x.free()
    z = a.alloc(10 * MB)  # reuses x's space
    assert a.stats()["reserved"] == 32 * MB
    del y, z


def test_data_survives_churn(gpu):
    a = Arena(slab_bytes=32 * MB)
    rng = np.random.default_rng(3)
    keep = a.alloc(4 * MB)
    pat = rng.integers(0, 2**32, size=MB, dtype=np.uint32)
    upload(keep.ptr, pat)
    live = []
    for it in range(200):
        if live and rng.random() < 0.5:
            live.pop(int(rng.integers(len(live)))).free()
        else:
            b = a.alloc(int(rng.integers(1, 3 * MB)))
            upload(b.ptr, np.full(b.nbytes // 4, it, np.uint32))
            live.append(b)
    np.testing.assert_array_equal(download(keep.ptr, MB), pat)
    spans = sorted((b.ptr, b.ptr + b.nbytes) for b in live + [keep])
    for (s0, e0), (s1, _) in zip(spans, spans[1:]):
        assert e0 <= s1
