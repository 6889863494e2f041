"""bk_field_stats_noalloc (no device allocation: its pass-1 partials live in
the library's static per-GPU array) against bk_field_stats on the same
fields, through the C-ABI: two calls in a row, the second over a smaller
input, plus a call with the results left in device memory by the first
still in place.  Results must agree to the last bit: both entry points
launch the same two kernels with the same block count, so the FP64 sums are
formed in the same order."""
import ctypes as C

import numpy as np
import pytest

from bricklib_amd import BK_BF16, BK_F32, BK_F64, _lib
from bricklib_amd._lib import check

pytestmark = pytest.mark.gpu


def _vp(x):
    return C.c_void_p(int(x))


def _stats_alloc(lib, brick, grid, nlist, stream):
    out = np.zeros(3)
    check(lib.bk_field_stats(brick.dtype, _vp(brick.ptr), C.c_size_t(brick.step),
                             _vp(grid.d_interior.ptr), C.c_uint32(nlist), C.c_size_t(512),
                             out.ctypes.data_as(_lib.dblp), _vp(stream.handle)))
    return out


def _stats_noalloc(gpu, lib, brick, grid, nlist, d_out, stream):
    check(lib.bk_field_stats_noalloc(brick.dtype, _vp(brick.ptr), C.c_size_t(brick.step),
                                     _vp(grid.d_interior.ptr), C.c_uint32(nlist),
                                     C.c_size_t(512), _vp(d_out.ptr), _vp(stream.handle)))
    stream.sync()
    return d_out.download(np.float64, 3)


@pytest.mark.parametrize("dtype", [BK_F64, BK_F32, BK_BF16], ids=["f64", "f32", "bf16"])
def test_noalloc_stats_match_allocating_entry_point(gpu, dtype):
    lib = gpu.load()
    stream = gpu.Stream()
    d_out = gpu.DeviceBuffer(3 * 8)
    for nb, seed in (((12, 10, 9), 3), ((3, 2, 5), 4)):
        print(f"[case] noalloc stats dtype={dtype} nb={nb}", flush=True)
        g = gpu.PeriodicGrid(nb)
        a = gpu.Brick(g.info, g.field_storage(dtype), 0)
        g.fill_uniform(a, seed)
        gpu.device_sync()
        n = g.interior.size
        for nlist in (n, max(1, n // 3)):
            want = _stats_alloc(lib, a, g, nlist, stream)
            got = _stats_noalloc(gpu, lib, a, g, nlist, d_out, stream)
            np.testing.assert_array_equal(got, want)
            assert want[1] > 0.0
        del a, g


def test_noalloc_stats_rejects_bad_arguments(gpu):
    lib = gpu.load()
    d_out = gpu.DeviceBuffer(3 * 8)
    assert lib.bk_field_stats_noalloc(7, None, C.c_size_t(512), None, C.c_uint32(0),
                                      C.c_size_t(512), _vp(d_out.ptr), None) != 0
    assert lib.bk_field_stats_noalloc(BK_F64, None, C.c_size_t(512), None, C.c_uint32(0),
                                      C.c_size_t(512), None, None) != 0
    assert lib.bk_field_stats_noalloc(BK_F64, None, C.c_size_t(256), None, C.c_uint32(0),
                                      C.c_size_t(512), _vp(d_out.ptr), None) != 0
