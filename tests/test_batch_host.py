"""Host side of the batched path (no GPU): the packed request records that
bk_batch_set reads must match the C layout of bk_batch_req field by field."""
import ctypes as C

import numpy as np

from bricklib_amd._lib import bk_batch_req, dblp
from bricklib_amd.batch import REQ_DTYPE


def test_record_dtype_matches_struct():
    assert REQ_DTYPE.itemsize == C.sizeof(bk_batch_req)
    for name, _ in bk_batch_req._fields_:
        assert REQ_DTYPE.fields[name][1] == getattr(bk_batch_req, name).offset


def test_packed_records_round_trip():
    co = [np.arange(7, dtype=np.float64) + q for q in range(3)]
    recs = []
    for q in range(3):
        r = bk_batch_req()
        r.stencil, r.dtype = q, 1
        r.d_adj, r.d_list, r.nlist = 0x1000 + q, 0x2000 + q, 10 + q
        r.d_f0, r.d_f1, r.step = 0x3000 + q, 0x4000 + q, 512
        r.h_coeff = co[q].ctypes.data_as(dblp)
        r.theta, r.cond_mode = 0.25 * q, 1
        recs.append(bytes(r))
    arr = np.frombuffer(bytearray(b"".join(recs)), dtype=REQ_DTYPE)
    arr["ticks"] = [5, 6, 7]
    arr["parity"] = [1, 0, 1]
    p = arr.ctypes.data_as(C.POINTER(bk_batch_req))
    for q in range(3):
        assert (p[q].stencil, p[q].dtype, p[q].nlist, p[q].step) == (q, 1, 10 + q, 512)
        assert (p[q].d_adj, p[q].d_list, p[q].d_f0, p[q].d_f1) == (
            0x1000 + q, 0x2000 + q, 0x3000 + q, 0x4000 + q)
        assert p[q].h_coeff[6] == 6 + q and p[q].theta == 0.25 * q
        assert (p[q].ticks, p[q].parity) == (5 + q, (q + 1) % 2)


def test_serving_lanes_pair_cubes_with_the_other_dtypes_stars():
    from bricklib_amd._lib import (BK_BF16, BK_COND7, BK_CUBE125, BK_F32, BK_F64, BK_STAR7,
                                   BK_STAR25)
    from bricklib_amd.serve import lane_of
    assert lane_of(BK_CUBE125, BK_F64, 2) == 0
    assert lane_of(BK_STAR7, BK_F32, 2) == 0 and lane_of(BK_STAR25, BK_BF16, 2) == 0
    assert lane_of(BK_CUBE125, BK_F32, 2) == 1
    assert lane_of(BK_STAR7, BK_F64, 2) == 1 and lane_of(BK_COND7, BK_F64, 2) == 1
    assert all(lane_of(k, d, 1) == 0 for k in range(5) for d in (BK_F64, BK_F32, BK_BF16))


def test_import_raises_the_hardware_queue_count_to_eight():
    """The serving lanes need their streams on separate hardware queues;
    importing the package raises GPU_MAX_HW_QUEUES to 8 when it is unset or
    lower, and keeps a higher value."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = "import os, bricklib_amd; print(os.environ['GPU_MAX_HW_QUEUES'])"
    env = {k: v for k, v in os.environ.items() if k != "GPU_MAX_HW_QUEUES"}
    env["PYTHONPATH"] = root
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True).stdout.strip()
    assert out == "8"
    for given, want in (("4", "8"), ("16", "16"), ("bogus", "8")):
        env["GPU_MAX_HW_QUEUES"] = given
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                             text=True, check=True).stdout.strip()
        assert out == want, (given, out)
