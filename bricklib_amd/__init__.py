"""bricklib_amd -- MI355X-native brick-layout stencil sweeps.

The product is the HIP library ``libbricks.so`` behind the C-ABI of
``include/bk_api.h`` and the C++ headers beside it.  This package is its
Python host mirror (ctypes), used by the tests, the benchmark and the
batched/serving runtimes.  Importing it does not load the library; the
first call that needs it does, and fails loudly when it is not built.
"""
import os as _os

# HIP gives a process GPU_MAX_HW_QUEUES hardware queues (4 unless set) and
# shares them among its streams beyond that count.  The serving loop runs
# two streams concurrently (plus a capture stream per batch); with four
# queues, whether the two land on separate queues depends on how many
# streams the process already holds (measured: one stream more, and the
# two lanes ran one after the other, p50 latency 10 -> 13.5 ms).  Eight
# queues keep them apart, so a value below 8 (or none) is raised to 8; a
# higher one is kept.  Takes effect only when HIP is initialised after this
# import.
try:
    _queues = int(_os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _queues = 0
if _queues < 8:
    _os.environ["GPU_MAX_HW_QUEUES"] = "8"

from ._lib import (BK_BF16, BK_COND7, BK_COND_REFERENCE, BK_COND_THRESHOLD, BK_CUBE125,
                   BK_F32, BK_F64, BK_LAYOUT_4x8x16, BK_LAYOUT_4x8x16_F88, BK_LAYOUT_8x8x8,
                   BK_STAR7, BK_STAR13, BK_STAR25, BrickError, load)
from .bricks import (Brick, BrickInfo, BrickStorage, DeviceBuffer, Event, PeriodicGrid,
                     Plan, Stream, plan_host, copyFromBrick, copyToBrick, device_count, device_sync, init_grid,
                     movBrickInfo, movBrickStorage, sweep, sweeps, tune, cand_name)
from .stencils import Stencil
from . import precision  # noqa: E402

__all__ = [
    "BK_F64", "BK_F32", "BK_BF16", "BK_STAR7", "BK_STAR13", "BK_STAR25", "BK_CUBE125",
    "BK_COND7", "BK_COND_REFERENCE", "BK_COND_THRESHOLD", "BrickError", "load", "Brick",
    "BrickInfo", "BrickStorage", "DeviceBuffer", "Event", "PeriodicGrid", "Plan", "Stream",
    "plan_host",
    "copyFromBrick", "copyToBrick", "device_count", "device_sync", "init_grid", "movBrickInfo",
    "movBrickStorage", "sweep", "sweeps", "Stencil", "tune", "cand_name", "BK_LAYOUT_8x8x8",
    "BK_LAYOUT_4x8x16", "BK_LAYOUT_4x8x16_F88",
]
