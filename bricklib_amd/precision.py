"""Precision forms: FP32 bricks with FP32 arithmetic and BF16 storage with
FP32 accumulation, measured against the FP64 sweep on the same GPU.

From the same seeded U[-1,1] input and the same coefficients, run S sweeps
(each with its periodic ghost refresh) in FP64 and in the low-precision
form, then compare the fields in FP64 on the GPU (bk_field_compare):

  max_rel   max |a - r| / max(|r|, rms(r))   (the floor is the reference
            field's RMS, so points where the field crosses zero do not
            dominate; the field decays over the sweeps)
  rms_rel   ||a - r||_2 / ||r||_2
  l2        the FP64 global L2 norm of each field (bk_field_stats)
  underflow True when the FP64 field's RMS is below the smallest normal
            FP32 value: the low-precision form cannot represent it, and
            its relative error says nothing about the arithmetic (the
            125-point cube with L1-normalised random coefficients contracts
            the field by about 0.17 per sweep and reaches ~1e-60 after 100)
"""
from __future__ import annotations

import numpy as np

from . import bricks as B
from ._lib import BK_BF16, BK_F32, BK_F64
from .stencils import Stencil

DTYPE_NAMES = {BK_F64: "f64", BK_F32: "f32", BK_BF16: "bf16"}


def run_sweeps(grid: B.PeriodicGrid, st: Stencil, dtype: int, nsweeps: int, seed: int,
               stream=None, plan="auto"):
    """nsweeps ping-pong sweeps from the seeded input; returns (result field,
    the two storages kept alive)."""
    a = B.Brick(grid.info, grid.field_storage(dtype), 0)
    b = B.Brick(grid.info, grid.field_storage(dtype), 0)
    grid.fill_uniform(a, seed, stream=stream)
    res = B.sweeps(st, a, b, grid, nsweeps, stream=stream, plan=plan)
    return res, (a, b)


def compare(grid: B.PeriodicGrid, a: B.Brick, ref: B.Brick, stream=None) -> dict:
    s = grid.stats(ref, stream)
    rms = float(np.sqrt(s[1] / grid.cells))
    c = grid.compare(a, ref, floor=rms, stream=stream)
    return {"max_abs": float(c["max_abs"]), "max_rel": float(c["max_rel"]),
            "rms_rel": float(c["rms_rel"]), "ref_rms": rms,
            "underflow": bool(rms < float(np.finfo(np.float32).tiny)),
            "l2_ref": float(np.sqrt(s[1])), "l2": float(np.sqrt(grid.stats(a, stream)[1]))}


def study(n: int, stencil, dtypes=(BK_F32, BK_BF16), sweeps=(1, 10, 100), seed: int = 7,
          stream=None) -> dict:
    """Errors of each low-precision form against FP64 after each sweep count
    on an n^3 periodic grid."""
    grid = B.PeriodicGrid((n // 8,) * 3, stream=stream)
    st = Stencil.make(stencil, seed=seed)
    out = {}
    for S in sweeps:
        ref, keep_r = run_sweeps(grid, st, BK_F64, S, seed, stream)
        for dt in dtypes:
            lo, keep = run_sweeps(grid, st, dt, S, seed, stream)
            out[f"{DTYPE_NAMES[dt]}_{S}"] = compare(grid, lo, ref, stream)
            del keep, lo
        del keep_r, ref
    return out
