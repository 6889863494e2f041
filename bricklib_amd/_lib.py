"""ctypes binding of libbricks.so (the C-ABI declared in include/bk_api.h).

Loading fails loudly: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbricks.so")
API_VERSION = 3  # BK_API_VERSION of include/bk_api.h

BK_OK = 0
BK_ERR_ARG = 1
BK_ERR_NOMEM = 2
BK_ERR_HIP = 3
BK_ERR_UNSUPPORTED = 4
BK_ERR_NODEV = 5

BK_F64, BK_F32, BK_BF16 = 0, 1, 2
BK_STAR7, BK_STAR13, BK_STAR25, BK_CUBE125, BK_COND7 = 0, 1, 2, 3, 4
BK_COND_REFERENCE, BK_COND_THRESHOLD = 0, 1
BK_PLAN_COLUMN, BK_PLAN_QUAD = 1, 4
BK_LAYOUT_8x8x8, BK_LAYOUT_4x8x16, BK_LAYOUT_4x8x16_F88 = 0, 1, 2
BK_TUNE_ALL_LAYOUTS = 0x7
BK_TUNE_BATCH_SCHEDULES = 0x100

u32p = C.POINTER(C.c_uint32)
dblp = C.POINTER(C.c_double)
longp = C.POINTER(C.c_long)
vp = C.c_void_p


class bk_sweep_desc(C.Structure):
    _fields_ = [
        ("stencil", C.c_int),
        ("dtype", C.c_int),
        ("d_adj", vp),
        ("d_list", vp),
        ("nlist", C.c_uint32),
        ("d_in", vp),
        ("in_step", C.c_size_t),
        ("d_out", vp),
        ("out_step", C.c_size_t),
        ("h_coeff", dblp),
        ("theta", C.c_double),
        ("cond_mode", C.c_int),
        ("plan", vp),
        ("layout", C.c_int),
        ("wg_bricks", C.c_int),
        ("lane_shift", C.c_int),
    ]


class bk_serve_config(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_double),
        ("chunk", C.c_uint32),
        ("max_active", C.c_uint32),
        ("lanes", C.c_int),
        ("use_graph", C.c_int),
        ("ghost", C.c_int),
        ("pipeline", C.c_int),
        ("tune", C.c_int),
        ("slab_bytes", C.c_size_t),
        ("star_plan", C.c_int),
    ]


class bk_serve_req(C.Structure):
    _fields_ = [
        ("nb", C.c_long * 3),
        ("stencil", C.c_int),
        ("dtype", C.c_int),
        ("steps", C.c_uint32),
        ("seed", C.c_uint64),
        ("h_coeff", dblp),
        ("theta", C.c_double),
        ("cond_mode", C.c_int),
        ("arrival", C.c_double),
        ("keep", C.c_int),
    ]


class bk_serve_info(C.Structure):
    _fields_ = [
        ("state", C.c_int),
        ("arrival", C.c_double),
        ("admitted", C.c_double),
        ("finished", C.c_double),
        ("sweeps_done", C.c_uint32),
        ("field_bytes", C.c_size_t),
        ("d_result", vp),
        ("plan_shape", C.c_int),
        ("kcol", C.c_uint32),
    ]


class bk_serve_stats(C.Structure):
    _fields_ = [
        ("requests", C.c_uint32),
        ("p50_s", C.c_double),
        ("p99_s", C.c_double),
        ("mean_s", C.c_double),
        ("elapsed_s", C.c_double),
        ("stencil_updates", C.c_double),
        ("gstencil_s", C.c_double),
        ("peak_hbm_bytes", C.c_double),
        ("arena_reserved_bytes", C.c_double),
        ("chunks", C.c_uint64),
        ("tuned_classes", C.c_uint32),
    ]


class bk_tune_cand(C.Structure):
    _fields_ = [
        ("layout", C.c_int),
        ("plan_shape", C.c_int),
        ("kcol", C.c_uint32),
        ("wg_bricks", C.c_int),
        ("lane_shift", C.c_int),
        ("ms", C.c_float),
    ]


class bk_decomp_region(C.Structure):
    _fields_ = [
        ("neighbor", C.c_int32 * 3),
        ("ghost_pos", C.c_uint32),
        ("skin_pos", C.c_uint32),
        ("len", C.c_uint32),
        ("skin_st", C.c_uint32),
        ("skin_ed", C.c_uint32),
        ("first_pad", C.c_uint32),
        ("last_pad", C.c_uint32),
    ]


class bk_batch_req(C.Structure):
    _fields_ = [
        ("stencil", C.c_int),
        ("dtype", C.c_int),
        ("d_adj", vp),
        ("d_list", vp),
        ("nlist", C.c_uint32),
        ("plan", vp),
        ("d_pairs", vp),
        ("npairs", C.c_uint32),
        ("d_f0", vp),
        ("d_f1", vp),
        ("step", C.c_size_t),
        ("h_coeff", dblp),
        ("theta", C.c_double),
        ("cond_mode", C.c_int),
        ("parity", C.c_uint32),
        ("ticks", C.c_uint32),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "bk_version": (C.c_int, []),
    "bk_last_error": (C.c_char_p, []),
    "bk_stencil_ncoeff": (C.c_int, [C.c_int]),
    "bk_stencil_radius": (C.c_int, [C.c_int]),
    "bk_dtype_size": (C.c_int, [C.c_int]),
    "bk_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bk_set_device": (C.c_int, [C.c_int]),
    "bk_device_sync": (C.c_int, []),
    "bk_pointer_is_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "bk_storage_alloc": (C.c_int, [C.c_int, C.c_long, C.c_size_t, C.POINTER(vp)]),
    "bk_free": (C.c_int, [vp]),
    "bk_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
    "bk_alloc_async": (C.c_int, [C.c_size_t, C.POINTER(vp), vp]),
    "bk_free_async": (C.c_int, [vp, vp]),
    "bk_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "bk_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "bk_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "bk_zero": (C.c_int, [vp, C.c_size_t, vp]),
    "bk_stream_copy": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "bk_valu_probe": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint32, vp, C.c_size_t,
                                C.POINTER(C.c_double), vp]),
    "bk_stream_create": (C.c_int, [C.POINTER(vp)]),
    "bk_stream_destroy": (C.c_int, [vp]),
    "bk_stream_sync": (C.c_int, [vp]),
    "bk_event_create": (C.c_int, [C.POINTER(vp)]),
    "bk_event_destroy": (C.c_int, [vp]),
    "bk_event_record": (C.c_int, [vp, vp]),
    "bk_event_sync": (C.c_int, [vp]),
    "bk_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "bk_init_grid": (C.c_int, [longp, C.c_int, u32p, u32p]),
    "bk_layout_extents": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "bk_fill_uniform_layout": (C.c_int, [C.c_int, C.c_int, vp, C.c_size_t, longp, C.c_int,
                                         C.c_uint64, C.c_double, C.c_double, vp]),
    "bk_tiled_list": (C.c_int, [longp, C.c_int, C.c_uint32, C.c_uint32, u32p, u32p]),
    "bk_tune_sweep": (C.c_int, [C.c_int, C.c_int, dblp, C.c_double, C.c_int, longp, C.c_int,
                                C.c_int, C.c_int, C.c_double, C.c_uint, C.POINTER(bk_tune_cand),
                                C.c_uint32, u32p, vp]),
    "bk_tune_lookup": (C.c_int, [C.c_int, C.c_int, longp, C.c_int, C.c_uint,
                                 C.POINTER(bk_tune_cand)]),
    "bk_server_create": (C.c_int, [C.POINTER(bk_serve_config), C.POINTER(vp)]),
    "bk_server_destroy": (C.c_int, [vp]),
    "bk_server_submit": (C.c_int, [vp, C.POINTER(bk_serve_req), u32p]),
    "bk_server_prewarm": (C.c_int, [vp, C.POINTER(bk_serve_req), C.c_uint32]),
    "bk_server_run": (C.c_int, [vp, C.c_int]),
    "bk_server_start": (C.c_int, [vp]),
    "bk_server_stop": (C.c_int, [vp]),
    "bk_server_wait": (C.c_int, [vp, C.c_uint32, C.c_double]),
    "bk_server_info": (C.c_int, [vp, C.c_uint32, C.POINTER(bk_serve_info)]),
    "bk_server_result": (C.c_int, [vp, C.c_uint32, vp, C.c_size_t]),
    "bk_server_release": (C.c_int, [vp, C.c_uint32]),
    "bk_server_stats": (C.c_int, [vp, C.POINTER(bk_serve_stats)]),
    "bk_server_reset_stats": (C.c_int, [vp]),
    "bk_ghost_pairs": (C.c_int, [longp, u32p, u32p]),
    "bk_interior_list": (C.c_int, [longp, u32p, u32p]),
    "bk_sweep": (C.c_int, [C.POINTER(bk_sweep_desc), vp]),
    "bk_ghost_refresh": (C.c_int, [C.c_int, vp, C.c_size_t, C.c_size_t, vp, C.c_uint32, vp]),
    "bk_sweeps_deep": (C.c_int, [C.POINTER(bk_sweep_desc), vp, C.c_uint32, vp, vp, C.c_uint32,
                                 C.c_int, C.c_int, C.c_int, vp]),
    "bk_sweeps": (C.c_int, [C.POINTER(bk_sweep_desc), vp, C.c_uint32, C.c_int, C.c_int, vp]),
    "bk_sweeps_fused": (C.c_int, [C.POINTER(bk_sweep_desc), vp, vp, vp, C.c_uint32, C.c_int,
                                  C.c_int, vp]),
    "bk_ghost_images": (C.c_int, [longp, u32p, u32p, C.POINTER(C.c_uint32)]),
    "bk_decomp_layout": (C.c_int, [longp, C.c_int, C.c_size_t, C.c_size_t, u32p, u32p,
                                   C.POINTER(C.c_uint32), u32p, C.POINTER(bk_decomp_region),
                                   C.POINTER(C.c_uint32)]),
    "bk_decomp_layout_skin": (C.c_int, [longp, C.c_int, C.c_size_t, C.c_size_t,
                                        C.POINTER(C.c_int32), C.c_int, u32p, u32p,
                                        C.POINTER(C.c_uint32), u32p, C.POINTER(bk_decomp_region),
                                        C.POINTER(C.c_uint32), u32p]),
    "bk_graph_cache_clear": (C.c_int, []),
    "bk_fill_uniform": (C.c_int, [C.c_int, vp, C.c_size_t, longp, C.c_uint64, C.c_double,
                                  C.c_double, vp]),
    "bk_arena_create": (C.c_int, [C.c_size_t, C.c_size_t, C.POINTER(vp)]),
    "bk_arena_destroy": (C.c_int, [vp]),
    "bk_arena_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    "bk_arena_free": (C.c_int, [vp, vp]),
    "bk_arena_stats": (C.c_int, [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_size_t), C.POINTER(C.c_uint32)]),
    "bk_fill_uniform_grid": (C.c_int, [C.c_int, vp, C.c_size_t, longp, C.c_int, C.c_uint64,
                                       C.c_double, C.c_double, vp]),
    "bk_field_stats": (C.c_int, [C.c_int, vp, C.c_size_t, vp, C.c_uint32, C.c_size_t, dblp, vp]),
    "bk_field_stats_noalloc": (C.c_int, [C.c_int, vp, C.c_size_t, vp, C.c_uint32, C.c_size_t,
                                         vp, vp]),
    "bk_field_compare": (C.c_int, [C.c_int, vp, C.c_size_t, C.c_int, vp, C.c_size_t, vp,
                                   C.c_uint32, C.c_size_t, C.c_double, C.c_double, C.c_double,
                                   dblp, vp]),
    "bk_plan_create": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, C.c_int, C.c_uint32,
                                 C.POINTER(vp)]),
    "bk_plan_destroy": (C.c_int, [vp]),
    "bk_plan_query": (C.c_int, [vp, u32p, u32p, C.POINTER(C.c_int), u32p]),
    "bk_plan_download": (C.c_int, [vp, u32p, u32p, u32p, u32p]),
    "bk_plan_build_host": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, C.c_int, C.c_uint32,
                                     u32p, u32p, u32p, u32p, u32p, u32p]),
    "bk_batch_create": (C.c_int, [C.c_uint32, C.POINTER(vp)]),
    "bk_batch_destroy": (C.c_int, [vp]),
    "bk_batch_set": (C.c_int, [vp, C.POINTER(bk_batch_req), C.c_uint32, vp]),
    "bk_batch_run": (C.c_int, [vp, C.c_uint32, C.c_int, vp]),
    "bk_copy_to_brick": (C.c_int, [C.c_int, longp, longp, longp, vp, vp, vp, C.c_size_t, vp]),
    "bk_copy_from_brick": (C.c_int, [C.c_int, longp, longp, longp, vp, vp, vp, C.c_size_t, vp]),
}

_lib = None


def load() -> C.CDLL:
    """Load libbricks.so, raising ImportError with the build command when it
    is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m bricklib_amd.build` "
            "(the HIP library is required: there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    # the structures above follow one version of include/bk_api.h
    if lib.bk_version() != API_VERSION:
        raise ImportError(
            f"{LIB_PATH} implements C-ABI version {lib.bk_version()}, this binding expects "
            f"{API_VERSION}; rebuild with `python -m bricklib_amd.build`")
    _lib = lib
    return lib


class BrickError(RuntimeError):
    """A libbricks call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[bk status {code}] {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != BK_OK:
        msg = load().bk_last_error()
        raise BrickError(rc, msg.decode() if msg else "")


def longs(v) -> C.Array:
    return (C.c_long * 3)(*[int(x) for x in v])
