"""ctypes binding of libcnq.so (the C ABI in include/cnq.h).

The library is loaded from this package's directory; there is no fallback:
without it every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcnq.so"
NONE = 0xFFFFFFFF

OK, NO_DEVICE, INVALID_ARG, CAPACITY, OOM, HIP_ERROR = range(6)


class CnqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cnq error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("hierarchy", C.c_uint32),
        ("prefilter", C.c_float),
        ("wide_extent", C.c_float),
        ("light_cone_cap", C.c_uint32),
        ("wide_cone_cap", C.c_uint32),
        ("profile_stage", C.c_uint32),
        ("object_cov", C.c_uint32),
        ("nearest_first", C.c_uint32),
    ]


class SkyOptions(C.Structure):
    _fields_ = [
        ("max_combine", C.c_uint32),
        ("horizon_cull", C.c_uint32),
        ("prefilter_scale", C.c_float),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("tris", C.c_void_p), ("n_tris", C.c_uint64),
        ("ranges", C.c_void_p), ("n_objects", C.c_uint32),
        ("blas", C.c_void_p), ("n_blas", C.c_uint32),
        ("blas_items", C.c_void_p), ("n_blas_items", C.c_uint64),
        ("blas_cov", C.c_void_p),
        ("nodes", C.c_void_p), ("n_nodes", C.c_uint32),
        ("items", C.c_void_p), ("n_items", C.c_uint32),
        ("xforms", C.c_void_p), ("n_xforms", C.c_uint32),
        ("n_blas_cov", C.c_uint32),
    ]


class Placement(C.Structure):
    _fields_ = [("geometry", C.c_uint32), ("rotation", C.c_float * 9), ("scale", C.c_float),
                ("translation", C.c_float * 3), ("clear", C.c_float)]


class View(C.Structure):
    _fields_ = [("camera", C.c_float * 20), ("width", C.c_uint32), ("height", C.c_uint32),
                ("light_first", C.c_uint32), ("light_count", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("persistent_blocks", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("ray_stack", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("pixel_capacity", C.c_uint64), ("query_capacity", C.c_uint64), ("max_lights", C.c_uint32),
                ("max_view_lights", C.c_uint32), ("max_views_per_launch", C.c_uint32), ("slots", C.c_uint32),
                ("tuning", Tuning), ("launch_queries", C.c_uint64), ("express_queries", C.c_uint64)]


class MoverTree(C.Structure):
    _fields_ = [("fixed_root", C.c_uint32 * 12), ("first_node", C.c_uint32), ("nodes", C.c_uint32),
                ("first_item", C.c_uint32), ("n_objects", C.c_uint32), ("objects", C.c_void_p)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32), ("n_lights", C.c_uint32),
                ("launch", C.c_uint64), ("queued_ms", C.c_double), ("latency_ms", C.c_double),
                ("admitted_at_ms", C.c_double), ("done_at_ms", C.c_double)]


REQ_QUEUED, REQ_ADMITTED, REQ_DONE, REQ_RELEASED = range(4)


class VcacheSurfaces(C.Structure):
    _fields_ = [("tri_chart", C.c_void_p), ("tri_uv", C.c_void_p), ("normals", C.c_void_p),
                ("n_charts", C.c_uint32), ("chart_size", C.c_void_p), ("chart_grid", C.c_void_p)]


class VcacheConfig(C.Structure):
    _fields_ = [("pages", C.c_uint32), ("append_pages", C.c_uint32), ("candidates", C.c_uint32),
                ("base_texel", C.c_float), ("light_ratio", C.c_uint32), ("visible_frames", C.c_uint32)]


class VcacheShadeParams(C.Structure):
    _fields_ = [("pixels_per_texel", C.c_float), ("refine_levels", C.c_uint32), ("request_tile", C.c_uint32),
                ("pixel_angle", C.c_float)]


class MovedBox(C.Structure):
    _fields_ = [("object", C.c_uint32), ("old_lo", C.c_float * 3), ("old_hi", C.c_float * 3),
                ("new_lo", C.c_float * 3), ("new_hi", C.c_float * 3)]


# Every symbol include/cnq.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "cnq_options_default": (None, [C.POINTER(Options)]),
    "cnq_last_error": (C.c_char_p, []),
    "cnq_device_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "cnq_device_destroy": (None, [_P]),
    "cnq_flat_build": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(_P)]),
    "cnq_flat_sizes": (C.c_int, [_P, _P]),
    "cnq_flat_copy": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "cnq_flat_destroy": (None, [_P]),
    "cnq_flat_build_instanced": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint64, C.c_int, C.POINTER(_P)]),
    "cnq_flat_build_instanced_movers": (
        C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint32, C.c_int, C.POINTER(_P)]),
    "cnq_flat_mover_tree": (C.c_int, [_P, C.POINTER(MoverTree)]),
    "cnq_scene_set_movers": (C.c_int, [_P, C.POINTER(MoverTree), _P, C.c_uint32, _P]),
    "cnq_scene_set_transform": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "cnq_scene_apply_moves": (C.c_int, [_P, _P]),
    "cnq_scene_set_pose": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, _P]),
    "cnq_flat_instanced_sizes": (C.c_int, [_P, _P]),
    "cnq_flat_copy_instanced": (C.c_int, [_P, _P, _P]),
    "cnq_scene_create": (C.c_int, [_P, C.POINTER(SceneDesc), C.POINTER(_P)]),
    "cnq_scene_destroy": (None, [_P]),
    "cnq_scene_bytes": (C.c_uint64, [_P]),
    "cnq_irradiance": (C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint32, C.POINTER(Options), _P, _P, _P, _P]),
    "cnq_irradiance_device": (
        C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint32, C.POINTER(Options), _P, _P, _P, _P, _P]),
    "cnq_sky_options_default": (None, [C.POINTER(SkyOptions)]),
    "cnq_sky_irradiance": (C.c_int, [_P, _P, C.c_uint64, C.POINTER(Options), C.POINTER(SkyOptions), _P, _P, _P]),
    "cnq_sky_irradiance_device": (
        C.c_int, [_P, _P, C.c_uint64, C.POINTER(Options), C.POINTER(SkyOptions), _P, _P, _P, _P]),
    "cnq_sky_view_irradiance_device": (
        C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Options), C.POINTER(SkyOptions), _P, _P, _P,
                  _P]),
    "cnq_sky_irradiance_device_noalloc": (
        C.c_int, [_P, _P, C.c_uint64, C.POINTER(Options), C.POINTER(SkyOptions), _P, _P, _P]),
    "cnq_view_irradiance_device": (
        C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(Options), C.c_uint32, _P, _P, _P, _P]),
    "cnq_tuning_default": (None, [C.POINTER(Tuning)]),
    "cnq_batch_create": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, C.POINTER(_P)]),
    "cnq_batch_destroy": (None, [_P]),
    "cnq_batch_queries": (C.c_uint64, [_P]),
    "cnq_batch_run": (C.c_int, [_P, C.POINTER(Options), C.POINTER(Tuning), _P]),
    "cnq_batch_capture": (C.c_int, [_P, C.POINTER(Options), C.POINTER(Tuning)]),
    "cnq_batch_replay": (C.c_int, [_P, _P]),
    "cnq_batch_autotune": (C.c_int, [_P, C.POINTER(Options), _P, C.POINTER(Tuning), C.POINTER(C.c_float)]),
    "cnq_batch_read": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P]),
    "cnq_batch_set_lights": (C.c_int, [_P, _P, C.c_uint32]),
    "cnq_primary_device": (
        C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(Options), _P, _P, _P]),
    "cnq_render_view": (
        C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(Options), _P, _P, _P, _P]),
    "cnq_scene_refit_device": (C.c_int, [_P, _P, _P]),
    "cnq_scene_refit": (C.c_int, [_P, _P]),
    "cnq_scene_read_blas": (C.c_int, [_P, _P, _P]),
    "cnq_scene_counters": (C.c_int, [_P, C.c_int, _P, _P]),
    "cnq_scene_set_count_work": (C.c_int, [_P, C.c_int]),
    "cnq_scene_set_query_flags": (C.c_int, [_P, _P, C.c_uint64]),
    "cnq_scene_set_normals": (C.c_int, [_P, _P, C.c_uint64]),
    "cnq_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "cnq_server_create": (C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(Options), _P, C.c_uint32, C.POINTER(_P)]),
    "cnq_server_destroy": (None, [_P]),
    "cnq_server_submit_view": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(C.c_uint64)]),
    "cnq_server_submit_light": (C.c_int, [_P, C.c_uint32, _P, C.POINTER(C.c_uint64)]),
    "cnq_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(C.c_int)]),
    "cnq_server_info": (C.c_int, [_P, C.c_uint64, C.POINTER(RequestInfo)]),
    "cnq_server_read": (C.c_int, [_P, C.c_uint64, _P, _P, _P, _P]),
    "cnq_server_release": (C.c_int, [_P, C.c_uint64]),
    "cnq_server_counts": (C.c_int, [_P, _P]),
    "cnq_block_planes": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "cnq_cache_create": (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P, _P, C.POINTER(Options), C.POINTER(_P)]),
    "cnq_cache_destroy": (None, [_P]),
    "cnq_cache_update": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "cnq_cache_read": (C.c_int, [_P, _P, _P, _P]),
    "cnq_scene_last_moves": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "cnq_vcache_config_default": (None, [C.POINTER(VcacheConfig)]),
    "cnq_vcache_create": (C.c_int, [_P, C.POINTER(VcacheSurfaces), _P, C.c_uint32, C.POINTER(VcacheConfig),
                                    C.POINTER(Options), C.POINTER(_P)]),
    "cnq_vcache_destroy": (None, [_P]),
    "cnq_vcache_grid": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "cnq_vcache_request": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "cnq_vcache_evict": (C.c_int, [_P, _P, C.c_uint32]),
    "cnq_vcache_set_moved": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "cnq_vcache_frame": (C.c_int, [_P, _P, _P]),
    "cnq_vcache_read": (C.c_int, [_P, _P, C.c_uint32, _P, _P]),
    "cnq_vcache_state": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "cnq_vcache_shade": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, C.POINTER(VcacheShadeParams), _P, _P]),
    "cnq_vcache_verify": (C.c_int, [_P, _P, _P]),
    "cnq_vcache_set_lights": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, _P]),
    "cnq_vcache_plan": (C.c_int, [_P, C.c_uint32, _P]),
    "cnq_vcache_set_moving_lights": (C.c_int, [_P, _P, C.c_uint32, C.c_float]),
    "cnq_vcache_move_light": (C.c_int, [_P, C.c_uint32, _P]),
    "cnq_vcache_ledger_stats": (C.c_int, [_P, _P]),
    "cnq_vcache_ledger_state": (C.c_int, [_P, _P, _P]),
    "cnq_vcache_verify_ledger": (C.c_int, [_P, C.c_uint32, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m continuum_amd.build`")
        # PyTorch ships its own HIP runtime under the same soname. Loading it
        # first makes the process hold one runtime, shared with the library,
        # so torch's device buffers and streams are valid here (loaded the
        # other way round, torch then finds no device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise CnqError(rc, lib().cnq_last_error().decode(errors="replace"))


def default_sky_options(**kw) -> SkyOptions:
    o = SkyOptions()
    lib().cnq_sky_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_options(**kw) -> Options:
    o = Options()
    lib().cnq_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
