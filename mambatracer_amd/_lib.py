"""ctypes binding of libmambart.so (the C ABI declared in include/mambart.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmambart.so")

MT_OK = 0
ERRORS = {
    -1: "MT_ERR_ARG",
    -2: "MT_ERR_HIP",
    -3: "MT_ERR_ALLOC",
    -4: "MT_ERR_DEPTH",
    -5: "MT_ERR_CAPACITY",
    -6: "MT_ERR_STATE",
}
MT_TRACE_PACKED = 0
MT_TRACE_REFERENCE = 1
MT_TRACE_PACKED_NO_STACK = 2
MT_TRACE_PACKED_STACK = 3
MT_TRACE_PACKED_REFILL = 4
MT_TRACE_PACKED_SORTED = 5
MT_TRACE_PACKED_POOL = 6
MT_TRACE_PACKED_BINNED = 7
MT_BATCH_TRACE_SPLIT = 8  # batch engine only: pools for continuing paths, one ray per lane for new ones
MT_BSDF_WORDS = 20


class MambartError(RuntimeError):
    def __init__(self, code: int, func: str, msg: str):
        super().__init__(f"{func} -> {ERRORS.get(code, code)}: {msg}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("normals", C.c_void_p),
        ("uvs", C.c_void_p),
        ("faces", C.c_void_p),
        ("nodes", C.c_void_p),
        ("bounds", C.c_void_p),
        ("num_tris", C.c_int32),
        ("num_points", C.c_int32),
        ("max_depth", C.c_int32),
    ]


class ObjCounts(C.Structure):
    _fields_ = [("num_points", C.c_int32), ("num_normals", C.c_int32), ("num_uvs", C.c_int32),
                ("num_tris", C.c_int32), ("num_groups", C.c_int32)]


class Payload(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("elem_bytes", C.c_int32)]


class RenderDesc(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("look", C.c_float * 3),
        ("up", C.c_float * 3),
        ("vfov_deg", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp", C.c_int32),
        ("max_bounces", C.c_int32),
        ("rr_start", C.c_int32),
        ("seed", C.c_uint32),
        ("mode", C.c_int32),
        ("flags", C.c_int32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("extension_rays", C.c_int64),
        ("shadow_rays", C.c_int64),
        ("paths", C.c_int64),
        ("bounce_counts", C.c_int32 * 64),
        ("closest_launches", C.c_int32),
        ("any_launches", C.c_int32),
        ("closest_ms", C.c_float),
        ("any_ms", C.c_float),
        ("closest_ilp", C.c_int64 * 3),
        ("any_ilp", C.c_int64 * 3),
    ]


MT_RENDER_PICKED_LIGHT = 1
MT_RENDER_PROFILE = 2
MT_RENDER_TRAVERSAL_STATS = 4
MT_RENDER_LIGHTGRID_ENV = 8
MT_RENDER_LIGHTGRID_TILE_PDF = 16
MT_RENDER_TAIL_SHIFT = 8


def MT_RENDER_TAIL_FROM(b: int) -> int:
    """Render flag: bounces b and later in one launch (mambart.h)."""
    return int(b) << MT_RENDER_TAIL_SHIFT


class JobDesc(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("materials", C.c_void_p),
        ("env", C.c_void_p),
        ("light_faces", C.c_void_p),
        ("d_num_lights", C.c_void_p),
        ("render", RenderDesc),
        ("image", C.c_void_p),
        ("tag", C.c_int64),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("extension_rays", C.c_int64),
        ("shadow_rays", C.c_int64),
        ("jobs_submitted", C.c_int64),
        ("jobs_admitted", C.c_int64),
        ("jobs_completed", C.c_int64),
        ("jobs_cancelled", C.c_int64),
        ("active_jobs", C.c_int32),
        ("pending_jobs", C.c_int32),
        ("arena_used", C.c_uint64),
        ("arena_peak", C.c_uint64),
    ]


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
U32 = C.c_uint32
U64 = C.c_uint64
F32 = C.c_float
PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int32 status unless listed in _RESTYPES)
SIGNATURES = {
    "mt_last_error": [C.c_char_p, I32],
    "mt_version": [],
    "mt_ctx_create": [I32, P, I32, PP],
    "mt_ctx_destroy": [P],
    "mt_ctx_sync": [P],
    "mt_buf_alloc": [P, U64, PP],
    "mt_buf_free": [P, P],
    "mt_buf_upload": [P, P, P, U64],
    "mt_buf_download": [P, P, P, U64],
    "mt_buf_fill_i32": [P, P, I32, U64],
    "mt_bvh_build": [I32, P, P, P, P, P, I32],
    "mt_gen_soup": [U64, I32, F32, F32, F32, F32, I32, I32, P, P],
    "mt_gen_room": [I32, P, P],
    "mt_gen_camera_rays": [P, P, P, F32, I32, I32, P, P],
    "mt_gen_random_rays": [U64, I64, I64, P, P, P, P, I32],
    "mt_gen_spheres": [U64, I32, F32, F32, F32, F32, F32, F32, P, P, P],
    "mt_bvh_build_device": [P, I32, P, P, P, P, C.POINTER(C.c_int32)],
    "mt_obj_parse": [C.c_char_p, U64, C.POINTER(ObjCounts), P, P, P, P],
    "mt_hdr_decode": [C.c_char_p, U64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), P],
    "mt_scene_create": [P, C.POINTER(SceneDesc), PP],
    "mt_scene_destroy": [P],
    "mt_trace_closest": [P, P, P, P, P, I32, P, P, P, I32],
    "mt_trace_any": [P, P, P, P, P, I32, P, P, I32],
    "mt_shade_hits": [P, P, P, P, P, P, P, I32, P, P, P, P, P],
    "mt_intersect_primitives": [P, P, P, P, P, I32, I32],
    "mt_intersect_occlusion": [P, P, P, P, P, I32, I32],
    "mt_shadow_rays": [P, P, P, P, P, P, P, I32, P, P, P],
    "mt_scan_excl_i32": [P, P, P, P, I32, P],
    "mt_compact": [P, P, I32, P, I32, P, P, C.POINTER(Payload), I32],
    "mt_compact_noalloc": [P, P, I32, P, I32, P, P, C.POINTER(Payload), I32],
    "mt_scan_check": [P],
    "mt_compact_isects": [P, P, P, P, P, I32, P],
    "mt_mesh_lights": [P, P, P, P, P],
    "mt_lightgrid_create": [P, PP],
    "mt_lightgrid_destroy": [P],
    "mt_lightgrid_arrays": [P, PP, PP, PP],
    "mt_lightgrid_clear": [P, P],
    "mt_lightgrid_set_tile_func": [P, P, P],
    "mt_lightgrid_update": [P, P],
    "mt_lightgrid_build_sat": [P, P],
    "mt_lightgrid_sample": [P, P, I32, P, P, P, P, P],
    "mt_lightgrid_cells": [P, I32, P, P, P],
    "mt_sat_tiles": [P, P, P, I32, I32, I32, I32],
    "mt_env_lum_sat": [P, P, I32, I32, P, P],
    "mt_env_create": [P, P, P, P, I32, I32, PP],
    "mt_env_destroy": [P],
    "mt_sat_sample_region": [P, P, P, I32, I32, I32, I32, I32, I32, I32, P, P, P, P],
    "mt_env_illuminate": [P, P, I32, P, P, P, P, P],
    "mt_env_radiance": [P, P, I32, P, P, P, P],
    "mt_lightgrid_set_camera": [P, P],
    "mt_env_illuminate_adaptive": [P, P, P, I32, P, P, P, P, P, P, P],
    "mt_env_radiance_adaptive": [P, P, P, I32, P, P, I32, P, P, P, P],
    "mt_env_attach_lightgrid": [P, P],
    "mt_bsdf_setup": [P, I32, P, P, P, P, P, P, P],
    "mt_bsdf_sample": [P, I32, P, P, P, P, P, P, P, P],
    "mt_bsdf_eval": [P, I32, P, P, P, P, P, P, P],
    "mt_trace_autotune": [P, P, P, P, P, I32, I32, P, P, P, I32, C.POINTER(C.c_int32), C.POINTER(C.c_float)],
    "mt_render": [P, P, P, P, P, P, C.POINTER(RenderDesc), P, C.POINTER(RenderStats)],
    "mt_render_graph_create": [P, P, P, P, P, P, C.POINTER(RenderDesc), P, PP],
    "mt_render_graph_launch": [P, C.c_uint32],
    "mt_render_graph_stats": [P, C.POINTER(RenderStats)],
    "mt_render_graph_destroy": [P],
    "mt_image_init": [P, P, P, I32],
    "mt_image_process": [P, P, F32, I32, F32, F32, I32, P, P, P, P],
    "mt_scan_excl_f32": [P, P, P, P, I32, P],
    "mt_batch_create": [P, I32, U64, I32, PP],
    "mt_batch_destroy": [P],
    "mt_batch_job_bytes": [C.POINTER(RenderDesc), C.POINTER(C.c_uint64)],
    "mt_batch_submit": [P, C.POINTER(JobDesc)],
    "mt_batch_step": [P, I32, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "mt_batch_get_stats": [P, C.POINTER(BatchStats)],
    "mt_batch_cancel": [P, C.c_int64, P],
    "mt_batch_set_iterations": [P, I32],
    "mt_batch_set_graphs": [P, I32],
    "mt_batch_set_trace_mode": [P, I32],
    "mt_batch_set_policy": [P, I32],
}
MT_BATCH_OLDEST_FIRST = 0
MT_BATCH_SMALLEST_FIRST = 1
_RESTYPES = {"mt_version": C.c_char_p}

_lib = None


def load() -> C.CDLL:
    """Loads libmambart.so and declares every entry point; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m mambatracer_amd.build` "
            "(or __graft_entry__.build()) first. There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int32)
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(512)
    load().mt_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> None:
    """Calls an entry point and raises MambartError on a non-zero status."""
    rc = getattr(load(), name)(*args)
    if rc != MT_OK:
        raise MambartError(rc, name, last_error())
