"""ctypes view of include/luxion_mi355x.h (the C ABI of libluxion_mi355x.so).

The structure layouts below follow the header field by field; tests check
that every symbol the header declares is exported.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libluxion_mi355x.so")

LX_OK = 0
STATUS_NAMES = {
    0: "LX_OK", 1: "LX_ERR_INVALID", 2: "LX_ERR_OOM", 3: "LX_ERR_HIP", 4: "LX_ERR_IO",
    5: "LX_ERR_FORMAT", 6: "LX_ERR_BUDGET", 7: "LX_ERR_CANCELLED", 8: "LX_ERR_TIMEOUT",
    9: "LX_ERR_UNSUPPORTED", 10: "LX_ERR_STATE", 11: "LX_ERR_NOT_FOUND",
}
LX_TEX_RGBA8, LX_TEX_RGBA16, LX_TEX_RGBA32F = 0, 1, 2
LX_ALPHA_OPAQUE, LX_ALPHA_MASK, LX_ALPHA_BLEND = 0, 1, 2
LX_CDF_FP64, LX_CDF_FP32_REFERENCE = 0, 1
LX_TONEMAP_REINHARD, LX_TONEMAP_AGX, LX_TONEMAP_ACES = 0, 1, 2
LX_TIMING_EVENTS, LX_TIMING_COUNT_WORK = 1, 2
LX_MATERIAL_MICROFACET, LX_MATERIAL_DIFFUSE, LX_MATERIAL_SPECULAR, LX_MATERIAL_EMISSIVE, LX_MATERIAL_GLASS = range(5)
LX_DEBUG_OFF, LX_DEBUG_ALBEDO, LX_DEBUG_WORLD_NORMAL, LX_DEBUG_NORMAL_MAP, LX_DEBUG_ROUGHNESS, LX_DEBUG_METALLIC, \
    LX_DEBUG_ALPHA = range(7)
LX_SHADING_UBER, LX_SHADING_LEGACY = 0, 1
LX_SORT_NONE, LX_SORT_RAYS, LX_SORT_MATERIALS = 0, 1, 2

F2 = C.c_float * 2
F3 = C.c_float * 3
F12 = C.c_float * 12


class lx_tex_transform(C.Structure):
    _fields_ = [("offset", F2), ("scale", F2), ("rotation", C.c_float)]


class lx_mesh_desc(C.Structure):
    _fields_ = [("num_verts", C.c_uint32), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("uvs", C.c_void_p), ("num_tris", C.c_uint32), ("indices", C.c_void_p)]


class lx_instance_desc(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("material", C.c_uint32), ("transform", F12)]


class lx_material_desc(C.Structure):
    _fields_ = [("base_color", F3), ("albedo_tex", C.c_int32), ("albedo_xf", lx_tex_transform),
                ("roughness", C.c_float), ("metallic", C.c_float), ("metallic_roughness_tex", C.c_int32),
                ("metallic_roughness_xf", lx_tex_transform), ("normal_tex", C.c_int32),
                ("normal_xf", lx_tex_transform), ("emission_color", F3), ("emission_strength", C.c_float),
                ("emissive_tex", C.c_int32), ("emissive_xf", lx_tex_transform), ("ior", C.c_float),
                ("transmission", C.c_float), ("alpha", C.c_float), ("alpha_mode", C.c_int32),
                ("alpha_cutoff", C.c_float), ("material_type", C.c_int32)]


class lx_texture_desc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_int32), ("data", C.c_void_p)]


class lx_camera_desc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("position", F3), ("view", F3), ("up", F3),
                ("right", F3), ("pixel_length", F2)]


class lx_scene_desc(C.Structure):
    _fields_ = [("meshes", C.POINTER(lx_mesh_desc)), ("num_meshes", C.c_uint32),
                ("instances", C.POINTER(lx_instance_desc)), ("num_instances", C.c_uint32),
                ("materials", C.POINTER(lx_material_desc)), ("num_materials", C.c_uint32),
                ("textures", C.POINTER(lx_texture_desc)), ("num_textures", C.c_uint32),
                ("envmap", C.POINTER(lx_texture_desc)), ("camera", lx_camera_desc), ("trace_depth", C.c_int32)]


class lx_build_opts(C.Structure):
    _fields_ = [("bvh_width", C.c_int32), ("leaf_size", C.c_int32), ("builder", C.c_int32),
                ("sah_bins", C.c_int32), ("cdf_precision", C.c_int32)]


class lx_scene_info(C.Structure):
    _fields_ = [("num_meshes", C.c_uint32), ("num_instances", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_textures", C.c_uint32), ("num_unique_tris", C.c_uint64), ("num_instanced_tris", C.c_uint64),
                ("num_emissive_instances", C.c_uint32), ("total_emissive_area", C.c_float),
                ("env_width", C.c_int32), ("env_height", C.c_int32), ("device_bytes", C.c_uint64),
                ("bvh_nodes", C.c_uint32), ("bvh_depth", C.c_int32), ("build_ms", C.c_double),
                ("sah_cost", C.c_double), ("name", C.c_char * 128)]


class lx_render_opts(C.Structure):
    _fields_ = [("max_depth", C.c_int32), ("dmis", C.c_int32), ("emis", C.c_int32),
                ("envmap_intensity", C.c_float), ("seed", C.c_uint32), ("color_mode", C.c_int32),
                ("camera", C.POINTER(lx_camera_desc)), ("material_debug_mode", C.c_int32),
                ("shading_model", C.c_int32)]


class lx_render_stats(C.Structure):
    _fields_ = [("radiance_rays", C.c_uint64), ("dl_shadow_rays", C.c_uint64), ("env_shadow_rays", C.c_uint64),
                ("path_bounces", C.c_uint64), ("iterations", C.c_uint32), ("trace_ms", C.c_double),
                ("shade_ms", C.c_double), ("other_ms", C.c_double), ("trace_launches", C.c_uint32),
                ("shade_launches", C.c_uint32), ("closest_nodes", C.c_uint64), ("closest_tris", C.c_uint64),
                ("closest_insts", C.c_uint64), ("any_nodes", C.c_uint64), ("any_tris", C.c_uint64),
                ("any_insts", C.c_uint64), ("tail_closest_rays", C.c_uint64), ("tail_any_rays", C.c_uint64),
                ("tail_closest_nodes", C.c_uint64), ("tail_closest_tris", C.c_uint64),
                ("tail_closest_insts", C.c_uint64), ("tail_any_nodes", C.c_uint64), ("tail_any_tris", C.c_uint64),
                ("tail_any_insts", C.c_uint64), ("tail_ms", C.c_double), ("tail_launches", C.c_uint32),
                ("tail_bounces", C.c_uint64)]


class lx_ray(C.Structure):
    _fields_ = [("origin", F3), ("tmin", C.c_float), ("dir", F3), ("tmax", C.c_float)]


class lx_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("instance", C.c_uint32), ("prim", C.c_uint32), ("u", C.c_float),
                ("v", C.c_float)]


class lx_job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", lx_camera_desc), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("dmis", C.c_int32), ("emis", C.c_int32), ("envmap_intensity", C.c_float), ("seed", C.c_uint32),
                ("first_iter", C.c_uint32), ("deadline_ms", C.c_double), ("material_debug_mode", C.c_int32),
                ("shading_model", C.c_int32)]


class lx_job_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("rgb_sum", C.c_void_p), ("spp_done", C.c_uint32), ("rays", C.c_uint64),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("finish_ms", C.c_double)]


class lx_tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("bvh_width", C.c_int32), ("queue_chunk", C.c_int32),
                ("waves_per_cu", C.c_int32), ("use_graph", C.c_int32), ("max_paths_per_round", C.c_uint32),
                ("wavefront_bounces", C.c_int32), ("tail_min_waves", C.c_int32), ("sort_mode", C.c_int32),
                ("pixel_tile", C.c_int32), ("tile_band", C.c_int32), ("overlap_rounds", C.c_int32),
                ("round_frames", C.c_int32)]


class lx_server_opts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_jobs_in_flight", C.c_int32),
                ("spp_per_round", C.c_int32), ("max_paths_per_round", C.c_uint32), ("max_depth", C.c_int32)]


class lx_server_stats(C.Structure):
    _fields_ = [("jobs_done", C.c_uint64), ("jobs_failed", C.c_uint64), ("jobs_rejected", C.c_uint64),
                ("rounds", C.c_uint64), ("rays", C.c_uint64), ("busy_ms", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("peak_bytes", C.c_uint64), ("peak_jobs_in_flight", C.c_uint32)]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
SIGNATURES = {
    "lx_last_error": (C.c_char_p, []),
    "lx_abi_version": (C.c_int, []),
    "lx_now_ms": (C.c_double, []),
    "lx_context_create": (C.c_int, [C.c_int, C.c_uint64, _PP]),
    "lx_context_destroy": (None, [_P]),
    "lx_context_bytes_in_use": (C.c_uint64, [_P]),
    "lx_scene_create": (C.c_int, [_P, C.POINTER(lx_scene_desc), _PP]),
    "lx_scene_load_glb": (C.c_int, [_P, C.c_char_p, C.c_char_p, _PP]),
    "lx_render_save_png": (C.c_int, [_P, C.c_char_p]),
    "lx_debug_write_png": (C.c_int, [C.c_char_p, _P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "lx_debug_parse_glb": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(lx_scene_desc), _P, C.c_uint32]),
    "lx_scene_build_accel": (C.c_int, [_P, C.POINTER(lx_build_opts)]),
    "lx_scene_get_camera": (C.c_int, [_P, C.POINTER(lx_camera_desc), C.POINTER(C.c_int32)]),
    "lx_scene_get_info": (C.c_int, [_P, C.POINTER(lx_scene_info)]),
    "lx_scene_read_env_cdfs": (C.c_int, [_P, _P, _P]),
    "lx_scene_read_emitter_cdfs": (C.c_int, [_P, _P, _P, C.POINTER(C.c_uint32)]),
    "lx_scene_read_instance_tri_cdf": (C.c_int, [_P, C.c_uint32, _P, C.POINTER(C.c_uint32)]),
    "lx_scene_destroy": (None, [_P]),
    "lx_render_create": (C.c_int, [_P, C.POINTER(lx_render_opts), _PP]),
    "lx_render_iterate": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "lx_render_sync": (C.c_int, [_P]),
    "lx_render_read_accum": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "lx_render_read_ldr": (C.c_int, [_P, _P, C.c_int32]),
    "lx_render_reset": (C.c_int, [_P]),
    "lx_render_stats_get": (C.c_int, [_P, C.POINTER(lx_render_stats)]),
    "lx_render_set_timing": (C.c_int, [_P, C.c_int32]),
    "lx_render_destroy": (None, [_P]),
    "lx_trace_closest": (C.c_int, [_P, _P, C.c_uint32, C.c_int32, C.c_uint32, _P]),
    "lx_trace_occluded": (C.c_int, [_P, _P, C.c_uint32, C.c_int32, C.c_uint32, _P]),
    "lx_trace_emitting": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P]),
    "lx_trace_closest_noalloc": (C.c_int, [_P, _P, C.c_uint32, C.c_int32, C.c_uint32, _P]),
    "lx_debug_bounce": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P]),
    "lx_debug_math_eval": (C.c_int, [_P, C.c_int32, _P, _P, _P, C.c_uint32]),
    "lx_debug_copy_bandwidth": (C.c_int, [_P, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
    "lx_tuning_get": (C.c_int, [_P, C.POINTER(lx_tuning)]),
    "lx_tuning_set": (C.c_int, [_P, C.POINTER(lx_tuning)]),
    "lx_autotune": (C.c_int, [_P, _P, C.c_uint32, C.c_char_p, C.POINTER(lx_tuning)]),
    "lx_tuning_load": (C.c_int, [_P, C.c_char_p]),
    "lx_batch_run": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "lx_server_start": (C.c_int, [_P, C.POINTER(lx_server_opts), _PP]),
    "lx_server_submit": (C.c_int, [_P, C.POINTER(lx_job), C.POINTER(lx_job_result), C.POINTER(C.c_uint64)]),
    "lx_server_poll": (C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "lx_server_cancel": (C.c_int, [_P, C.c_uint64]),
    "lx_server_stats_get": (C.c_int, [_P, C.POINTER(lx_server_stats)]),
    "lx_server_inject_budget_failure": (C.c_int, [_P, C.c_uint32]),
    "lx_server_stop": (C.c_int, [_P]),
}

_lib = None


class LuxionError(RuntimeError):
    def __init__(self, status, msg):
        self.status = status
        self.status_name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{self.status_name}: {msg}")


def load():
    """Load the native library; raises if it has not been built (there is no
    fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C luxion_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != LX_OK:
        raise LuxionError(status, load().lx_last_error().decode())
