"""ctypes mirror of the structs and functions of include/rayon_hip.h."""
from __future__ import annotations

import ctypes as C

f3 = C.c_float * 3
d3 = C.c_double * 3

STATUS = {0: "RAYON_OK", 1: "RAYON_ERR_INVALID", 2: "RAYON_ERR_HIP", 3: "RAYON_ERR_OOM",
          4: "RAYON_ERR_BUDGET", 5: "RAYON_ERR_IO", 6: "RAYON_ERR_PARSE", 7: "RAYON_ERR_UNSUPPORTED"}


class _NS:
    def __init__(self, **kw):
        self.__dict__.update(kw)


MAT = _NS(LAMBERTIAN=0, METAL=1, MIRROR=2, ROUGH_MIRROR=3, GLASS=4, DIELECTRIC=5, LIGHT=6, CONSTANT=7,
          SHOW_NORMALS=8, SDF_MATERIAL=9, ANISOTROPIC_METAL=10, THIN_FILM=11, CLEAR_COAT=12,
          GGX_CONDUCTOR=13, GGX_DIELECTRIC=14)
GEOM = _NS(SPHERE=0, RECTANGLE=1, DISPLACED_SPHERE=3, TRIANGLE=4)
SDF = _NS(SPHERE=0, ROUNDED_BOX=1, TORUS=2)
PATTERN = _NS(NONE=0, FIBONACCI_DOTS=1)
FLAG = _NS(REF_QUIRKS=1, REF_RNG=2, NEE=4, ADAPTIVE_REF=8, ADAPTIVE_STDERR=16, DOF=32, COUNT_TRAVERSAL=64)
GEN = _NS(BOOK_SPHERES=0, CORNELL_ICOSPHERE=1, SDF_TERRAIN=2)
BVH = _NS(DEFAULT=0, QUANTIZED=1, DEVICE=2, BUILDER_HOST=0, BUILDER_DEVICE=1)
YAML = _NS(LIFT_CAPS=1)
JOB = _NS(QUEUED=0, RUNNING=1, DONE=2, FAILED=3, CANCELLED=4)
PATH = _NS(AUTO=0, WAVEFRONT=1, MEGAKERNEL=2)
OPT = _NS(WAVE_CAPACITY=0, WAVE_WG_PER_CU=1, GRAPH=2, STAGE_TIMING=3, COUNT_TRAVERSAL=4, WAVE_CHUNK=5, SHADE_CHUNK=6,
          TRACE_REFILL=7, TRACE_LEAF_EXIT=8)


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", f3), ("emission", f3), ("roughness", C.c_float),
                ("metallic", C.c_float), ("refractive_index", C.c_float), ("transmission", C.c_float),
                ("anisotropy", C.c_float), ("eta", f3), ("k", f3), ("film_thickness", C.c_float),
                ("film_ior", C.c_float), ("pattern", C.c_int32), ("pattern_color", f3),
                ("pattern_param1", C.c_float), ("pattern_param2", C.c_float)]


class Geometry(C.Structure):
    _fields_ = [("type", C.c_int32), ("material_id", C.c_int32), ("visible", C.c_int32),
                ("has_normals", C.c_int32), ("p", C.c_float * 9), ("n", C.c_float * 9)]


class Mesh(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_int32)), ("num_vertices", C.c_int64), ("num_triangles", C.c_int64),
                ("material_id", C.c_int32), ("visible", C.c_int32)]


class SdfPrim(C.Structure):
    _fields_ = [("type", C.c_int32), ("center", f3), ("params", C.c_float * 4)]


class SceneDesc(C.Structure):
    _fields_ = [("materials", C.POINTER(Material)), ("num_materials", C.c_int32),
                ("geometries", C.POINTER(Geometry)), ("num_geometries", C.c_int64),
                ("meshes", C.POINTER(Mesh)), ("num_meshes", C.c_int32),
                ("sdf_prims", C.POINTER(SdfPrim)), ("num_sdf_prims", C.c_int32),
                ("sdf_smooth_k", C.c_float), ("sdf_material_id", C.c_int32), ("sdf_max_steps", C.c_int32),
                ("sdf_eps", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("num_prims", C.c_int64), ("num_nodes", C.c_int64), ("num_qnodes", C.c_int64),
                ("max_depth", C.c_int32), ("num_lights", C.c_int32), ("device_bytes", C.c_int64),
                ("build_ms", C.c_double), ("upload_ms", C.c_double), ("builder", C.c_int32), ("wide_depth", C.c_int32)]


class Camera(C.Structure):
    _fields_ = [("center", d3), ("pixel00", d3), ("delta_u", d3), ("delta_v", d3), ("u", d3), ("v", d3),
                ("w", d3)]


class JobParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("camera_center", f3), ("pixel00", f3),
                ("delta_u", f3), ("delta_v", f3), ("cam_u", f3), ("cam_v", f3), ("max_depth", C.c_int32),
                ("spp", C.c_int32), ("seed", C.c_uint64), ("flags", C.c_uint32),
                ("light_intensity", C.c_float), ("background_intensity", C.c_float),
                ("metal_fuzziness", C.c_float), ("glass_ior", C.c_float), ("dof_aperture", C.c_float),
                ("dof_focus_distance", C.c_float), ("min_adaptive_samples", C.c_int32),
                ("adaptive_threshold", C.c_float)]


class JobStats(C.Structure):
    _fields_ = [("samples_done", C.c_int64), ("primary_samples", C.c_int64), ("segments", C.c_int64),
                ("shadow_segments", C.c_int64), ("converged_pixels", C.c_int64), ("done", C.c_int32),
                ("pad_", C.c_int32), ("node_fetches", C.c_int64), ("prim_fetches", C.c_int64),
                ("material_fetches", C.c_int64), ("light_fetches", C.c_int64)]


class RenderStats(C.Structure):
    _fields_ = [("samples_per_pixel", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("max_depth", C.c_int32), ("rays_traced", C.c_int64), ("image_size_bytes", C.c_int64),
                ("render_time_ns", C.c_int64)]


class WaveStats(C.Structure):
    _fields_ = [("stage_ms", C.c_double * 8), ("launches", C.c_int64 * 8), ("rounds", C.c_int64),
                ("extend_node_fetches", C.c_int64), ("extend_prim_fetches", C.c_int64),
                ("shadow_node_fetches", C.c_int64), ("shadow_prim_fetches", C.c_int64),
                ("material_fetches", C.c_int64), ("light_fetches", C.c_int64), ("sdf_marches", C.c_int64),
                ("sdf_steps", C.c_int64), ("sdf_prim_evals", C.c_int64), ("sdf_wave_slots", C.c_int64)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active_jobs", C.c_int32),
                ("first_pass_samples", C.c_int32), ("max_pass_samples", C.c_int32), ("pad_", C.c_int32),
                ("target_pass_ms", C.c_double), ("max_pass_pixels", C.c_int64)]


class ServerJobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("first_kept", C.c_int32), ("samples_done", C.c_int64), ("segments", C.c_int64),
                ("queue_ms", C.c_double), ("ttff_ms", C.c_double), ("ttc_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("passes", C.c_int64), ("first_passes", C.c_int64), ("completed", C.c_int64),
                ("failed", C.c_int64), ("queued", C.c_int64), ("active", C.c_int64), ("segments", C.c_int64),
                ("used_bytes", C.c_int64), ("busy_s", C.c_double), ("wall_s", C.c_double),
                ("pass_samples", C.c_int32), ("pass_jobs", C.c_int32), ("pass_ms", C.c_double),
                ("first_ms", C.c_double), ("check_ms", C.c_double), ("pass_paths", C.c_double),
                ("first_kept", C.c_int64), ("first_kept_paths", C.c_int64)]


class TuneResult(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("chunk", C.c_int32), ("wg_per_cu", C.c_int32), ("best_ms", C.c_double),
                ("start_ms", C.c_double), ("candidates", C.c_int32), ("cached", C.c_int32),
                ("mega_scenes", C.c_int32), ("wave_scenes", C.c_int32), ("shade_chunk", C.c_int32),
                ("refill", C.c_int32)]


class ProgressiveConfig(C.Structure):
    _fields_ = [("target_fps", C.c_double), ("max_batch", C.c_int32), ("first_batch", C.c_int32),
                ("adaptive_depth", C.c_int32), ("channels", C.c_int32), ("gamma", C.c_float), ("pad_", C.c_int32)]


class ProgressiveFrame(C.Structure):
    _fields_ = [("samples", C.c_int64), ("segments", C.c_int64), ("frame_ms", C.c_double),
                ("convergence", C.c_double), ("batch", C.c_int32), ("next_batch", C.c_int32),
                ("depth", C.c_int32), ("done", C.c_int32)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("seed", C.c_uint64), ("detail", C.c_int32), ("count", C.c_int32)]


def declare(lib: C.CDLL) -> C.CDLL:
    vp, i32, i64, f32, st = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_int
    P = C.POINTER
    sig = {
        "rayon_last_error": (C.c_char_p, []),
        "rayon_abi_version": (i32, []),
        "rayon_camera_frame": (st, [P(C.c_double), P(C.c_double), P(C.c_double), C.c_double, i32, i32,
                                    P(Camera)]),
        "rayon_job_params_default": (None, [P(JobParams)]),
        "rayon_job_params_set_camera": (None, [P(JobParams), P(Camera)]),
        "rayon_ctx_create": (st, [i32, P(vp)]),
        "rayon_ctx_destroy": (st, [vp]),
        "rayon_ctx_stream": (vp, [vp]),
        "rayon_ctx_sync": (st, [vp]),
        "rayon_scene_create": (st, [vp, P(SceneDesc), C.c_uint32, P(vp)]),
        "rayon_scene_info_get": (st, [vp, P(SceneInfo)]),
        "rayon_bvh_check": (st, [P(SceneDesc), P(SceneInfo)]),
        "rayon_scene_check": (st, [vp, P(SceneDesc), P(i32)]),
        "rayon_scene_prim_order": (st, [vp, P(i64), i64]),
        "rayon_scene_destroy": (st, [vp]),
        "rayon_job_create": (st, [vp, vp, P(JobParams), P(vp)]),
        "rayon_job_destroy": (st, [vp]),
        "rayon_job_reset": (st, [vp]),
        "rayon_job_render_plain": (st, [vp, vp, i32]),
        "rayon_render_batched": (st, [vp, P(vp), i32, i32]),
        "rayon_ctx_set_option": (st, [vp, i32, i64]),
        "rayon_ctx_get_option": (st, [vp, i32, P(i64)]),
        "rayon_ctx_wave_stats": (st, [vp, P(WaveStats), i32]),
        "rayon_job_resolve_linear": (st, [vp, vp, P(f32)]),
        "rayon_job_resolve_u8": (st, [vp, vp, P(C.c_uint8), i32, f32]),
        "rayon_job_resolve_u8_noalloc": (st, [vp, vp, P(C.c_uint8), i32, f32]),
        "rayon_job_sample_counts": (st, [vp, vp, P(i32)]),
        "rayon_job_stats_get": (st, [vp, vp, P(JobStats)]),
        "rayon_trace_closest": (st, [vp, vp, P(f32), P(f32), i64, f32, f32, P(i64), P(f32)]),
        "rayon_job_primary_rays": (st, [vp, vp, i32, P(f32), P(f32)]),
        "rayon_scatter_probe": (st, [vp, P(Material), C.c_uint32, f32, f32, f32, P(f32), P(f32), P(f32), P(i32),
                                     P(f32), i32, i64, P(f32), P(f32), P(f32), P(i32), P(i32)]),
        "rayon_accumulate_probe": (st, [vp, P(f32), i64, P(f32), P(f32)]),
        "rayon_scene_set_batch_path": (st, [vp, i32]),
        "rayon_scene_batch_path": (i32, [vp]),
        "rayon_progressive_config_default": (None, [P(ProgressiveConfig)]),
        "rayon_progressive_create": (st, [vp, vp, P(JobParams), P(ProgressiveConfig), P(vp)]),
        "rayon_progressive_step": (st, [vp, P(ProgressiveFrame)]),
        "rayon_progressive_display": (st, [vp, P(C.c_uint8), P(i64)]),
        "rayon_progressive_reset": (st, [vp, P(Camera)]),
        "rayon_progressive_job": (vp, [vp]),
        "rayon_progressive_destroy": (st, [vp]),
        "rayon_autotune": (st, [vp, P(vp), i32, i32, P(TuneResult)]),
        "rayon_server_config_default": (None, [P(ServerConfig)]),
        "rayon_server_create": (st, [vp, P(ServerConfig), P(vp)]),
        "rayon_server_submit": (st, [vp, vp, P(JobParams), P(C.c_uint64)]),
        "rayon_server_job_info_get": (st, [vp, C.c_uint64, P(ServerJobInfo)]),
        "rayon_server_result": (st, [vp, C.c_uint64, P(f32)]),
        "rayon_server_first_frame": (st, [vp, C.c_uint64, P(C.c_uint8)]),
        "rayon_server_cancel": (st, [vp, C.c_uint64]),
        "rayon_server_release": (st, [vp, C.c_uint64]),
        "rayon_server_drain": (st, [vp, C.c_double]),
        "rayon_server_stats_get": (st, [vp, P(ServerStats)]),
        "rayon_server_destroy": (st, [vp]),
        "rayon_generate": (st, [P(GenParams), P(vp)]),
        "rayon_load_yaml": (st, [C.c_char_p, P(vp)]),
        "rayon_load_yaml_ex": (st, [C.c_char_p, C.c_uint32, P(vp)]),
        "rayon_save_yaml": (st, [vp, C.c_char_p]),
        "rayon_generated_settings": (st, [vp, P(f32), P(f32), P(f32), P(i32), P(i32)]),
        "rayon_generated_desc": (P(SceneDesc), [vp]),
        "rayon_generated_camera": (st, [vp, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]),
        "rayon_generated_background_intensity": (f32, [vp]),
        "rayon_generated_free": (None, [vp]),
        "rayon_write_png": (st, [C.c_char_p, P(C.c_uint8), i32, i32, i32]),
        "rayon_job_write_png": (st, [vp, vp, C.c_char_p, f32]),
        "rayon_write_render_stats": (st, [C.c_char_p, P(RenderStats)]),
        "rayon_format_duration": (st, [i64, C.c_char_p, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
