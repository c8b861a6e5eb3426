"""Python binding of librayon_hip.so (the C ABI in include/rayon_hip.h).

The library is the product: a C++/HIP path tracer for MI355X. This module only wraps its
entry points with ctypes for tests, the benchmark and scripting. It never falls back to
anything else: if the shared library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .abi import (  # noqa: F401  (re-exported)
    Camera, GenParams, Geometry, JobParams, JobStats, Material, Mesh, RenderStats, SceneDesc, SceneInfo, SdfPrim,
    WaveStats,
    ServerConfig, ServerJobInfo, ServerStats, TuneResult, ProgressiveConfig, ProgressiveFrame, PATH, YAML, MAT, GEOM, SDF, PATTERN, FLAG, GEN, OPT, BVH, JOB, STATUS, declare,
)

LIB_PATH = Path(__file__).resolve().parent / "librayon_hip.so"
_lib = None


class RayonError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Loads librayon_hip.so (built in-tree by rayon_amd.build)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RayonError(f"{LIB_PATH} is not built; run `python -m rayon_amd.build`")
        _lib = declare(C.CDLL(str(LIB_PATH)))
    return _lib


def _check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().rayon_last_error().decode(errors="replace")
        raise RayonError(f"{what}: {STATUS.get(status, status)}: {msg}")


def camera_frame(look_from, look_at, vup, vfov: float, width: int, height: int) -> Camera:
    cam = Camera()
    _check(lib().rayon_camera_frame((C.c_double * 3)(*look_from), (C.c_double * 3)(*look_at),
                                    (C.c_double * 3)(*vup), float(vfov), int(width), int(height),
                                    C.byref(cam)), "rayon_camera_frame")
    return cam


def bvh_check(desc_ptr) -> SceneInfo:
    """Host-only BVH build + invariant check (no GPU needed)."""
    info = SceneInfo()
    _check(lib().rayon_bvh_check(desc_ptr, C.byref(info)), "rayon_bvh_check")
    return info


def job_params(width: int, height: int, camera: Camera, *, spp: int = 64, max_depth: int = 16,
               seed: int = 1234, flags: int = 0, **kw) -> JobParams:
    p = JobParams()
    lib().rayon_job_params_default(C.byref(p))
    p.width, p.height, p.spp, p.max_depth, p.seed, p.flags = width, height, spp, max_depth, seed, flags
    lib().rayon_job_params_set_camera(C.byref(p), C.byref(camera))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def write_png(path: str, image: np.ndarray):
    """Writes an 8-bit image (H x W, or H x W x C with C in 1..4) as a PNG file."""
    a = np.ascontiguousarray(image, dtype=np.uint8)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ValueError("image must be H x W or H x W x C")
    h, w, c = a.shape
    _check(lib().rayon_write_png(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_uint8)), w, h, c),
           "rayon_write_png")


def write_render_stats(image_path: str, *, samples_per_pixel: int, width: int, height: int, max_depth: int,
                       rays_traced: int, render_time_ns: int, image_size_bytes: int = -1):
    """Writes the .txt and .json statistics of a render next to its image."""
    s = RenderStats(samples_per_pixel, width, height, max_depth, rays_traced, image_size_bytes, render_time_ns)
    _check(lib().rayon_write_render_stats(str(image_path).encode(), C.byref(s)), "rayon_write_render_stats")


def format_duration(duration_ns: int) -> str:
    buf = C.create_string_buffer(64)
    _check(lib().rayon_format_duration(int(duration_ns), buf, 64), "rayon_format_duration")
    return buf.value.decode()


class Generated:
    """A synthetic or YAML-loaded scene description held by the library."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def generate(cls, kind: int, seed: int, detail: int = 0, count: int = 0) -> "Generated":
        gp = GenParams(kind=kind, seed=seed, detail=detail, count=count)
        h = C.c_void_p()
        _check(lib().rayon_generate(C.byref(gp), C.byref(h)), "rayon_generate")
        return cls(h)

    @classmethod
    def load_yaml(cls, path: str, flags: int = 0) -> "Generated":
        h = C.c_void_p()
        _check(lib().rayon_load_yaml_ex(str(path).encode(), flags, C.byref(h)), "rayon_load_yaml")
        return cls(h)

    def save_yaml(self, path: str):
        _check(lib().rayon_save_yaml(self._h, str(path).encode()), "rayon_save_yaml")

    def settings(self) -> dict:
        col = (C.c_float * 3)()
        bi, amb = C.c_float(), C.c_float()
        bvh, ada = C.c_int32(), C.c_int32()
        _check(lib().rayon_generated_settings(self._h, col, C.byref(bi), C.byref(amb), C.byref(bvh), C.byref(ada)),
               "rayon_generated_settings")
        return dict(background_color=list(col), background_intensity=bi.value, ambient_light=amb.value,
                    use_bvh=bool(bvh.value), adaptive_sampling=bool(ada.value))

    @property
    def desc(self) -> SceneDesc:
        return lib().rayon_generated_desc(self._h).contents

    def desc_ptr(self):
        return lib().rayon_generated_desc(self._h)

    def camera(self):
        lf, la, up = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        fov = C.c_double()
        _check(lib().rayon_generated_camera(self._h, lf, la, up, C.byref(fov)), "rayon_generated_camera")
        return list(lf), list(la), list(up), fov.value

    def background_intensity(self) -> float:
        return lib().rayon_generated_background_intensity(self._h)

    def close(self):
        if self._h:
            lib().rayon_generated_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().rayon_ctx_create(device, C.byref(h)), "rayon_ctx_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        return lib().rayon_ctx_stream(self._h) or 0

    def sync(self):
        _check(lib().rayon_ctx_sync(self._h), "rayon_ctx_sync")

    def scene(self, desc_ptr, bvh_flags: int = 0) -> "Scene":
        h = C.c_void_p()
        _check(lib().rayon_scene_create(self._h, desc_ptr, bvh_flags, C.byref(h)), "rayon_scene_create")
        return Scene(self, h)

    def job(self, scene: "Scene", params: JobParams) -> "Job":
        h = C.c_void_p()
        _check(lib().rayon_job_create(self._h, scene._h, C.byref(params), C.byref(h)), "rayon_job_create")
        return Job(self, scene, h, params)

    def render_batched(self, jobs, samples: int):
        arr = (C.c_void_p * len(jobs))(*[j._h for j in jobs])
        _check(lib().rayon_render_batched(self._h, arr, len(jobs), samples), "rayon_render_batched")

    def set_option(self, opt: int, value: int):
        _check(lib().rayon_ctx_set_option(self._h, opt, value), "rayon_ctx_set_option")

    def get_option(self, opt: int) -> int:
        v = C.c_int64()
        _check(lib().rayon_ctx_get_option(self._h, opt, C.byref(v)), "rayon_ctx_get_option")
        return v.value

    def scatter_probe(self, m: Material, dirs, normals, tangents, front, rnd, flags: int = 0, fuzz: float = 0.8,
                      glass_ior: float = 1.5, light: float = 1.0):
        """The device's scatter_material for n given hits and random-number rows (rnd: n x k).
        Returns (scattered, dir, atten, emitted, used)."""
        f32p = C.POINTER(C.c_float)
        i32p = C.POINTER(C.c_int32)
        d = np.ascontiguousarray(dirs, np.float32)
        nn = np.ascontiguousarray(normals, np.float32)
        t = np.ascontiguousarray(tangents, np.float32)
        fr = np.ascontiguousarray(front, np.int32)
        r = np.ascontiguousarray(rnd, np.float32)
        n, k = r.shape
        od, at, em = (np.empty((n, 3), np.float32) for _ in range(3))
        ok, used = np.empty(n, np.int32), np.empty(n, np.int32)
        _check(lib().rayon_scatter_probe(self._h, C.byref(m), flags, fuzz, glass_ior, light, d.ctypes.data_as(f32p),
                                         nn.ctypes.data_as(f32p), t.ctypes.data_as(f32p), fr.ctypes.data_as(i32p),
                                         r.ctypes.data_as(f32p), k, n, od.ctypes.data_as(f32p),
                                         at.ctypes.data_as(f32p), em.ctypes.data_as(f32p), ok.ctypes.data_as(i32p),
                                         used.ctypes.data_as(i32p)), "rayon_scatter_probe")
        return ok.astype(bool), od, at, em, used

    def accumulate_probe(self, rgb):
        """The device's compensated per-pixel accumulation over the series rgb (n x 3):
        (sum rgb + luminance^2, compensation), each 4 floats."""
        v = np.ascontiguousarray(rgb, np.float32)
        s, c = np.empty(4, np.float32), np.empty(4, np.float32)
        f32p = C.POINTER(C.c_float)
        _check(lib().rayon_accumulate_probe(self._h, v.ctypes.data_as(f32p), v.shape[0], s.ctypes.data_as(f32p),
                                            c.ctypes.data_as(f32p)), "rayon_accumulate_probe")
        return s, c

    def autotune(self, jobs, samples: int) -> TuneResult:
        """Times wavefront launch shapes on clones of `jobs` and keeps the fastest on this context."""
        arr = (C.c_void_p * len(jobs))(*[j._h for j in jobs])
        r = TuneResult()
        _check(lib().rayon_autotune(self._h, arr, len(jobs), samples, C.byref(r)), "rayon_autotune")
        return r

    def progressive(self, scene: "Scene", params: JobParams, **cfg) -> "Progressive":
        return Progressive(self, scene, params, **cfg)

    def server(self, **cfg) -> "Server":
        return Server(self, **cfg)

    def wave_stats(self, reset: bool = False) -> WaveStats:
        s = WaveStats()
        _check(lib().rayon_ctx_wave_stats(self._h, C.byref(s), int(reset)), "rayon_ctx_wave_stats")
        return s

    def close(self):
        if self._h:
            lib().rayon_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Scene:
    def __init__(self, ctx: Context, h):
        self.ctx, self._h = ctx, h

    def info(self) -> SceneInfo:
        i = SceneInfo()
        _check(lib().rayon_scene_info_get(self._h, C.byref(i)), "rayon_scene_info_get")
        return i

    def set_batch_path(self, path: int):
        _check(lib().rayon_scene_set_batch_path(self._h, path), "rayon_scene_set_batch_path")

    @property
    def batch_path(self) -> int:
        return lib().rayon_scene_batch_path(self._h)

    def check(self, desc_ptr) -> int:
        """Verifies the scene's binary tree against its description; returns the tree depth."""
        depth = C.c_int32(0)
        _check(lib().rayon_scene_check(self._h, desc_ptr, C.byref(depth)), "rayon_scene_check")
        return depth.value

    def prim_order(self) -> np.ndarray:
        n = self.info().num_prims
        out = np.empty(n, np.int64)
        _check(lib().rayon_scene_prim_order(self._h, out.ctypes.data_as(C.POINTER(C.c_int64)), n),
               "rayon_scene_prim_order")
        return out

    def trace_closest(self, origins: np.ndarray, dirs: np.ndarray, t_min: float = 0.001,
                      t_max: float = 3.4e38):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        prim = np.empty(n, np.int64)
        t = np.empty(n, np.float32)
        _check(lib().rayon_trace_closest(self.ctx._h, self._h, o.ctypes.data_as(C.POINTER(C.c_float)),
                                         d.ctypes.data_as(C.POINTER(C.c_float)), n, t_min, t_max,
                                         prim.ctypes.data_as(C.POINTER(C.c_int64)),
                                         t.ctypes.data_as(C.POINTER(C.c_float))), "rayon_trace_closest")
        return prim, t

    def close(self):
        if self._h:
            lib().rayon_scene_destroy(self._h)
            self._h = None


class Job:
    def __init__(self, ctx: Context, scene: Scene, h, params: JobParams):
        self.ctx, self.scene, self._h, self.params = ctx, scene, h, params

    @property
    def shape(self):
        return self.params.height, self.params.width

    def render_plain(self, samples: int):
        _check(lib().rayon_job_render_plain(self.ctx._h, self._h, samples), "rayon_job_render_plain")

    def reset(self):
        _check(lib().rayon_job_reset(self._h), "rayon_job_reset")

    def resolve_linear(self) -> np.ndarray:
        h, w = self.shape
        out = np.empty((h, w, 3), np.float32)
        _check(lib().rayon_job_resolve_linear(self.ctx._h, self._h, out.ctypes.data_as(C.POINTER(C.c_float))),
               "rayon_job_resolve_linear")
        return out

    def resolve_u8(self, channels: int = 3, gamma: float = 2.0) -> np.ndarray:
        h, w = self.shape
        out = np.empty((h, w, channels), np.uint8)
        _check(lib().rayon_job_resolve_u8(self.ctx._h, self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          channels, gamma), "rayon_job_resolve_u8")
        return out

    def resolve_u8_noalloc(self, channels: int = 3, gamma: float = 2.0) -> np.ndarray:
        """As resolve_u8, staged in the library's fixed device array (no device allocation)."""
        h, w = self.shape
        out = np.empty((h, w, channels), np.uint8)
        _check(lib().rayon_job_resolve_u8_noalloc(self.ctx._h, self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                  channels, gamma), "rayon_job_resolve_u8_noalloc")
        return out

    def write_png(self, path: str, gamma: float = 2.0):
        """Resolves the image (gamma-corrected RGB) and writes it as a PNG file."""
        _check(lib().rayon_job_write_png(self.ctx._h, self._h, str(path).encode(), gamma),
               "rayon_job_write_png")

    def sample_counts(self) -> np.ndarray:
        h, w = self.shape
        out = np.empty((h, w), np.int32)
        _check(lib().rayon_job_sample_counts(self.ctx._h, self._h, out.ctypes.data_as(C.POINTER(C.c_int32))),
               "rayon_job_sample_counts")
        return out

    def stats(self) -> JobStats:
        s = JobStats()
        _check(lib().rayon_job_stats_get(self.ctx._h, self._h, C.byref(s)), "rayon_job_stats_get")
        return s

    def primary_rays(self, sample: int = 0):
        h, w = self.shape
        o = np.empty((h * w, 3), np.float32)
        d = np.empty((h * w, 3), np.float32)
        _check(lib().rayon_job_primary_rays(self.ctx._h, self._h, sample, o.ctypes.data_as(C.POINTER(C.c_float)),
                                            d.ctypes.data_as(C.POINTER(C.c_float))), "rayon_job_primary_rays")
        return o, d

    def close(self):
        if self._h:
            lib().rayon_job_destroy(self._h)
            self._h = None


def server_config(**kw) -> ServerConfig:
    c = ServerConfig()
    lib().rayon_server_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


class Server:
    """Continuous-batching render server on one context (a native thread in the library)."""

    def __init__(self, ctx: Context, **cfg):
        self.ctx = ctx
        self.config = server_config(**cfg)
        h = C.c_void_p()
        _check(lib().rayon_server_create(ctx._h, C.byref(self.config), C.byref(h)), "rayon_server_create")
        self._h = h
        self._shape = {}

    def submit(self, scene: Scene, params: JobParams) -> int:
        i = C.c_uint64()
        _check(lib().rayon_server_submit(self._h, scene._h, C.byref(params), C.byref(i)), "rayon_server_submit")
        self._shape[i.value] = (params.height, params.width)
        return i.value

    def info(self, job_id: int) -> ServerJobInfo:
        s = ServerJobInfo()
        _check(lib().rayon_server_job_info_get(self._h, job_id, C.byref(s)), "rayon_server_job_info_get")
        return s

    def result(self, job_id: int) -> np.ndarray:
        h, w = self._shape[job_id]
        out = np.empty((h, w, 3), np.float32)
        _check(lib().rayon_server_result(self._h, job_id, out.ctypes.data_as(C.POINTER(C.c_float))),
               "rayon_server_result")
        return out

    def first_frame(self, job_id: int) -> np.ndarray:
        h, w = self._shape[job_id]
        out = np.empty((h, w, 4), np.uint8)
        _check(lib().rayon_server_first_frame(self._h, job_id, out.ctypes.data_as(C.POINTER(C.c_uint8))),
               "rayon_server_first_frame")
        return out

    def cancel(self, job_id: int):
        _check(lib().rayon_server_cancel(self._h, job_id), "rayon_server_cancel")

    def release(self, job_id: int):
        _check(lib().rayon_server_release(self._h, job_id), "rayon_server_release")

    def drain(self, timeout_s: float = 60.0):
        _check(lib().rayon_server_drain(self._h, timeout_s), "rayon_server_drain")

    def stats(self) -> ServerStats:
        s = ServerStats()
        _check(lib().rayon_server_stats_get(self._h, C.byref(s)), "rayon_server_stats_get")
        return s

    def close(self):
        if self._h:
            lib().rayon_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Progressive:
    """Headless progressive renderer of one job (frame-time batch controller, optional
    progressive depth, display resolved on a second stream into pinned memory)."""

    def __init__(self, ctx: Context, scene: Scene, params: JobParams, **cfg):
        self.ctx, self.params = ctx, params
        c = ProgressiveConfig()
        lib().rayon_progressive_config_default(C.byref(c))
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise AttributeError(k)
            setattr(c, k, v)
        self.config = c
        h = C.c_void_p()
        _check(lib().rayon_progressive_create(ctx._h, scene._h, C.byref(params), C.byref(c), C.byref(h)),
               "rayon_progressive_create")
        self._h = h
        self.job = Job(ctx, scene, C.c_void_p(lib().rayon_progressive_job(h)), params)

    def step(self) -> ProgressiveFrame:
        f = ProgressiveFrame()
        _check(lib().rayon_progressive_step(self._h, C.byref(f)), "rayon_progressive_step")
        return f

    def display(self):
        out = np.empty((self.params.height, self.params.width, self.config.channels), np.uint8)
        n = C.c_int64()
        _check(lib().rayon_progressive_display(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n)),
               "rayon_progressive_display")
        return out, n.value

    def reset(self, camera: Camera = None):
        _check(lib().rayon_progressive_reset(self._h, C.byref(camera) if camera is not None else None),
               "rayon_progressive_reset")

    def close(self):
        if self._h:
            self.job._h = None  # owned by the renderer
            lib().rayon_progressive_destroy(self._h)
            self._h = None
