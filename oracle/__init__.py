"""ctypes binding of liboracle.so — the FP64 CPU restatement used as the checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from rayon_amd.abi import Camera, JobParams, Material, SceneDesc

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
# The benchmark's CPU baseline: the same restatement built in FP32 with the GPU path's tree
# rule and the reference CPU renderer's row-chunk schedule (rayon_amd/build.py). Timed by
# bench.py; never the checker.
CPU_BASE_PATH = Path(__file__).resolve().parent / "libcpubase.so"
_libs = {}


def cpu_baseline_lib() -> C.CDLL:
    return lib(CPU_BASE_PATH)


def lib(path: Path = LIB_PATH) -> C.CDLL:
    if path not in _libs:
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run `python -m rayon_amd.build --oracle`")
        L = C.CDLL(str(path))
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        P = C.POINTER
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [P(SceneDesc)]
        L.oracle_scene_free.argtypes = [vp]
        L.oracle_scene_num_prims.restype = i64
        L.oracle_scene_num_prims.argtypes = [vp]
        L.oracle_scene_num_lights.restype = i32
        L.oracle_scene_num_lights.argtypes = [vp]
        L.oracle_trace_closest.argtypes = [vp, P(C.c_float), P(C.c_float), i64, C.c_double, C.c_double, i32,
                                           P(i64), P(C.c_double)]
        L.oracle_intersect_prim.restype = C.c_double
        L.oracle_intersect_prim.argtypes = [vp, i64, P(C.c_float), P(C.c_float), C.c_double, C.c_double]
        L.oracle_trace_any.argtypes = [vp, P(C.c_float), P(C.c_float), i64, C.c_double, C.c_double, P(i32)]
        L.oracle_render.argtypes = [vp, P(JobParams), P(Camera), P(i64), i64, i32, i32, i32, P(C.c_double),
                                    P(C.c_uint32), P(i64)]
        L.oracle_render_samples.argtypes = [vp, P(JobParams), P(Camera), P(i64), i64, i32, i32, i32,
                                            P(C.c_double)]
        L.oracle_ref_rng_init.restype = C.c_uint32
        L.oracle_ref_rng_init.argtypes = [i64]
        L.oracle_ref_rng_next.restype = C.c_float
        L.oracle_ref_rng_next.argtypes = [P(C.c_uint32)]
        L.oracle_counter_rng_init.restype = C.c_uint32
        L.oracle_counter_rng_init.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.oracle_counter_rng_next.restype = C.c_float
        L.oracle_counter_rng_next.argtypes = [P(C.c_uint32)]
        L.oracle_camera_frame.argtypes = [P(C.c_double), P(C.c_double), P(C.c_double), C.c_double, i32, i32,
                                          P(Camera)]
        L.oracle_adaptive_ref.restype = i32
        L.oracle_adaptive_ref.argtypes = [i32, P(C.c_double), P(C.c_double), i32, i32, C.c_double]
        L.oracle_adaptive_stderr.restype = i32
        L.oracle_adaptive_stderr.argtypes = [C.c_double, C.c_double, i64, i32, C.c_double]
        L.oracle_gamma_u8.restype = C.c_uint8
        L.oracle_gamma_u8.argtypes = [C.c_double, i64, C.c_double]
        L.oracle_scatter_probe.restype = i32
        L.oracle_scatter_probe.argtypes = [P(Material), C.c_uint32, C.c_float, C.c_float, C.c_float,
                                           P(C.c_double), P(C.c_double), i32, P(C.c_double), P(C.c_float), i32,
                                           P(C.c_double), P(C.c_double), P(C.c_double), P(i32)]
        d3 = P(C.c_double)
        L.oracle_set_fault.argtypes = [i32]
        L.oracle_sample_vndf.argtypes = [d3, C.c_double, C.c_double, d3, i64, d3]
        L.oracle_ggx_lambda.restype = C.c_double
        L.oracle_ggx_lambda.argtypes = [d3, C.c_double, C.c_double]
        L.oracle_fresnel_dielectric.restype = C.c_double
        L.oracle_fresnel_dielectric.argtypes = [C.c_double, C.c_double]
        L.oracle_fresnel_conductor.restype = C.c_double
        L.oracle_fresnel_conductor.argtypes = [C.c_double, C.c_double, C.c_double]
        L.oracle_unit_disk.restype = i64
        L.oracle_unit_disk.argtypes = [P(C.c_float), i64, i64, d3]
        L.oracle_scatter_batch.argtypes = [P(Material), C.c_uint32, C.c_float, C.c_float, C.c_float, d3, d3, i32,
                                           d3, P(C.c_float), i32, i64, d3, d3, P(i32)]
        _libs[path] = L
    return _libs[path]


class fault:
    """Context manager planting one of the oracle's known mistakes (tests of the tests)."""

    VNDF_NO_WARP, VNDF_SWAP_ALPHA, CONDUCTOR_RP_SIGN, DIELECTRIC_RP, DISK_NO_REJECT, LAMBDA_HALF = 1, 2, 3, 4, 5, 6

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        lib().oracle_set_fault(self.which)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_fault(0)
        return False


def sample_vndf(wo, ax, ay, u):
    """VNDF normals for the (n, 2) uniforms u."""
    w = np.ascontiguousarray(wo, np.float64)
    uu = np.ascontiguousarray(u, np.float64)
    out = np.empty((uu.shape[0], 3), np.float64)
    lib().oracle_sample_vndf(_dp(w), ax, ay, _dp(uu), uu.shape[0], _dp(out))
    return out


def unit_disk(rnd, n):
    r = np.ascontiguousarray(rnd, np.float32)
    out = np.empty((n, 2), np.float64)
    used = lib().oracle_unit_disk(_fp(r), r.size, n, _dp(out))
    return out, used


def scatter_batch(m: Material, dir_, normal, rnd, front=True, tangent=(1.0, 0.0, 0.0), flags=0, fuzz=0.8,
                  glass_ior=1.5, light=1.0):
    """Scatters `len(rnd)` samples of one incident direction; rnd is (count, k) uniforms."""
    r = np.ascontiguousarray(rnd, np.float32)
    count, k = r.shape
    d = np.ascontiguousarray(dir_, np.float64)
    nn = np.ascontiguousarray(normal, np.float64)
    t = np.ascontiguousarray(tangent, np.float64)
    od = np.empty((count, 3), np.float64)
    at = np.empty((count, 3), np.float64)
    ok = np.empty(count, np.int32)
    lib().oracle_scatter_batch(C.byref(m), flags, fuzz, glass_ior, light, _dp(d), _dp(nn), int(front), _dp(t),
                               _fp(r), k, count, _dp(od), _dp(at), ok.ctypes.data_as(C.POINTER(C.c_int32)))
    return ok.astype(bool), od, at


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleScene:
    def __init__(self, desc_ptr, library: C.CDLL = None):
        self._L = library or lib()
        self._h = self._L.oracle_scene_create(desc_ptr)

    @property
    def num_prims(self) -> int:
        return self._L.oracle_scene_num_prims(self._h)

    @property
    def num_lights(self) -> int:
        return self._L.oracle_scene_num_lights(self._h)

    def trace_closest(self, origins, dirs, t_min=0.001, t_max=3.4e38, brute=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        n = o.shape[0]
        prim = np.empty(n, np.int64)
        t = np.empty(n, np.float64)
        self._L.oracle_trace_closest(self._h, _fp(o), _fp(d), n, t_min, t_max, int(brute),
                                   prim.ctypes.data_as(C.POINTER(C.c_int64)), _dp(t))
        return prim, t

    def intersect_prim(self, prim, origin, d, t_min=0.001, t_max=3.4e38) -> float:
        o = np.ascontiguousarray(origin, np.float32)
        dd = np.ascontiguousarray(d, np.float32)
        return self._L.oracle_intersect_prim(self._h, int(prim), _fp(o), _fp(dd), t_min, t_max)

    def trace_any(self, origins, dirs, t_min, t_max):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.empty(o.shape[0], np.int32)
        self._L.oracle_trace_any(self._h, _fp(o), _fp(d), o.shape[0], t_min, t_max,
                               out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def render(self, params: JobParams, pixels=None, sample0=0, nsamples=1, threads=0, camera: Camera = None,
               rng_state=None):
        """Sums of nsamples samples per listed pixel (all pixels if None): (npix, 3) float64."""
        if pixels is None:
            pix = None
            npix = params.width * params.height
        else:
            pix = np.ascontiguousarray(pixels, np.int64)
            npix = pix.shape[0]
        sums = np.zeros((npix, 3), np.float64)
        segs = C.c_int64()
        rs = None
        if rng_state is not None:
            rs = rng_state.ctypes.data_as(C.POINTER(C.c_uint32))
        self._L.oracle_render(self._h, C.byref(params), C.byref(camera) if camera is not None else None,
                            pix.ctypes.data_as(C.POINTER(C.c_int64)) if pix is not None else None, npix, sample0,
                            nsamples, threads, _dp(sums), rs, C.byref(segs))
        return sums, segs.value

    def render_samples(self, params: JobParams, pixels, sample0, nsamples, threads=0):
        pix = np.ascontiguousarray(pixels, np.int64)
        out = np.zeros((pix.shape[0], nsamples, 3), np.float64)
        self._L.oracle_render_samples(self._h, C.byref(params), None, pix.ctypes.data_as(C.POINTER(C.c_int64)),
                                    pix.shape[0], sample0, nsamples, threads, _dp(out))
        return out

    def close(self):
        if self._h:
            self._L.oracle_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
