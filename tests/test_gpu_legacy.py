"""GPU tests of the reference-named entry points (include/rayon_legacy.h), driven with the
reference's own call sequences: the offline mode-2 loop of renderer_cuda_host.hpp:35-60
(five batches, host float3 accumulation buffer) and the progressive mode-3 pattern
(device-resident sums, GPU gamma conversion, adaptive count buffer)."""
import ctypes as C

import numpy as np
import pytest

import rayon_amd as ra
from oracle import lib as olib

pytestmark = pytest.mark.gpu

D = C.c_double


def legacy():
    L = C.CDLL(str(ra.LIB_PATH))
    L.rayon_legacy_scene_create.restype = C.c_void_p
    L.rayon_legacy_scene_create.argtypes = [C.POINTER(ra.SceneDesc)]
    L.rayon_legacy_scene_destroy.argtypes = [C.c_void_p]
    L.renderPixelsCUDAAccumulative.restype = C.c_ulonglong
    L.renderPixelsCUDAAccumulative.argtypes = (
        [C.c_void_p, C.POINTER(C.c_float), C.c_void_p, C.c_int, C.c_int] + [D] * 12 + [C.c_int] * 3 +
        [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)] + [D] * 6 + [C.c_void_p, C.c_int, C.c_float])
    L.freeDeviceRandomStates.argtypes = [C.c_void_p]
    L.freeDeviceAccumBuffer.argtypes = [C.c_void_p]
    L.resetDeviceAccumBuffer.argtypes = [C.c_void_p, C.c_int]
    L.convertAccumToDisplayCUDA.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_float, C.c_void_p]
    L.allocateAdaptiveBuffer.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    L.resetAdaptiveBuffer.argtypes = [C.c_void_p, C.c_int]
    L.freeAdaptiveBuffer.argtypes = [C.c_void_p]
    L.countConvergedPixels.restype = C.c_int
    L.countConvergedPixels.argtypes = [C.c_void_p, C.c_int]
    L.renderSampleHeatmapCUDA.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_int, C.c_int]
    for n in ("setLightIntensity", "setBackgroundIntensity", "setMetalFuzziness", "setGlassRefractionIndex",
              "setDOFAperture", "setDOFFocusDistance"):
        getattr(L, n).argtypes = [C.c_float]
    L.setDOFEnabled.argtypes = [C.c_bool]
    return L


def call_render(L, scene, cam, W, H, n, depth, states, accp, host=None, counts=None, min_s=32, thr=0.01):
    args = [None, host.ctypes.data_as(C.POINTER(C.c_float)) if host is not None else None, scene, W, H,
            *cam.center, *cam.pixel00, *cam.delta_u, *cam.delta_v, n, 0, depth, C.byref(states), C.byref(accp),
            *cam.u, *cam.v, counts, min_s, thr]
    return L.renderPixelsCUDAAccumulative(*args)


@pytest.fixture(scope="module")
def book():
    g = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 1984)
    return g


def test_mode2_batches_equal_plain_reference_job(ctx, book):
    L = legacy()
    W, H, spp, depth = 80, 45, 10, 16
    scene = L.rayon_legacy_scene_create(book.desc_ptr())
    assert scene
    lf, la, up, fov = book.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    states, accp = C.c_void_p(), C.c_void_p()
    host = np.zeros(W * H * 3, np.float32)
    per = max(1, spp // 5)  # renderer_cuda_host.hpp:35
    done, rays = 0, 0
    while done < spp:
        k = min(per, spp - done)
        done += k
        rays += call_render(L, scene, cam, W, H, k, depth, states, accp, host=host)
    assert rays == W * H * spp
    # the same work as one plain job with the reference's stream and quirks
    sc = ctx.scene(book.desc_ptr())
    job = ctx.job(sc, ra.job_params(W, H, cam, spp=spp, max_depth=depth, flags=ra.FLAG.REF_RNG | ra.FLAG.REF_QUIRKS))
    job.render_plain(spp)
    ref = job.resolve_linear().reshape(-1) * spp
    np.testing.assert_allclose(host, ref, rtol=2e-6, atol=2e-6)
    # the gamma conversion of the reference (render_acc_kernel.cu:30-40) on the device sums
    img = np.empty(W * H * 3, np.uint8)
    L.convertAccumToDisplayCUDA(accp, img.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, 3, spp, 2.0, None)
    exp = np.array([olib().oracle_gamma_u8(float(v), spp, 2.0) for v in host], np.uint8)
    assert np.mean(np.abs(img.astype(int) - exp.astype(int)) <= 1) == 1.0
    L.freeDeviceRandomStates(states)
    L.freeDeviceAccumBuffer(accp)
    job.close()
    sc.close()
    L.rayon_legacy_scene_destroy(scene)


def test_mode3_adaptive_and_reset(ctx, book):
    L = legacy()
    W, H = 64, 36
    scene = L.rayon_legacy_scene_create(book.desc_ptr())
    lf, la, up, fov = book.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    states, accp, counts = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.allocateAdaptiveBuffer(C.byref(counts), W * H)
    assert counts.value
    L.setMetalFuzziness(1.0)  # mode 3 sets 1.0 (renderer_cuda_progressive_host.hpp:98)
    for _ in range(8):
        call_render(L, scene, cam, W, H, 16, 16, states, accp, counts=counts, min_s=32, thr=0.02)
    conv = L.countConvergedPixels(counts, W * H)
    assert 0 < conv <= W * H
    img = np.empty(W * H * 4, np.uint8)
    L.convertAccumToDisplayCUDA(accp, img.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, 4, 128, 2.0, counts)
    assert (img[3::4] == 255).all() and img[0::4].max() > 0
    heat = np.empty(W * H * 3, np.uint8)
    L.renderSampleHeatmapCUDA(counts, heat.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, 3, 128)
    assert heat.max() > 0
    # camera change: reset sums and counts (renderer_cuda_progressive_host.hpp:603-616)
    L.resetDeviceAccumBuffer(accp, W * H)
    L.resetAdaptiveBuffer(counts, W * H)
    assert L.countConvergedPixels(counts, W * H) == 0
    L.convertAccumToDisplayCUDA(accp, img.ctypes.data_as(C.POINTER(C.c_uint8)), W, H, 4, 1, 2.0, None)
    assert img[0::4].max() == 0
    L.setMetalFuzziness(0.8)
    L.freeAdaptiveBuffer(counts)
    L.freeDeviceRandomStates(states)
    L.rayon_legacy_scene_destroy(scene)


def test_null_scene_returns_zero():
    L = legacy()
    states, accp = C.c_void_p(), C.c_void_p()
    cam = ra.camera_frame((0, 0, 1), (0, 0, 0), (0, 1, 0), 40, 8, 8)
    assert call_render(L, None, cam, 8, 8, 4, 4, states, accp) == 0
    assert not states.value


@pytest.mark.parametrize("dof", [False, True])
def test_mode2_loop_matches_oracle_reference_stream(book, dof):
    """The offline mode-2 loop of renderer_cuda_host.hpp:35-60 (five batches into a host
    float3 buffer) against the FP64 oracle run with the reference's sequential stream
    (cuda_utils.cu:26, cuda_utils.cuh:24-36) and material quirks (REF_RNG | REF_QUIRKS), with
    the lens model of render_acc_kernel.cu:209-221 off and on. Tolerance: relative RMSE of
    the summed image <= 1% (measured ~1e-3: FP32 against FP64 on identical random numbers)."""
    from oracle import OracleScene
    L = legacy()
    W, H, spp, depth = 96, 54, 10, 16
    scene = L.rayon_legacy_scene_create(book.desc_ptr())
    lf, la, up, fov = book.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    aperture, focus = 0.25, 10.0
    L.setDOFEnabled(dof)
    L.setDOFAperture(aperture)
    L.setDOFFocusDistance(focus)
    try:
        states, accp = C.c_void_p(), C.c_void_p()
        host = np.zeros(W * H * 3, np.float32)
        per, done = max(1, spp // 5), 0
        while done < spp:
            k = min(per, spp - done)
            call_render(L, scene, cam, W, H, k, depth, states, accp, host=host)
            done += k
        L.freeDeviceRandomStates(states)
        L.freeDeviceAccumBuffer(accp)
    finally:
        L.setDOFEnabled(False)
        L.rayon_legacy_scene_destroy(scene)
    flags = ra.FLAG.REF_RNG | ra.FLAG.REF_QUIRKS | (ra.FLAG.DOF if dof else 0)
    P = ra.job_params(W, H, cam, spp=spp, max_depth=depth, flags=flags, light_intensity=1.0,
                      background_intensity=1.0, metal_fuzziness=0.8, glass_ior=1.5, dof_aperture=aperture,
                      dof_focus_distance=focus)
    osc = OracleScene(book.desc_ptr())
    rng = np.array([olib().oracle_ref_rng_init(i) for i in range(W * H)], np.uint32)
    sums, _ = osc.render(P, nsamples=spp, rng_state=rng)  # the FP32 camera of P, as the device casts it
    ref = sums.reshape(-1)
    err = float(np.sqrt(((host - ref) ** 2).sum() / (ref ** 2).sum()))
    print(f"mode 2 dof={dof}: rel RMSE vs oracle {err:.5f}", flush=True)
    assert err <= 0.01
    if dof:  # the lens really moved the rays: the pinhole oracle image differs
        P0 = ra.job_params(W, H, cam, spp=spp, max_depth=depth, flags=ra.FLAG.REF_RNG | ra.FLAG.REF_QUIRKS,
                           metal_fuzziness=0.8, glass_ior=1.5)
        rng0 = np.array([olib().oracle_ref_rng_init(i) for i in range(W * H)], np.uint32)
        pin, _ = osc.render(P0, nsamples=spp, rng_state=rng0)
        assert float(np.sqrt(((pin.reshape(-1) - ref) ** 2).sum() / (ref ** 2).sum())) > 0.05
