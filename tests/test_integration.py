"""The drop-in boundary from the reference's side: its own offline GPU renderer, unchanged.

oracle/build_ref.sh compiles the reference's RendererCUDA (src/rayon/gpu_renderers/
renderer_cuda_host.hpp:23-74, which calls renderPixelsCUDAAccumulative and the free* helpers
declared in renderer_cuda_device.cuh:8-67) and its scene-file loader from the reference
sources, together with the scene shim a RayON maintainer adds
(rayon_amd/integration/scene_builder_hip.cc, in place of scene_builder_cuda.cu), and links
them against librayon_hip.so into oracle/_ref/ref_cuda_host_drive. No stand-ins: the
reference headers are used as they are.

* CPU: the program links, and every symbol of the library it needs is one that
  include/rayon_legacy.h declares and librayon_hip.so exports.
* GPU: the program renders a scene file through the reference's host code and the library;
  its 8-bit image equals the library's native path on the same file (rayon_load_yaml, a job
  with the reference stream and quirks, five batches, rayon_job_resolve_u8) to within one
  level in at most 0.1% of the channel values (the host gamma conversion of
  render_utils.hpp:41-74 against the device's), and it reports W*H*spp rays.
"""
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import rayon_amd as ra

ROOT = Path(__file__).resolve().parent.parent
DRIVE = ROOT / "oracle" / "_ref" / "ref_cuda_host_drive"
# built by the same recipe from the reference's sources: present exactly where they were
REF_LOADER = ROOT / "oracle" / "_ref" / "ref_yaml_dump"

SCENE = """
camera:
  position: [0, 2.2, 7]
  look_at: [0, 0.8, 0]
  up: [0, 1, 0]
  fov: 40
materials:
  - { name: ground, type: lambertian, albedo: [0.5, 0.5, 0.5] }
  - { name: red, type: lambertian, albedo: [0.7, 0.15, 0.1] }
  - { name: steel, type: metal, albedo: [0.8, 0.8, 0.85], roughness: 0.0 }
  - { name: glass, type: glass, refractive_index: 1.5 }
  - { name: gold, type: anisotropic_metal, preset: gold, roughness: 0.35, anisotropy: 0.6 }
  - { name: coat, type: clear_coat, albedo: [0.1, 0.3, 0.8], roughness: 0.1 }
  - { name: lamp, type: light, color: [1.0, 0.9, 0.8], emission_intensity: 6 }
  - name: dots
    type: lambertian
    albedo: [0.9, 0.9, 0.9]
    pattern:
      type: fibonacci_dots
      color: [0.05, 0.05, 0.05]
      dot_count: 40
      dot_radius: 0.15
geometry:
  - { type: sphere, material: ground, center: [0, -1000, 0], radius: 1000 }
  - { type: sphere, material: glass, center: [0, 1, 0], radius: 1 }
  - { type: sphere, material: steel, center: [-2.2, 1, -0.5], radius: 1 }
  - { type: sphere, material: gold, center: [2.2, 1, -0.5], radius: 1 }
  - { type: displaced_sphere, material: dots, center: [-1.0, 0.45, 1.6], radius: 0.45 }
  - { type: sphere, material: coat, center: [1.1, 0.4, 1.7], radius: 0.4 }
  - { type: rectangle, material: lamp, corner: [-1.5, 4, -1.5], u: [3, 0, 0], v: [0, 0, 3] }
  - { type: triangle, material: red, v0: [-3, 0, -3], v1: [3, 0, -3], v2: [0, 3, -3.5] }
"""


def _need_driver():
    if not DRIVE.exists():
        if REF_LOADER.exists():
            pytest.fail(f"{DRIVE} was not built although the reference's loader was (run build())")
        pytest.skip("the reference's host renderer is built only where its sources exist")


def test_reference_host_renderer_links_against_the_library():
    _need_driver()
    ldd = subprocess.run(["ldd", str(DRIVE)], capture_output=True, text=True, check=True).stdout
    assert re.search(r"librayon_hip\.so => .*rayon_amd/librayon_hip\.so", ldd), ldd
    undef = subprocess.run(["nm", "-u", str(DRIVE)], capture_output=True, text=True, check=True).stdout.split()
    exported = set(subprocess.run(["nm", "-D", "--defined-only", str(ra.LIB_PATH)], capture_output=True, text=True,
                                  check=True).stdout.split())
    header = (ROOT / "include" / "rayon_legacy.h").read_text()
    needed = [s for s in undef if s in exported]
    assert {"renderPixelsCUDAAccumulative", "freeDeviceRandomStates", "freeDeviceAccumBuffer",
            "rayon_legacy_scene_create", "rayon_legacy_scene_destroy"} <= set(needed), needed
    for s in needed:
        assert re.search(rf"\b{s}\(", header), f"{s} is not declared in include/rayon_legacy.h"


@pytest.mark.gpu
def test_reference_host_renderer_renders_through_the_library(tmp_path):
    _need_driver()
    scene = tmp_path / "scene.yaml"
    scene.write_text(SCENE)
    W, H, spp = 160, 90, 20
    out = tmp_path / "ref.rgb"
    print(f"reference RendererCUDA through librayon_hip: {W}x{H}, {spp} spp", flush=True)
    p = subprocess.run([str(DRIVE), str(scene), str(W), str(H), str(spp), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    rays = int(re.search(r"rays (\d+)", p.stdout).group(1))
    assert rays == W * H * spp
    ref_img = np.fromfile(out, np.uint8).reshape(H, W, 3)
    # the library's native path on the same file, with the reference's stream and quirks
    g = ra.Generated.load_yaml(str(scene))
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    P = ra.job_params(W, H, cam, spp=spp, max_depth=16, seed=0, flags=ra.FLAG.REF_RNG | ra.FLAG.REF_QUIRKS,
                      light_intensity=1.0, background_intensity=1.0, metal_fuzziness=0.8, glass_ior=1.5)
    with ra.Context(0) as ctx:
        sc = ctx.scene(g.desc_ptr())
        job = ctx.job(sc, P)
        per, done = max(1, spp // 5), 0
        while done < spp:  # the five batches of renderer_cuda_host.hpp:35-50
            k = min(per, spp - done)
            job.render_plain(k)
            done += k
        img = job.resolve_u8(3, 2.0)
        job.close()
        sc.close()
    diff = np.abs(img.astype(int) - ref_img.astype(int))
    print(f"  max level difference {diff.max()}, differing values {(diff > 0).mean():.5f}", flush=True)
    assert ref_img.max() > 0 and ref_img.std() > 10  # a real image
    assert diff.max() <= 1 and (diff > 0).mean() <= 1e-3
