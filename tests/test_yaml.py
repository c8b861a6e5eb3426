"""Scene-file loader (rayon_load_yaml) against the reference's own loader.

Where the reference's sources are present, oracle/build_ref.sh compiles its loader
(yaml_scene_loader.cc with obj_loader.hpp) into oracle/_ref/ref_yaml_dump, which writes
what the reference loads. Every field must match: materials exactly after the FP32 cast
the reference applies at upload (scene_builder_cuda.cu:28-65), geometry coordinates the
same way, camera and settings exactly. The scenes are the reference's shipped files
(read in place) and synthetic files written here to exercise the parser's corner cases.
Without the reference (as on a machine that only has this repository), the synthetic
cases still run against expectations derived from the format's rules.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import rayon_amd as ra

ROOT = Path(__file__).resolve().parents[1]
REF_TOOL = ROOT / "oracle" / "_ref" / "ref_yaml_dump"
REF_SCENES = Path(os.environ.get("RAYON_REFERENCE", "/root/reference")) / "resources" / "scenes"

f32 = np.float32


def ref_load(path, tmp_path):
    out = tmp_path / "ref.bin"
    subprocess.run([str(REF_TOOL), str(path), str(out)], check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, timeout=600)
    a = np.fromfile(out, np.float64)
    nm, ng, ok = int(a[0]), int(a[1]), a[2] == 1.0
    cam = a[3:20]
    mats = a[20:20 + 26 * nm].reshape(nm, 26)
    geos = a[20 + 26 * nm:].reshape(ng, 22)
    return ok, cam, mats, geos


def ours(g):
    d = g.desc
    mats = np.zeros((d.num_materials, 26))
    for i in range(d.num_materials):
        m = d.materials[i]
        mats[i] = [m.type, *m.albedo, *m.emission, m.roughness, m.metallic, m.refractive_index, m.transmission,
                   m.anisotropy, *m.eta, *m.k, m.film_thickness, m.film_ior, m.pattern, *m.pattern_color,
                   m.pattern_param1, m.pattern_param2]
    n = d.num_geometries
    geos = np.zeros((n, 22))
    if n:
        raw = np.ctypeslib.as_array(C.cast(d.geometries, C.POINTER(C.c_byte)), shape=(n * C.sizeof(ra.Geometry),))
        rec = raw.view(np.dtype([("type", "<i4"), ("mat", "<i4"), ("vis", "<i4"), ("hn", "<i4"), ("p", "<f4", 9),
                                 ("n", "<f4", 9)]))
        geos[:, 0], geos[:, 1], geos[:, 2], geos[:, 3] = rec["type"], rec["mat"], rec["vis"], rec["hn"]
        geos[:, 4:13], geos[:, 13:22] = rec["p"], rec["n"]
    lf, la, up, fov = g.camera()
    s = g.settings()
    cam = np.array([*lf, *la, *up, fov, *s["background_color"], s["background_intensity"], s["ambient_light"],
                    s["use_bvh"], s["adaptive_sampling"]])
    return cam, mats, geos


def assert_same(path, tmp_path, flags=0):
    ok, rcam, rmats, rgeos = ref_load(path, tmp_path)
    if not ok:
        with pytest.raises(ra.RayonError):
            ra.Generated.load_yaml(str(path), flags)
        return None
    g = ra.Generated.load_yaml(str(path), flags)
    cam, mats, geos = ours(g)
    # camera position/look_at/up are doubles in both; fov and settings are floats
    np.testing.assert_array_equal(cam[:9], rcam[:9])
    np.testing.assert_array_equal(cam[9:].astype(f32), rcam[9:].astype(f32))
    assert mats.shape == rmats.shape, (mats.shape, rmats.shape)
    np.testing.assert_array_equal(mats.astype(f32), rmats.astype(f32))
    assert geos.shape == rgeos.shape, (geos.shape, rgeos.shape)
    np.testing.assert_array_equal(geos[:, :4], rgeos[:, :4])
    np.testing.assert_array_equal(geos[:, 4:].astype(f32), rgeos[:, 4:].astype(f32))
    return g


needs_ref = pytest.mark.skipif(not REF_TOOL.exists() or not REF_SCENES.is_dir(),
                               reason="the reference's loader (oracle/_ref) or its scenes are not available here")


@needs_ref
@pytest.mark.parametrize("name", sorted(p.name for p in REF_SCENES.glob("*.yaml")) if REF_SCENES.is_dir() else [])
def test_shipped_scene_matches_reference_loader(name, tmp_path):
    g = assert_same(REF_SCENES / name, tmp_path)
    assert g is not None and g.desc.num_geometries > 0


CORNER_CASES = {
    "nested_pattern_block": """
materials:
  - name: "dots"
    type: "clear_coat"   # the nested block below replaces this type
    albedo: [0.9, 0.1, 0.1]
    pattern:
      type: "fibonacci_dots"
      color: [0.02, 0.02, 0.02]
      dot_count: 12
  - { name: flow_dots, type: lambertian, pattern.type: checkerboard, pattern.color: [1, 0, 0] }
geometry:
  - { type: sphere, material: dots, center: [0, 1, 0], radius: 0.5 }
  - { type: sphere, material: flow_dots, center: [0, 2, 0] }
""",
    "unknown_entries_and_visibility": """
camera:
  position: [1, 2, 3]
  fov: 33.3
settings:
  background_intensity: 0.25
  use_bvh: "true"
  adaptive_sampling: 0
materials:
  - name: a
    type: not_a_type
    emission: [1, 2]
  - name: light
    type: light
    color: [1.0, 0.5, 0.25]
    emission_intensity: 8
  - name: a
    type: mirror
geometry:
  - type: sphere
    material: nothing_by_that_name
    visible: false
  - type: rectangle
    material: light
    corner: [0, 4, 0]
    u: [1, 0, 0]
    v: [0, 0, 1]
  - type: cube
    material: a
    visible: false
  - type: triangle
    material: a
    v0: [0, 0, 0]
    v1: [1, 0, 0]
    v2: [0, 1, 0]
    n0: [0, 0, 1]
""",
    "presets_and_defaults": """
materials:
  - { name: g, type: anisotropic_metal, preset: gold, roughness: 0.3, anisotropy: 0.5 }
  - { name: s, type: anisotropic_metal, preset: "silver" }
  - { name: c, type: lambertian, preset: copper }
  - { name: t, type: thin_film, film_thickness: 420.5, film_ior: 1.4, refractive_index: 1.0 }
  - name: "quoted"
    type: "glass"
    refractive_index: 1.52e0
    transmission: 1
geometry:
  - { type: displaced_sphere, material: g, center: [0.1, 0.2, 0.3], radius: 0.7, displacement_scale: 0.1 }
  - { type: sphere, material: t }
  - { type: sphere, material: quoted, center: [1e-3, -2.5, 7] }
""",
}


OBJ = """# a quad with normals, a triangle without normal indices, a face with negative and
# out-of-range indices
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
vn 0 0 -1
vt 0 0
f 1//1 2//1 3//1 4//1
f 1 2 3
f -4/1/2 -3/1/2 -2/1/2
f 1 2 9
"""


def write_case(tmp_path, name, text):
    p = tmp_path / f"{name}.yaml"
    p.write_text(text)
    return p


@pytest.mark.parametrize("name", sorted(CORNER_CASES))
def test_corner_cases_match_reference_loader(name, tmp_path):
    p = write_case(tmp_path, name, CORNER_CASES[name])
    if REF_TOOL.exists():
        assert_same(p, tmp_path)
    else:
        ra.Generated.load_yaml(str(p))


def test_obj_faces_and_visibility_of_last_triangle(tmp_path):
    (tmp_path / "m.obj").write_text(OBJ)
    p = write_case(tmp_path, "obj", """
materials:
  - { name: m, type: lambertian }
geometry:
  - { type: obj, material: m, file: missing.obj, visible: false }
  - { type: obj, material: m, file: m.obj, position: [10, 0, 0], scale: [2, 2, 2], visible: false }
""")
    g = assert_same(p, tmp_path) if REF_TOOL.exists() else ra.Generated.load_yaml(str(p))
    _, _, geos = ours(g)
    # quad -> 2 triangles, plain face -> 1 (normal index 0 stands for a missing one), negative
    # indices -> 1, out-of-range face -> dropped
    assert len(geos) == 4
    assert list(geos[:, 3]) == [1, 1, 1, 1]
    np.testing.assert_array_equal(geos[0, 4:13], [10, 0, 0, 12, 0, 0, 12, 2, 0])
    np.testing.assert_array_equal(geos[3, 13:16], [0, 0, -1])
    # an entry's visible flag lands on the last geometry so far: for the OBJ, its last
    # triangle only (the missing file's flag found no geometry to land on)
    assert list(geos[:, 2]) == [1, 1, 1, 0]


def test_caps_and_lifting_them(tmp_path):
    mats = "\n".join(f"  - {{ name: m{i}, type: lambertian, albedo: [{i / 200}, 0, 0] }}" for i in range(120))
    geos = "\n".join(f"  - {{ type: sphere, material: m{i % 120}, center: [{i}, 0, 0], radius: 0.5 }}"
                     for i in range(1100))
    p = write_case(tmp_path, "big", f"materials:\n{mats}\ngeometry:\n{geos}\n")
    g = ra.Generated.load_yaml(str(p))
    # 100 materials; spheres naming m100..m119 are skipped; 1000 entries read at most
    assert g.desc.num_materials == 100
    assert g.desc.num_geometries == sum(1 for i in range(1000) if i % 120 < 100)
    if REF_TOOL.exists():
        assert_same(p, tmp_path)
    h = ra.Generated.load_yaml(str(p), ra.YAML.LIFT_CAPS)
    assert h.desc.num_materials == 120 and h.desc.num_geometries == 1100


def test_errors(tmp_path):
    with pytest.raises(ra.RayonError, match="IO"):
        ra.Generated.load_yaml(str(tmp_path / "absent.yaml"))
    p = write_case(tmp_path, "bad_number", "materials:\n  - { name: m, roughness: abc }\ngeometry:\n"
                                           "  - { type: sphere, material: m }\n")
    with pytest.raises(ra.RayonError, match="PARSE"):
        ra.Generated.load_yaml(str(p))
    p = write_case(tmp_path, "no_geometry", "materials:\n  - { name: m }\n")
    with pytest.raises(ra.RayonError, match="PARSE"):
        ra.Generated.load_yaml(str(p))


@needs_ref
def test_save_and_reload_is_lossless(tmp_path):
    for name in ("default_scene.yaml", "05_material_laboratory.yaml", "09_color_bleed_box.yaml"):
        g = ra.Generated.load_yaml(str(REF_SCENES / name))
        out = tmp_path / f"re_{name}"
        g.save_yaml(str(out))
        h = ra.Generated.load_yaml(str(out), ra.YAML.LIFT_CAPS)
        for a, b in zip(ours(g), ours(h)):
            np.testing.assert_array_equal(a.astype(f32), b.astype(f32))
