"""CPU checks of the C-ABI library: it loads, exports every function its headers declare,
and its host-side parts (scene generators, camera, argument validation) behave. No GPU
compute is called here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import rayon_amd as ra

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "for", "while", "sizeof", "return", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


@pytest.mark.parametrize("header", sorted((ROOT / "include").glob("*.h")), ids=lambda p: p.name)
def test_library_exports_every_declared_function(header):
    lib = C.CDLL(str(ra.LIB_PATH))
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header.name} declares but the library does not export: {missing}"


def test_abi_version():
    assert ra.lib().rayon_abi_version() == 1


def test_book_scene_generator_is_deterministic_and_shaped():
    a = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 1984)
    b = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 1984)
    da, db = a.desc, b.desc
    n = da.num_geometries
    assert n == db.num_geometries and 400 < n < 490
    ga = [tuple(da.geometries[i].p[:4]) for i in range(n)]
    gb = [tuple(db.geometries[i].p[:4]) for i in range(n)]
    assert ga == gb
    assert ga[0] == (0.0, -1000.0, 0.0, 1000.0)
    small = [g for g in ga if g[3] == pytest.approx(0.2)]
    assert len(small) == n - 4
    types = [da.materials[da.geometries[i].material_id].type for i in range(n)]
    frac_lambert = types.count(ra.MAT.LAMBERTIAN) / n
    assert 0.7 < frac_lambert < 0.9
    c = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 7)
    assert [tuple(c.desc.geometries[i].p[:4]) for i in range(4)] != ga[:4]


def test_cornell_generator_counts():
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=3)
    d = g.desc
    assert d.num_geometries == 12  # 10 wall triangles + the quad light + the sphere
    assert d.num_meshes == 1 and d.meshes[0].num_triangles == 20 * 4 ** 3
    assert d.materials[d.geometries[10].material_id].type == ra.MAT.LIGHT
    # one GGX conductor and one GGX dielectric object, roughness each in [0.05, 0.8]
    m_ico = d.materials[d.meshes[0].material_id]
    m_ball = d.materials[d.geometries[11].material_id]
    assert d.geometries[11].type == ra.GEOM.SPHERE
    assert {m_ico.type, m_ball.type} == {ra.MAT.GGX_CONDUCTOR, ra.MAT.GGX_DIELECTRIC}
    assert 0.05 <= m_ico.roughness <= 0.8 and 0.05 <= m_ball.roughness <= 0.8
    ball = np.array(d.geometries[11].p[:4])
    assert np.linalg.norm(ball[:3] - np.array([278, 200, 278])) > 150 * 1.05 + ball[3]  # apart from the icosphere
    pos = np.ctypeslib.as_array(d.meshes[0].positions, (d.meshes[0].num_vertices * 3,)).reshape(-1, 3)
    r = np.linalg.norm(pos - np.array([278, 200, 278]), axis=1)
    assert 150 * 0.95 - 1e-3 <= r.min() and r.max() <= 150 * 1.05 + 1e-3  # displacement 0.05 * [-1,1]
    assert r.std() > 0.5


def test_sdf_terrain_generator_counts():
    g = ra.Generated.generate(ra.GEN.SDF_TERRAIN, 11, detail=317, count=64)
    d = g.desc
    assert d.num_sdf_prims == 64 and d.meshes[0].num_triangles == 2 * 316 * 316
    assert d.sdf_smooth_k == pytest.approx(0.1) and d.sdf_max_steps == 256


def test_generate_rejects_bad_kind():
    with pytest.raises(ra.RayonError, match="unknown kind"):
        ra.Generated.generate(99, 1)


def test_camera_frame_rejects_degenerate():
    with pytest.raises(ra.RayonError):
        ra.camera_frame((1, 1, 1), (1, 1, 1), (0, 1, 0), 40, 8, 8)


@pytest.mark.parametrize("kind,seed,detail", [(ra.GEN.BOOK_SPHERES, 1984, 0), (ra.GEN.CORNELL_ICOSPHERE, 7, 6),
                                              (ra.GEN.SDF_TERRAIN, 11, 101)])
def test_bvh_invariants(kind, seed, detail):
    g = ra.Generated.generate(kind, seed, detail=detail)
    info = ra.bvh_check(g.desc_ptr())  # also verifies the 8-wide quantized tree
    assert 0 < info.max_depth <= 32
    assert 0 < info.num_qnodes <= info.num_nodes
    assert info.num_prims > 0 and info.num_nodes == info.num_prims - 1 or info.num_nodes > 0


def test_bvh_depth_bound_on_degenerate_input():
    # 20000 coincident spheres: no split is possible by centroid, median splits keep the
    # depth within the device stack
    import ctypes as C
    from test_oracle import make_desc, mat, sphere
    desc, keep = make_desc([mat(0)], [sphere((1, 2, 3), 0.5, 0) for _ in range(20000)])
    info = ra.bvh_check(C.pointer(desc))
    assert info.max_depth <= 32 and info.num_prims == 20000
