"""Pins the FP64 oracle (oracle/oracle.cpp) before it is trusted as the checker.

The reference ships no golden vectors for this path (its only test, src/rayon/tests.cpp,
checks ONB orthogonality). Every check here therefore comes from what the mathematics
fixes: closed-form intersections, reflection/refraction laws, Fresnel limits, the
white-furnace identity, estimator agreement (NEE vs BSDF sampling), brute force vs BVH,
and the reference's documented RNG seeding. Each check is chosen so that a plausible
mistake (dropped term, wrong sign, swapped operand) fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import rayon_amd as ra
from oracle import OracleScene, lib as olib


# ----------------------------------------------------------------------------- helpers
def make_desc(mats, geoms, meshes=(), keep=None):
    keep = keep if keep is not None else []
    M = (ra.Material * len(mats))(*mats)
    G = (ra.Geometry * max(1, len(geoms)))(*geoms)
    Ms = (ra.Mesh * max(1, len(meshes)))(*meshes)
    d = ra.SceneDesc()
    d.materials, d.num_materials = M, len(mats)
    d.geometries, d.num_geometries = G, len(geoms)
    d.meshes, d.num_meshes = Ms, len(meshes)
    d.sdf_smooth_k, d.sdf_max_steps, d.sdf_eps = 0.1, 256, 1e-4
    keep += [M, G, Ms]
    return d, keep


def mat(type_, albedo=(0.5, 0.5, 0.5), **kw):
    m = ra.Material()
    m.type = type_
    m.albedo[:] = albedo
    m.refractive_index = 1.0
    m.film_thickness, m.film_ior = 400.0, 1.33
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


def sphere(c, r, m, visible=1):
    g = ra.Geometry()
    g.type, g.material_id, g.visible = ra.GEOM.SPHERE, m, visible
    g.p[:4] = [*c, r]
    return g


def tri(a, b, c, m):
    g = ra.Geometry()
    g.type, g.material_id, g.visible = ra.GEOM.TRIANGLE, m, 1
    g.p[:] = [*a, *b, *c]
    return g


def rect(q, u, v, m):
    g = ra.Geometry()
    g.type, g.material_id, g.visible = ra.GEOM.RECTANGLE, m, 1
    g.p[:] = [*q, *u, *v]
    return g


def trace(scene_geoms, o, d, brute=True, tmin=0.001, mats=None):
    desc, keep = make_desc(mats or [mat(0)], scene_geoms)
    s = OracleScene(C.pointer(desc))
    p, t = s.trace_closest(np.array([o], np.float32), np.array([d], np.float32), t_min=tmin, brute=brute)
    return int(p[0]), float(t[0])


def probe(m, d, n, front=1, rnd=(), flags=0, tangent=(1, 0, 0), fuzz=0.8, gior=1.5, light=1.0):
    od, at, em = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
    used = C.c_int32()
    r = (C.c_float * max(1, len(rnd)))(*rnd)
    ok = olib().oracle_scatter_probe(C.byref(m), flags, fuzz, gior, light, (C.c_double * 3)(*d),
                                     (C.c_double * 3)(*n), front, (C.c_double * 3)(*tangent), r, len(rnd), od, at,
                                     em, C.byref(used))
    return bool(ok), np.array(od), np.array(at), np.array(em), used.value


# ----------------------------------------------------------------------------- RNG
def test_ref_rng_seeding_matches_reference_formula():
    # cuda_utils.cu:26 — state = (uint32)(1234 + idx * 747796405)
    L = olib()
    assert L.oracle_ref_rng_init(0) == 1234
    assert L.oracle_ref_rng_init(1) == (1234 + 747796405) % 2**32
    assert L.oracle_ref_rng_init(7) == (1234 + 7 * 747796405) % 2**32


def _py_pcg(state):
    # independent restatement of cuda_utils.cuh:31-35 in Python integers
    state = (state * 747796405 + 2891336453) % 2**32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) % 2**32
    return state, (word >> 22) ^ word


def test_ref_rng_stream_known_answers():
    st = C.c_uint32(olib().oracle_ref_rng_init(3))
    s = (1234 + 3 * 747796405) % 2**32
    for _ in range(64):
        f = olib().oracle_ref_rng_next(C.byref(st))
        s, r = _py_pcg(s)
        assert st.value == s
        assert f == np.float32(np.float32(r) * np.float32(2.0**-32))


def test_counter_rng_uniformity_and_independence():
    L = olib()
    vals = []
    for pix in range(200):
        for smp in range(50):
            st = C.c_uint32(L.oracle_counter_rng_init(42, pix, smp))
            vals.append([L.oracle_counter_rng_next(C.byref(st)) for _ in range(2)])
    v = np.array(vals)
    assert v.min() >= 0.0 and v.max() < 1.0
    assert abs(v.mean() - 0.5) < 0.01
    assert abs(v.var() - 1 / 12) < 0.005
    hist, _ = np.histogram(v[:, 0], bins=20, range=(0, 1))
    chi2 = ((hist - hist.mean()) ** 2 / hist.mean()).sum()
    assert chi2 < 45  # 19 dof, p ~ 1e-3
    assert abs(np.corrcoef(v[:, 0], v[:, 1])[0, 1]) < 0.03


# ----------------------------------------------------------------------------- camera
def test_camera_frame_closed_form():
    cam = ra.Camera()
    olib().oracle_camera_frame((C.c_double * 3)(0, 0, 0), (C.c_double * 3)(0, 0, -1), (C.c_double * 3)(0, 1, 0),
                               90.0, 4, 2, C.byref(cam))
    np.testing.assert_allclose(list(cam.delta_u), [1, 0, 0], atol=1e-12)
    np.testing.assert_allclose(list(cam.delta_v), [0, -1, 0], atol=1e-12)
    np.testing.assert_allclose(list(cam.pixel00), [-1.5, 0.5, -1], atol=1e-12)
    # the product's host camera (rayon_camera_frame) restates the same frame
    pc = ra.camera_frame((0, 0, 0), (0, 0, -1), (0, 1, 0), 90.0, 4, 2)
    np.testing.assert_allclose(list(pc.pixel00), list(cam.pixel00), atol=1e-12)
    np.testing.assert_allclose(list(pc.delta_v), list(cam.delta_v), atol=1e-12)


# ----------------------------------------------------------------------------- intersections
def test_sphere_closed_form():
    assert trace([sphere((0, 0, 0), 1.0, 0)], (0, 0, -5), (0, 0, 1)) == (0, pytest.approx(4.0))
    assert trace([sphere((0, 0, 0), 1.0, 0)], (0, 0, -5), (0, 0, 2))[1] == pytest.approx(2.0)  # unnormalised dir
    assert trace([sphere((0, 0, 0), 1.0, 0)], (0, 0, 0), (0, 0, 1))[1] == pytest.approx(1.0)  # far root from inside
    assert trace([sphere((0, 0, 0), 1.0, 0)], (0, 1.001, -5), (0, 0, 1))[0] == -1


def test_triangle_and_rectangle_closed_form():
    T = [tri((0, 0, 0), (1, 0, 0), (0, 1, 0), 0)]
    assert trace(T, (0.25, 0.25, -1), (0, 0, 1)) == (0, pytest.approx(1.0))
    assert trace(T, (0.75, 0.75, -1), (0, 0, 1))[0] == -1  # u + v > 1
    assert trace(T, (0.25, 0.25, -1), (1, 0, 0))[0] == -1  # parallel
    R = [rect((0, 0, 0), (2, 0, 0), (0, 3, 0), 0)]
    assert trace(R, (1, 1, 5), (0, 0, -1)) == (0, pytest.approx(5.0))
    assert trace(R, (2.5, 1, 5), (0, 0, -1))[0] == -1


def test_exact_tie_goes_to_larger_index():
    T = [tri((0, 0, 0), (1, 0, 0), (0, 1, 0), 0), tri((0, 0, 0), (1, 0, 0), (0, 1, 0), 0),
         tri((0, 0, 5), (1, 0, 5), (0, 1, 5), 0)]
    assert trace(T, (0.2, 0.2, -1), (0, 0, 1))[0] == 1
    assert trace(T, (0.2, 0.2, -1), (0, 0, 1), brute=False)[0] == 1


def test_inclusive_t_range():
    # t exactly t_min is accepted (cuda_raytracer.cuh:215-220 rejects only root < t_min)
    assert trace([tri((0, 0, 1), (1, 0, 1), (0, 1, 1), 0)], (0.2, 0.2, 0), (0, 0, 1), tmin=1.0)[0] == 0


def test_bvh_equals_brute_force():
    rng = np.random.default_rng(5)
    geoms = []
    for _ in range(600):
        c = rng.uniform(-5, 5, 3)
        geoms.append(tri(c, c + rng.normal(0, 0.5, 3), c + rng.normal(0, 0.5, 3), 0))
    for _ in range(100):
        geoms.append(sphere(rng.uniform(-5, 5, 3), rng.uniform(0.05, 0.6), 0))
    for _ in range(20):
        # orthogonal edges: the reference's projection test (cuda_raytracer.cuh:256-261)
        # describes the parallelogram spanned by u, v only when u is perpendicular to v
        u = rng.normal(0, 1, 3)
        v = np.cross(u, rng.normal(0, 1, 3))
        geoms.append(rect(rng.uniform(-5, 5, 3), u, v, 0))
    desc, keep = make_desc([mat(0)], geoms)
    s = OracleScene(C.pointer(desc))
    o = rng.uniform(-8, 8, (20000, 3)).astype(np.float32)
    d = rng.normal(0, 1, (20000, 3)).astype(np.float32)
    p1, t1 = s.trace_closest(o, d, brute=True)
    p2, t2 = s.trace_closest(o, d, brute=False)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(t1, t2)
    assert (p1 >= 0).mean() > 0.15


# ----------------------------------------------------------------------------- materials
def test_lambert_direction_closed_form():
    u1, u2 = 0.3, 0.7
    ok, od, at, em, used = probe(mat(ra.MAT.LAMBERTIAN, (0.2, 0.4, 0.6)), (0, 0, -1), (0, 0, 1), rnd=(u1, u2))
    u1, u2 = float(np.float32(u1)), float(np.float32(u2))  # the oracle receives FP32 draws
    r, th = math.sqrt(u1), 2 * math.pi * u2
    # ONB of cuda_utils.cuh:99-107 for n = +z: u = (0,-1,0), v = n x u = (1,0,0)
    exp = r * math.cos(th) * np.array([0, -1, 0]) + r * math.sin(th) * np.array([1, 0, 0]) + \
        math.sqrt(1 - u1) * np.array([0, 0, 1])
    assert ok and used == 2
    np.testing.assert_allclose(od, exp, atol=1e-12)
    np.testing.assert_allclose(at, [0.2, 0.4, 0.6], rtol=1e-6)
    # cosine weighting: E[cos] = 2/3
    rng = np.random.default_rng(1)
    cs = [probe(mat(0), (0, 0, -1), (0, 0, 1), rnd=tuple(rng.random(2)))[1][2] for _ in range(4000)]
    assert abs(np.mean(cs) - 2 / 3) < 0.01


def test_mirror_reflection_law():
    d = np.array([1.0, -2.0, 0.5])
    ok, od, *_ = probe(mat(ra.MAT.MIRROR), d, (0, 1, 0))
    dn = d / np.linalg.norm(d)
    np.testing.assert_allclose(od, [dn[0], -dn[1], dn[2]], atol=1e-12)
    assert ok


def test_glass_snell_tir_and_schlick():
    g = mat(ra.MAT.DIELECTRIC, (1, 1, 1), refractive_index=1.5)
    th = math.radians(40)
    d = (math.sin(th), -math.cos(th), 0.0)
    ok, od, at, em, used = probe(g, d, (0, 1, 0), front=1, rnd=(0.999,))
    assert ok and used == 1
    st = math.sqrt(od[0] ** 2 + od[2] ** 2) / np.linalg.norm(od)
    assert st == pytest.approx(math.sin(th) / 1.5, rel=1e-9)  # Snell
    np.testing.assert_allclose(at, [1, 1, 1])
    # from inside at 60 degrees: 1.5 sin 60 > 1 -> total internal reflection, no draw (Q20)
    th = math.radians(60)
    ok, od, at, em, used = probe(g, (math.sin(th), -math.cos(th), 0), (0, 1, 0), front=0, rnd=(0.999,))
    assert used == 0 and od[1] > 0
    # Schlick at normal incidence: reflect iff rnd < r0
    r0 = ((1 - 1 / 1.5) / (1 + 1 / 1.5)) ** 2
    assert probe(g, (0, -1, 0), (0, 1, 0), rnd=(r0 * 0.99,))[1][1] > 0
    assert probe(g, (0, -1, 0), (0, 1, 0), rnd=(r0 * 1.01,))[1][1] < 0
    # spec mode uses the material IOR, ref quirks the global one (Q3)
    g2 = mat(ra.MAT.DIELECTRIC, (1, 1, 1), refractive_index=2.0)
    th = math.radians(40)
    d = (math.sin(th), -math.cos(th), 0.0)
    od = probe(g2, d, (0, 1, 0), rnd=(0.999,))[1]
    assert math.hypot(od[0], od[2]) / np.linalg.norm(od) == pytest.approx(math.sin(th) / 2.0, rel=1e-9)
    od = probe(g2, d, (0, 1, 0), rnd=(0.999,), flags=ra.FLAG.REF_QUIRKS, gior=1.5)[1]
    assert math.hypot(od[0], od[2]) / np.linalg.norm(od) == pytest.approx(math.sin(th) / 1.5, rel=1e-9)


def test_rough_mirror_quirks():
    m = mat(ra.MAT.METAL, (0.7, 0.6, 0.5), roughness=0.0)
    rnd = (0.25, 0.5)  # on_sphere: theta = pi/2, phi = acos(0) -> s = (0, 1, 0)
    ok, od, *_ = probe(m, (0, -1, 0), (0, 0, 1), rnd=rnd)
    np.testing.assert_allclose(od, [0, -1, 0], atol=1e-12)  # spec: roughness 0 is a mirror
    ok, od, *_ = probe(m, (0, -1, 1e-9), (0, 0, 1), rnd=rnd, flags=ra.FLAG.REF_QUIRKS, fuzz=0.8)
    pn = np.array([0, 0.24, 1.0]) / math.hypot(0.24, 1.0)  # 0.3 * 0.8 perturbation (Q1, Q2)
    dn = np.array([0, -1, 1e-9]) / np.linalg.norm([0, -1, 1e-9])
    np.testing.assert_allclose(od, dn - 2 * dn.dot(pn) * pn, atol=1e-9)


def test_conductor_fresnel_normal_incidence():
    eta, k = (0.2, 0.9, 1.4), (3.9, 2.4, 1.6)
    m = mat(ra.MAT.GGX_CONDUCTOR, (1, 1, 1), roughness=0.0, eta=eta, k=k)
    ok, od, at, em, used = probe(m, (0, 0, -1), (0, 0, 1))
    eta, k = [float(np.float32(x)) for x in eta], [float(np.float32(x)) for x in k]  # stored as FP32
    exp = [((e - 1) ** 2 + kk ** 2) / ((e + 1) ** 2 + kk ** 2) for e, kk in zip(eta, k)]
    np.testing.assert_allclose(at, exp, rtol=1e-9)
    assert used == 0 and od[2] > 0


@pytest.mark.parametrize("kind", [ra.MAT.GGX_CONDUCTOR, ra.MAT.GGX_DIELECTRIC, ra.MAT.ANISOTROPIC_METAL])
def test_ggx_energy_bounded(kind):
    # white furnace on one scattering event: E[weight] <= 1 (single-scatter GGX loses energy
    # only through G2/G1 < 1), close to 1 for smooth surfaces.
    rng = np.random.default_rng(2)
    for rough, lo in ((0.15, 0.93), (0.7, 0.6)):
        m = mat(kind, (1, 1, 1), roughness=rough, eta=(1e-3, 1e-3, 1e-3), k=(1e3, 1e3, 1e3), refractive_index=1.5,
                anisotropy=0.5)
        th = math.radians(35)
        ws = []
        for _ in range(3000):
            ok, od, at, em, used = probe(m, (math.sin(th), 0, -math.cos(th)), (0, 0, 1), rnd=tuple(rng.random(3)),
                                         tangent=(1, 0, 0))
            ws.append(at[0] if ok else 0.0)
        w = np.mean(ws)
        assert lo < w <= 1.0 + 1e-9, (kind, rough, w)


def test_ggx_dielectric_limits_smooth_glass():
    # roughness -> 0: refraction obeys Snell's law through the (near-flat) microfacet
    m = mat(ra.MAT.GGX_DIELECTRIC, (1, 1, 1), roughness=0.01, refractive_index=1.5)
    th = math.radians(30)
    ok, od, at, em, used = probe(m, (math.sin(th), 0, -math.cos(th)), (0, 0, 1), rnd=(0.3, 0.6, 0.999))
    assert ok and od[2] < 0 and used == 3
    assert math.hypot(od[0], od[1]) / np.linalg.norm(od) == pytest.approx(math.sin(th) / 1.5, rel=2e-3)


def test_light_and_constant_emit():
    ok, od, at, em, used = probe(mat(ra.MAT.LIGHT, emission=(2, 3, 4)), (0, 0, -1), (0, 0, 1), light=1.5)
    assert not ok and list(em) == [3, 4.5, 6]
    ok, od, at, em, used = probe(mat(ra.MAT.CONSTANT, (0.1, 0.2, 0.3)), (0, 0, -1), (0, 0, 1))
    assert not ok
    np.testing.assert_allclose(em, [0.1, 0.2, 0.3], rtol=1e-6)


# ----------------------------------------------------------------------------- whole path
def _render(desc, W, H, cam_args, spp, depth, flags=0, seed=7, **kw):
    s = OracleScene(C.pointer(desc))
    cam = ra.camera_frame(*cam_args, W, H)
    P = ra.job_params(W, H, cam, spp=spp, max_depth=depth, seed=seed, flags=flags, **kw)
    sums, segs = s.render(P, nsamples=spp, camera=cam)
    return sums.reshape(H, W, 3) / spp, segs


def test_white_furnace_blue_channel():
    # Lambert albedo 1 in blue under the sky, whose blue channel is exactly 1
    # (cuda_raytracer.cuh:924-929): every sample of every pixel returns blue == 1.
    desc, keep = make_desc([mat(0, (0.3, 0.6, 1.0))], [sphere((0, 0, 0), 1.0, 0)])
    img, _ = _render(desc, 24, 16, ((0, 0, 4), (0, 0, 0), (0, 1, 0), 40.0), 8, 16)
    np.testing.assert_allclose(img[..., 2], 1.0, atol=1e-12)
    assert img[..., 0].min() < 0.5  # the sphere is in view


def test_constant_material_and_sky_closed_form():
    desc, keep = make_desc([mat(ra.MAT.CONSTANT, (0.2, 0.5, 0.7))], [sphere((0, 0, 0), 50.0, 0)])
    img, _ = _render(desc, 8, 8, ((0, 0, 0.5), (0, 0, -1), (0, 1, 0), 30.0), 4, 8)
    np.testing.assert_allclose(img.reshape(-1, 3), np.tile([0.2, 0.5, 0.7], (64, 1)), rtol=1e-6)
    # empty scene except a tiny far sphere behind the camera: the image is the sky
    desc, keep2 = make_desc([mat(0)], [sphere((0, 0, 100), 0.1, 0)])
    img, _ = _render(desc, 4, 4, ((0, 0, 0), (0, 0, -1), (0, 1, 0), 1.0), 1, 4)
    t = 0.5 * (0 + 1)  # looking along -z, y ~ 0
    np.testing.assert_allclose(img[..., 0], 1 - 0.5 * t, atol=5e-3)
    np.testing.assert_allclose(img[..., 1], 1 - 0.3 * t, atol=5e-3)


def _cornell_tiny():
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=1)
    return g


def test_nee_agrees_with_bsdf_sampling():
    # two different estimators of the same integral: next-event estimation (spec) and
    # pure BSDF sampling (reference semantics) must agree in expectation.
    g = _cornell_tiny()
    s = OracleScene(g.desc_ptr())
    lf, la, up, fov = g.camera()
    W, H = 8, 8
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    res = []
    for flags in (0, ra.FLAG.NEE):
        P = ra.job_params(W, H, cam, spp=1, max_depth=4, seed=3, flags=flags)
        sums, _ = s.render(P, nsamples=6000, camera=cam)
        res.append(sums.sum(0) / (6000 * W * H))
    a, b = res
    np.testing.assert_allclose(a, b, rtol=0.05)
    assert b.min() > 0.01


def test_gamma_and_adaptive_closed_form():
    L = olib()
    assert L.oracle_gamma_u8(0.25 * 10, 10, 2.0) == 128
    assert L.oracle_gamma_u8(5.0, 1, 2.0) == 255  # clamped at 0.999
    assert L.oracle_gamma_u8(-1.0, 4, 2.0) == 0
    a0 = (C.c_double * 3)(10, 10, 10)
    a1 = (C.c_double * 3)(20.0001, 20.0001, 20.0001)
    assert L.oracle_adaptive_ref(10, a0, a1, 10, 16, 0.01) == -20  # mean unchanged -> converged
    a2 = (C.c_double * 3)(30, 30, 30)
    assert L.oracle_adaptive_ref(10, a0, a2, 10, 16, 0.01) == 20
    assert L.oracle_adaptive_ref(4, a0, a1, 4, 16, 0.01) == 8  # below the minimum
    assert L.oracle_adaptive_ref(-12, a0, a1, 4, 16, 0.01) == -12  # already converged
    # std error of n identical samples is 0 -> converged; of 0/1 samples ~ 0.5/sqrt(n)
    assert L.oracle_adaptive_stderr(16.0, 16.0, 16, 16, 0.005) == 1
    assert L.oracle_adaptive_stderr(8.0, 8.0, 16, 16, 0.005) == 0


def _mixed_emitter_scene(emissive=True):
    """Lambertian floor lit by a quad light (sampled by NEE), an emissive golf-ball sphere
    and a CONSTANT sphere (both reached by BSDF sampling only)."""
    mats = [mat(0, (0.7, 0.7, 0.7)), mat(ra.MAT.LIGHT, emission=(4, 4, 4)),
            mat(ra.MAT.LIGHT if emissive else 0, (0.5, 0.5, 0.5), emission=(3, 2, 1)),
            mat(ra.MAT.CONSTANT if emissive else 0, (0.2, 0.9, 0.3))]
    golf = sphere((1.5, 0.6, 0.0), 0.5, 2)
    golf.type = ra.GEOM.DISPLACED_SPHERE
    geoms = [rect((-5, 0, -5), (10, 0, 0), (0, 0, 10), 0), rect((-1, 4, -1), (2, 0, 0), (0, 0, 2), 1), golf,
             sphere((-1.5, 0.5, 0.5), 0.4, 3)]
    return make_desc(mats, geoms)


def test_nee_counts_emitters_it_does_not_sample():
    # With NEE on, only the emitters NEE samples may drop their BSDF-hit emission after a
    # Lambertian bounce; the golf-ball light and the CONSTANT sphere must still light the
    # floor, so NEE and pure BSDF sampling agree (an estimator that dropped them would be
    # 10% or more darker, checked below).
    W, H, n = 8, 8, 4000
    cam_args = ((0, 3, 6), (0, 0.3, 0), (0, 1, 0), 50.0)
    res = {}
    for emissive in (True, False):
        desc, keep = _mixed_emitter_scene(emissive)
        s = OracleScene(C.pointer(desc))
        assert s.num_lights == 1
        cam = ra.camera_frame(*cam_args, W, H)
        for flags in (0, ra.FLAG.NEE):
            P = ra.job_params(W, H, cam, spp=1, max_depth=4, seed=5, flags=flags, background_intensity=0.0)
            sums, _ = s.render(P, nsamples=n, camera=cam)
            res[emissive, flags] = sums.sum(0) / (n * W * H)
    np.testing.assert_allclose(res[True, 0], res[True, ra.FLAG.NEE], rtol=0.04)
    np.testing.assert_allclose(res[False, 0], res[False, ra.FLAG.NEE], rtol=0.04)
    assert (res[True, ra.FLAG.NEE] / res[False, ra.FLAG.NEE]).max() > 1.1
