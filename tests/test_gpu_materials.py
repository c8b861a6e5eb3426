"""GPU parity of the material models and of the compensated accumulation, sample by sample.

* Material probe: the device's scatter_material (run through rayon_scatter_probe on given
  hits and random numbers) against the oracle's FP64 restatement (oracle_scatter_probe),
  which tests/test_oracle_pins.py holds against closed forms and integrated BSDFs.
  Tolerances: the same random numbers drawn (count equal); scattered/absorbed decisions
  equal except at most 0.2% of samples, each of which must be a threshold tie (the
  oracle's decision flips when its random numbers move by 1e-5); directions within 2e-4
  (absolute, unit vectors) and attenuations / emission within 2e-4 relative (1e-5
  absolute for the small weights of grazing microfacet samples) otherwise.
* Accumulation: the device's per-pixel compensated (float-float) sums of a long series, and of whole
  renders of 2^20 identical samples, equal the exact sums to 2 ulp, where a plain FP32
  running sum drifts by hundreds of ulp.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import rayon_amd as ra
from oracle import lib as olib

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mat(kind, **kw):
    m = ra.Material()
    m.type = kind
    m.albedo[:] = kw.pop("albedo", (0.8, 0.6, 0.4))
    m.refractive_index = 1.5
    m.film_thickness, m.film_ior = 380.0, 1.33
    for k, v in kw.items():
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = v
        else:
            setattr(m, k, v)
    return m


MATERIALS = {
    "lambert": _mat(ra.MAT.LAMBERTIAN),
    "mirror": _mat(ra.MAT.MIRROR),
    "metal": _mat(ra.MAT.METAL, roughness=0.3),
    "rough_mirror": _mat(ra.MAT.ROUGH_MIRROR, roughness=0.6),
    "glass": _mat(ra.MAT.GLASS),
    "dielectric": _mat(ra.MAT.DIELECTRIC, refractive_index=1.7),
    "light": _mat(ra.MAT.LIGHT, emission=(3.0, 2.0, 1.0)),
    "constant": _mat(ra.MAT.CONSTANT),
    "show_normals": _mat(ra.MAT.SHOW_NORMALS),
    "aniso_metal": _mat(ra.MAT.ANISOTROPIC_METAL, roughness=0.5, anisotropy=0.7, eta=(0.2, 0.9, 1.1),
                        k=(3.9, 2.4, 2.6)),
    "ggx_conductor": _mat(ra.MAT.GGX_CONDUCTOR, roughness=0.509, eta=(0.27, 0.68, 1.22), k=(3.6, 2.6, 2.3)),
    "ggx_conductor_smooth": _mat(ra.MAT.GGX_CONDUCTOR, roughness=0.02, eta=(0.27, 0.68, 1.22), k=(3.6, 2.6, 2.3)),
    "ggx_dielectric": _mat(ra.MAT.GGX_DIELECTRIC, roughness=0.509),
    "ggx_dielectric_rough": _mat(ra.MAT.GGX_DIELECTRIC, roughness=0.8, refractive_index=2.4),
    "thin_film": _mat(ra.MAT.THIN_FILM),
    "clear_coat": _mat(ra.MAT.CLEAR_COAT, roughness=0.2),
}


def _hits(n, seed):
    """Random incident directions, normals facing them, tangents and front flags."""
    rng = np.random.default_rng(seed)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.where(((d * nrm).sum(1) > 0)[:, None], -d, d)  # the ray arrives against the normal
    d *= rng.uniform(0.5, 3.0, (n, 1))  # unnormalised, as camera rays are
    t = np.cross(nrm, rng.normal(size=(n, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    front = rng.integers(0, 2, n)
    return [x.astype(np.float32) for x in (d, nrm, t)] + [front.astype(np.int32)]


def _oracle_probe(m, d, n, t, front, rnd, flags, fuzz=0.8, ior=1.5, light=1.0):
    L = olib()
    N = len(d)
    ok = np.zeros(N, bool)
    od, at, em = np.zeros((N, 3)), np.zeros((N, 3)), np.zeros((N, 3))
    used = np.zeros(N, np.int32)
    dp = C.POINTER(C.c_double)
    for i in range(N):
        a, b, c = (np.ascontiguousarray(x[i], np.float64) for x in (d, n, t))
        r = np.ascontiguousarray(rnd[i], np.float32)
        o1, o2, o3 = np.zeros(3), np.zeros(3), np.zeros(3)
        u = C.c_int32()
        ok[i] = L.oracle_scatter_probe(C.byref(m), flags, fuzz, ior, light, a.ctypes.data_as(dp),
                                       b.ctypes.data_as(dp), int(front[i]), c.ctypes.data_as(dp),
                                       r.ctypes.data_as(C.POINTER(C.c_float)), r.size, o1.ctypes.data_as(dp),
                                       o2.ctypes.data_as(dp), o3.ctypes.data_as(dp), C.byref(u)) != 0
        od[i], at[i], em[i], used[i] = o1, o2, o3, u.value
    return ok, od, at, em, used


@pytest.mark.parametrize("flags", [0, ra.FLAG.REF_QUIRKS])
@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_scatter_probe_matches_oracle(ctx, name, flags):
    m = MATERIALS[name]
    N = 3000
    d, n, t, front = _hits(N, seed=sorted(MATERIALS).index(name))
    rnd = np.random.default_rng(7).random((N, 8)).astype(np.float32)
    log(f"scatter probe {name} flags={flags} n={N}")
    g_ok, g_d, g_a, g_e, g_u = ctx.scatter_probe(m, d, n, t, front, rnd, flags=flags)
    o_ok, o_d, o_a, o_e, o_u = _oracle_probe(m, d, n, t, front, rnd, flags)
    same = (g_ok == o_ok) & (g_u == o_u)
    bad = np.flatnonzero(~same)
    # every disagreement must be a threshold tie: the oracle changes its decision when its
    # random numbers move by 1e-5
    for i in bad:
        flipped = False
        for s in (-1, 1):
            r2 = np.clip(rnd[i:i + 1] + s * 1e-5, 0, 0.9999999).astype(np.float32)
            ok2, *_rest, u2 = _oracle_probe(m, d[i:i + 1], n[i:i + 1], t[i:i + 1], front[i:i + 1], r2, flags)
            flipped |= bool(ok2[0] != o_ok[i] or u2[0] != o_u[i])
        assert flipped, (name, i, g_ok[i], o_ok[i], g_u[i], o_u[i])
    assert len(bad) <= 0.002 * N, (name, len(bad))
    k = same & o_ok
    np.testing.assert_allclose(g_a[k], o_a[k], rtol=2e-4, atol=1e-5)
    dn = g_d[k] / np.linalg.norm(g_d[k], axis=1, keepdims=True)
    on = o_d[k] / np.linalg.norm(o_d[k], axis=1, keepdims=True)
    err = np.abs(dn - on).max() if k.any() else 0.0
    assert err < 2e-4, (name, err)
    np.testing.assert_allclose(g_e[same], o_e[same], rtol=2e-4, atol=1e-6)
    log(f"  {k.sum()} scattered samples agree, {len(bad)} threshold ties, max dir err {err:.2e}")


def _ulp(x):
    return np.spacing(np.abs(np.float32(x))).astype(np.float64)


def test_compensated_accumulation_series_is_exact(ctx):
    n = 1 << 20
    i = np.arange(n, dtype=np.float64)
    rgb = np.stack([np.full(n, 0.3), 0.1 + 1e-9 * i, 0.7 + 0.05 * np.sin(i)], 1).astype(np.float32)
    s, c = ctx.accumulate_probe(rgb)
    exact = rgb.astype(np.float64).sum(0)
    got = s[:3].astype(np.float64) + c[:3].astype(np.float64)
    naive = np.add.accumulate(rgb, axis=0, dtype=np.float32)[-1].astype(np.float64)
    log(f"exact {exact}, compensated {got}, naive FP32 {naive}")
    for ch in range(3):
        assert abs(float(np.float32(got[ch])) - exact[ch]) <= 2 * _ulp(exact[ch]), ch
        assert abs(naive[ch] - exact[ch]) > 100 * _ulp(exact[ch]), ch  # the plain sum drifts
    lum = (0.299 * rgb[:, 0].astype(np.float64) + 0.587 * rgb[:, 1] + 0.114 * rgb[:, 2])
    assert float(s[3]) + float(c[3]) == pytest.approx((lum * lum).sum(), rel=2e-6)


def _constant_scene():
    m = ra.Material()
    m.type = ra.MAT.CONSTANT
    m.albedo[:] = (0.3, 0.1, 0.7)
    g = ra.Geometry()
    g.type, g.material_id, g.visible = ra.GEOM.SPHERE, 0, 1
    g.p[:4] = [0.0, 0.0, 0.0, 100.0]  # the camera sits inside: every sample returns the colour
    M = (ra.Material * 1)(m)
    G = (ra.Geometry * 1)(g)
    Ms = (ra.Mesh * 1)()
    d = ra.SceneDesc()
    d.materials, d.num_materials = M, 1
    d.geometries, d.num_geometries = G, 1
    d.meshes, d.num_meshes = Ms, 0
    d.sdf_smooth_k, d.sdf_max_steps, d.sdf_eps = 0.1, 256, 1e-4
    return d, [M, G, Ms]


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_render_of_identical_samples_sums_exactly(ctx, path):
    desc, keep = _constant_scene()
    sc = ctx.scene(C.pointer(desc))
    sc.set_batch_path(ra.PATH.WAVEFRONT)
    W = H = 4
    cam = ra.camera_frame((0, 0, 0), (0, 0, -1), (0, 1, 0), 40.0, W, H)
    n = 1 << 20
    P = ra.job_params(W, H, cam, spp=n, max_depth=2, seed=3)
    job = ctx.job(sc, P)
    log(f"identical-sample render, {path}, {n} spp")
    for _ in range(4):
        if path == "plain":
            job.render_plain(n // 4)
        else:
            ctx.render_batched([job], n // 4)
    img = job.resolve_linear().reshape(-1, 3)
    exact = np.float32([0.3, 0.1, 0.7]).astype(np.float64)
    for ch in range(3):
        err = np.abs(img[:, ch].astype(np.float64) - exact[ch])
        assert err.max() <= 2 * _ulp(exact[ch]), (path, ch, err.max())
    naive = np.add.accumulate(np.full(n, np.float32(0.1)), dtype=np.float32)[-1] / n
    assert abs(naive - np.float32(0.1)) > 100 * _ulp(0.1)
    job.close()
    sc.close()
