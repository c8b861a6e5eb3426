"""Pins the oracle's microfacet and lens samplers to what the mathematics fixes.

The oracle (oracle/oracle.cpp) restates the reference's GGX model
(src/rayon/gpu_renderers/microfacet_ggx.cuh:21-111), the spec's exact dielectric Fresnel
and the lens sampler random_in_unit_disk (cuda_raytracer.cuh:114-122). The reference ships
no vectors for any of them, so each is held here against an independent statement:

* the VNDF sampler against the analytic density of visible normals
  D_wo(m) = G1(wo) max(0, wo.m) D(m) / wo.z (Heitz 2018), by a chi-square test;
* Smith's Lambda through the normalisation of that density (it integrates to one only
  with the correct G1);
* the dielectric Fresnel against the textbook sine/tangent laws, Brewster's angle and
  total internal reflection;
* the conductor Fresnel (FrComplex) against Fresnel's equations in complex arithmetic
  (numpy), which reduce to the dielectric ones at k = 0 and give 1 at grazing incidence;
* the unit-disk sampler against uniform area density and the 4/pi acceptance rate;
* the whole one-bounce GGX estimator (sampler, weight G2/G1, Fresnel) against the
  numerically integrated BRDF / BTDF: E[f cos / pdf] = integral of f cos over directions.

Each check is also run with a planted mistake in the oracle (oracle.fault): a missing
hemisphere warp, swapped alpha stretch, a wrong Rp term, a disk sampler without
rejection, a Lambda without its 1/2. Every planted mistake must fail at least one check.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

import rayon_amd as ra
from oracle import fault, lib as olib, sample_vndf, scatter_batch, unit_disk


# ----------------------------------------------------------------------------- analytic model
def ggx_d(m, ax, ay):
    """Anisotropic GGX normal distribution (Trowbridge-Reitz), per unit solid angle."""
    t = (m[..., 0] / ax) ** 2 + (m[..., 1] / ay) ** 2 + m[..., 2] ** 2
    return np.where(m[..., 2] > 0, 1.0 / (math.pi * ax * ay * t * t), 0.0)


def smith_lambda(w, ax, ay):
    return 0.5 * (np.sqrt(1.0 + ((ax * w[..., 0]) ** 2 + (ay * w[..., 1]) ** 2) / (w[..., 2] ** 2)) - 1.0)


def fresnel_complex(cos_i, eta, k):
    """Unpolarised reflectance of a surface of relative index eta + i k (Fresnel's equations)."""
    cos_i = np.clip(np.asarray(cos_i, np.float64), 0.0, 1.0)
    n = eta + 1j * k
    sin2 = 1.0 - cos_i ** 2
    cos_t = np.sqrt(1.0 - sin2 / (n * n) + 0j)
    rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
    rp = (n * cos_i - cos_t) / (n * cos_i + cos_t)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def fresnel_textbook(cos_i, eta):
    """Dielectric reflectance from the sine and tangent laws; eta = n_i / n_t."""
    ti = np.arccos(np.clip(np.asarray(cos_i, np.float64), 0.0, 1.0))
    st = eta * np.sin(ti)
    out = np.ones_like(ti)
    ok = st < 1.0
    tt = np.arcsin(st[ok])
    a, b = ti[ok] - tt, ti[ok] + tt
    with np.errstate(divide="ignore", invalid="ignore"):
        rs = np.where(b > 0, np.sin(a) ** 2 / np.sin(b) ** 2, ((1 - eta) / (1 + eta)) ** 2)
        rp = np.where(b > 0, np.tan(a) ** 2 / np.tan(b) ** 2, ((1 - eta) / (1 + eta)) ** 2)
    out[ok] = 0.5 * (rs + rp)
    return out


def sph(theta, phi):
    s = np.sin(theta)
    return np.stack([s * np.cos(phi), s * np.sin(phi), np.cos(theta)], -1)


def hemisphere_grid(n_theta, n_phi, lower=False):
    """Midpoint grid over a hemisphere: directions and their solid-angle weights."""
    th = (np.arange(n_theta) + 0.5) * (0.5 * math.pi / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2 * math.pi / n_phi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = sph(T, P)
    if lower:
        w = w * np.array([1.0, 1.0, -1.0])
    dw = np.sin(T) * (0.5 * math.pi / n_theta) * (2 * math.pi / n_phi)
    return w, dw


# ----------------------------------------------------------------------------- VNDF
VNDF_CASES = [  # (alpha_x, alpha_y, theta_o deg, phi_o deg)
    (0.259, 0.259, 30.0, 0.0),   # the mesh benchmark icosphere's roughness 0.509
    (0.5, 0.5, 70.0, 40.0),
    (0.15, 0.6, 50.0, 30.0),
    (0.6, 0.15, 60.0, 120.0),
    (0.05, 0.05, 80.0, 10.0),
]


def vndf_density(m, wo, ax, ay):
    g1 = 1.0 / (1.0 + smith_lambda(wo, ax, ay))
    return g1 * np.maximum(0.0, m @ wo) * ggx_d(m, ax, ay) / wo[2]


def vndf_bins(wo, ax, ay, n_theta=20, n_phi=24, sub=12):
    """Bin edges in (theta, phi) and the expected probability of each bin, by midpoint
    integration of the visible-normal density on a sub-grid."""
    tmax = min(0.5 * math.pi, math.atan(6.0 * max(ax, ay)))
    t_edges = np.concatenate([np.linspace(0.0, tmax, n_theta + 1), [0.5 * math.pi]])
    t_edges = np.unique(t_edges)
    p_edges = np.linspace(0.0, 2 * math.pi, n_phi + 1)
    probs = np.zeros((len(t_edges) - 1, n_phi))
    for i in range(len(t_edges) - 1):
        t0, t1 = t_edges[i], t_edges[i + 1]
        ts = t0 + (np.arange(sub) + 0.5) * (t1 - t0) / sub
        for j in range(n_phi):
            p0, p1 = p_edges[j], p_edges[j + 1]
            ps = p0 + (np.arange(sub) + 0.5) * (p1 - p0) / sub
            T, P = np.meshgrid(ts, ps, indexing="ij")
            m = sph(T, P)
            probs[i, j] = (vndf_density(m, wo, ax, ay) * np.sin(T)).sum() * (t1 - t0) / sub * (p1 - p0) / sub
    return t_edges, p_edges, probs


def vndf_chi2(ax, ay, th_deg, ph_deg, n=300_000, seed=1):
    """(p-value, total expected probability) of the oracle's VNDF samples."""
    wo = sph(math.radians(th_deg), math.radians(ph_deg))
    u = np.random.default_rng(seed).random((n, 2))
    m = sample_vndf(wo, ax, ay, u)
    t_edges, p_edges, probs = vndf_bins(wo, ax, ay)
    th = np.arccos(np.clip(m[:, 2], -1, 1))
    ph = np.mod(np.arctan2(m[:, 1], m[:, 0]), 2 * math.pi)
    obs, _, _ = np.histogram2d(th, ph, bins=[t_edges, p_edges])
    exp = probs * n
    # bins expected to hold fewer than 10 samples are pooled into one
    small = exp < 10
    o = np.concatenate([obs[~small], [obs[small].sum()]])
    e = np.concatenate([exp[~small], [exp[small].sum()]])
    keep = e > 0
    chi2 = (((o - e) ** 2)[keep] / e[keep]).sum() + (o[~keep].sum() > 0) * 1e9
    dof = keep.sum() - 1
    return float(stats.chi2.sf(chi2, dof)), float(probs.sum())


@pytest.mark.parametrize("case", VNDF_CASES)
def test_vndf_matches_visible_normal_density(case):
    p, total = vndf_chi2(*case)
    assert abs(total - 1.0) < 2e-3, total  # the density integrates to one (pins Lambda / G1)
    assert p > 1e-4, (case, p)


@pytest.mark.parametrize("which,case", [(fault.VNDF_NO_WARP, VNDF_CASES[1]),
                                        (fault.VNDF_SWAP_ALPHA, VNDF_CASES[2]),
                                        (fault.VNDF_SWAP_ALPHA, VNDF_CASES[3])])
def test_vndf_check_catches_planted_mistakes(which, case):
    with fault(which):
        p, _ = vndf_chi2(*case)
    assert p < 1e-8, (which, case, p)


def test_lambda_closed_form():
    L = olib()
    rng = np.random.default_rng(3)
    for _ in range(200):
        w = rng.normal(size=3)
        w[2] = abs(w[2]) + 1e-3
        w /= np.linalg.norm(w)
        ax, ay = rng.uniform(0.02, 1.0, 2)
        got = L.oracle_ggx_lambda(np.ascontiguousarray(w).ctypes.data_as(C.POINTER(C.c_double)), ax, ay)
        assert got == pytest.approx(float(smith_lambda(w, ax, ay)), rel=1e-12, abs=1e-15)


# ----------------------------------------------------------------------------- Fresnel
def test_dielectric_fresnel_closed_forms():
    L = olib()
    F = np.vectorize(L.oracle_fresnel_dielectric)
    for eta in (1 / 1.5, 1.5, 1 / 2.4, 1.33):
        # normal incidence
        assert F(1.0, eta) == pytest.approx(((1 - eta) / (1 + eta)) ** 2, rel=1e-12)
        # grazing incidence reflects everything
        assert F(0.0, eta) == pytest.approx(1.0, abs=1e-12)
        # every angle against the sine / tangent laws
        ci = np.cos(np.radians(np.linspace(0.5, 89.5, 90)))
        np.testing.assert_allclose(F(ci, eta), fresnel_textbook(ci, eta), rtol=1e-9, atol=1e-12)
        # Brewster's angle: Rp = 0, so R = Rs / 2 with Rs = sin^2(ti - tt)
        tb = math.atan(1.0 / eta)
        tt = 0.5 * math.pi - tb
        assert F(math.cos(tb), eta) == pytest.approx(0.5 * math.sin(tb - tt) ** 2, rel=1e-9)
    # beyond the critical angle of a denser incident medium: total internal reflection
    eta = 1.5
    tc = math.asin(1 / eta)
    for th in np.linspace(tc + 1e-6, 0.5 * math.pi, 20):
        assert F(math.cos(th), eta) == 1.0
    assert F(math.cos(tc - 1e-3), eta) < 1.0


def _dielectric_fresnel_ok():
    F = np.vectorize(olib().oracle_fresnel_dielectric)
    ci = np.cos(np.radians(np.linspace(0.5, 89.5, 90)))
    return all(np.allclose(F(ci, eta), fresnel_textbook(ci, eta), rtol=1e-9, atol=1e-12) for eta in (1 / 1.5, 1.5))


def test_dielectric_fresnel_check_catches_planted_mistake():
    assert _dielectric_fresnel_ok()
    with fault(fault.DIELECTRIC_RP):
        assert not _dielectric_fresnel_ok()


CONDUCTORS = [(0.2, 3.9), (0.27, 2.78), (1.1, 2.6), (2.9, 3.0), (0.05, 0.5), (1.5, 0.0), (0.7, 0.0)]


def _conductor_ok(tol=1e-9):
    F = np.vectorize(olib().oracle_fresnel_conductor)
    ci = np.cos(np.radians(np.linspace(0.0, 89.9, 120)))
    return all(np.allclose(F(ci, e, k), fresnel_complex(ci, e, k), rtol=tol, atol=1e-12) for e, k in CONDUCTORS)


def test_conductor_fresnel_oblique_angles():
    F = np.vectorize(olib().oracle_fresnel_conductor)
    assert _conductor_ok()
    for e, k in CONDUCTORS:
        assert F(0.0, e, k) == pytest.approx(1.0, abs=1e-12)  # grazing
        assert F(1.0, e, k) == pytest.approx(((e - 1) ** 2 + k ** 2) / ((e + 1) ** 2 + k ** 2), rel=1e-12)
    # k = 0 is a dielectric of relative index eta (incident side index 1)
    ci = np.cos(np.radians(np.linspace(0.0, 89.0, 60)))
    for e in (1.5, 1.33, 2.4):
        np.testing.assert_allclose(F(ci, e, 0.0), fresnel_textbook(ci, 1.0 / e), rtol=1e-9, atol=1e-12)


def test_conductor_fresnel_check_catches_planted_mistake():
    with fault(fault.CONDUCTOR_RP_SIGN):
        assert not _conductor_ok()


# ----------------------------------------------------------------------------- lens disk
def _disk_check(n=200_000, seed=5):
    rnd = np.random.default_rng(seed).random(8 * n).astype(np.float32)
    xy, used = unit_disk(rnd, n)
    assert used > 0
    r2 = (xy ** 2).sum(1)
    ph = np.mod(np.arctan2(xy[:, 1], xy[:, 0]), 2 * math.pi)
    obs, _, _ = np.histogram2d(r2, ph, bins=[np.linspace(0, 1, 11), np.linspace(0, 2 * math.pi, 13)])
    inside = r2 < 1
    obs_in = obs.sum()
    e = np.full(obs.shape, n / obs.size)
    chi2 = ((obs - e) ** 2 / e).sum() + (n - obs_in) * 1e3
    p = float(stats.chi2.sf(chi2, obs.size - 1))
    return p, used / n, bool(inside.all())


def test_unit_disk_uniform_area():
    p, draws, inside = _disk_check()
    assert inside and p > 1e-4, p
    # two draws per attempt, acceptance pi/4
    assert draws == pytest.approx(8 / math.pi, rel=0.01)


def test_unit_disk_check_catches_planted_mistake():
    with fault(fault.DISK_NO_REJECT):
        p, draws, inside = _disk_check()
    assert not inside or p < 1e-8


# ----------------------------------------------------------------------------- one bounce
def _mat(kind, rough, eta=(0.2, 0.27, 1.1), k=(3.9, 2.78, 2.6), ior=1.5, aniso=0.0):
    m = ra.Material()
    m.type = kind
    m.albedo[:] = (1.0, 1.0, 1.0)
    m.roughness = rough
    m.eta[:] = eta
    m.k[:] = k
    m.refractive_index = ior
    m.anisotropy = aniso
    return m


def _alphas(m):
    r = float(np.float32(m.roughness))
    if m.type == ra.MAT.ANISOTROPIC_METAL:
        aspect = math.sqrt(max(1e-4, 1 - 0.9 * float(np.float32(m.anisotropy))))
        return max(1e-4, r * r / aspect), max(1e-4, r * r * aspect)
    a = max(1e-4, r * r)
    return a, a


def _mc(m, th_deg, n, front=True, seed=11):
    wo = sph(math.radians(th_deg), math.radians(25.0))
    rnd = np.random.default_rng(seed).random((n, 3))
    ok, od, at = scatter_batch(m, -wo, (0.0, 0.0, 1.0), rnd, front=front, tangent=(1.0, 0.0, 0.0))
    return wo, ok, od, at


def conductor_albedo_numeric(m, wo):
    ax, ay = _alphas(m)
    wi, dw = hemisphere_grid(500, 1000)
    h = wo + wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    g2 = 1.0 / (1.0 + smith_lambda(wo, ax, ay) + smith_lambda(wi, ax, ay))
    base = ggx_d(h, ax, ay) * g2 / (4.0 * wo[2]) * dw
    coh = np.clip((h * wo).sum(-1), 0, 1)
    return np.array([(base * fresnel_complex(coh, float(np.float32(e)), float(np.float32(kk)))).sum()
                     for e, kk in zip(m.eta, m.k)])


CONDUCTOR_CASES = [(ra.MAT.GGX_CONDUCTOR, 0.509, 0.0, 35.0), (ra.MAT.GGX_CONDUCTOR, 0.8, 0.0, 70.0),
                   (ra.MAT.ANISOTROPIC_METAL, 0.6, 0.7, 50.0)]


def _conductor_albedo_err(kind, rough, aniso, th, n=200_000):
    m = _mat(kind, rough, aniso=aniso)
    wo, ok, od, at = _mc(m, th, n)
    w = np.where(ok[:, None], at, 0.0)
    mc, se = w.mean(0), w.std(0) / math.sqrt(n)
    num = conductor_albedo_numeric(m, wo)
    return np.abs(mc - num) / (5 * se + 2e-3), mc, num


@pytest.mark.parametrize("case", CONDUCTOR_CASES)
def test_ggx_conductor_estimator_matches_integrated_brdf(case):
    err, mc, num = _conductor_albedo_err(*case)
    assert (err < 1).all(), (case, mc, num)


def dielectric_albedo_numeric(m, wo, eta):
    a = _alphas(m)[0]
    F = fresnel_textbook
    lo = smith_lambda(wo, a, a)
    # reflection: f cos = D G2 F / (4 wo.z) over the upper hemisphere, h = (wo + wi) / |.|
    wi, dw = hemisphere_grid(500, 1000)
    h = wo + wi
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    g2 = 1.0 / (1.0 + lo + smith_lambda(wi, a, a))
    coh = (h * wo).sum(-1)
    refl = (ggx_d(h, a, a) * g2 * F(coh, eta) / (4 * wo[2]) * dw).sum()
    # transmission (radiance not scaled by eta^2): pdf(wi) = (1 - F) D_wo(m) |wi.m| / |eta wo + wi|^2
    # with m = +-(eta wo + wi) / |.| facing wo, times the weight G2 / G1
    wi, dw = hemisphere_grid(500, 1000, lower=True)
    v = eta * wo + wi
    m_ = v / np.linalg.norm(v, axis=-1, keepdims=True)
    m_ = np.where(m_[..., 2:3] < 0, -m_, m_)
    com, cim = (m_ * wo).sum(-1), (m_ * wi).sum(-1)
    valid = (com > 0) & (cim < 0)
    g2 = 1.0 / (1.0 + lo + smith_lambda(wi, a, a))
    jac = np.abs(cim) / (v * v).sum(-1)
    tr = np.where(valid, g2 * (1 - F(np.clip(com, 0, 1), eta)) * ggx_d(m_, a, a) * com / wo[2] * jac * dw, 0.0)
    return refl, tr.sum()


DIELECTRIC_CASES = [(0.509, True, 35.0), (0.509, False, 30.0), (0.3, True, 70.0), (0.8, True, 20.0)]


def _dielectric_err(rough, front, th, n=200_000):
    m = _mat(ra.MAT.GGX_DIELECTRIC, rough, ior=1.5)
    eta = 1 / 1.5 if front else 1.5
    wo, ok, od, at = _mc(m, th, n, front=front)
    w = np.where(ok, at[:, 0], 0.0)
    errs, vals = [], []
    num = dielectric_albedo_numeric(m, wo, eta)
    for sel, ref in ((od[:, 2] > 0, num[0]), (od[:, 2] < 0, num[1])):
        x = np.where(sel, w, 0.0)
        mc, se = x.mean(), x.std() / math.sqrt(n)
        errs.append(abs(mc - ref) / (5 * se + 2e-3))
        vals.append((mc, ref))
    return np.array(errs), vals


@pytest.mark.parametrize("case", DIELECTRIC_CASES)
def test_ggx_dielectric_estimator_matches_integrated_bsdf(case):
    err, vals = _dielectric_err(*case)
    assert (err < 1).all(), (case, vals)


@pytest.mark.parametrize("which", [fault.VNDF_NO_WARP, fault.VNDF_SWAP_ALPHA, fault.CONDUCTOR_RP_SIGN,
                                   fault.LAMBDA_HALF])
def test_conductor_estimator_check_catches_planted_mistakes(which):
    with fault(which):
        worst = max(float(_conductor_albedo_err(*c, n=100_000)[0].max()) for c in CONDUCTOR_CASES)
    assert worst > 1, which


# (the hemisphere warp barely moves a dielectric's total albedo; the chi-square test above
# and the conductor check catch that one)
@pytest.mark.parametrize("which", [fault.DIELECTRIC_RP, fault.LAMBDA_HALF])
def test_dielectric_estimator_check_catches_planted_mistakes(which):
    with fault(which):
        worst = max(float(_dielectric_err(*c, n=100_000)[0].max()) for c in DIELECTRIC_CASES)
    assert worst > 1, which
