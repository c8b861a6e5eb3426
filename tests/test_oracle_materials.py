"""Pins of the oracle's remaining surface functions against closed forms and definitions:
thin-film interference (cuda_raytracer.cuh:753-809), the clear-coat lobe choice
(:810-846), the Fibonacci-dots pattern (:377-416) and the golf-ball dimples
(shader_golf.cu:4-43, :53-140).

* Thin film at normal incidence with the same medium on both sides of the film: both
  interfaces reflect R = ((n0 - n1)/(n0 + n1))^2 and the Airy sum becomes
  2R(1 + cos d)/(1 + R^2 + 2R cos d), d = 4 pi n1 t / lambda. A quarter-wave film at 550 nm
  (cos d = -1) reflects nothing at 550 nm; a half-wave film (cos d = 1) reflects 4R/(1+R)^2.
  A sign slip on the cosine term swaps the two.
* Clear coat: the coat is chosen with the Schlick probability, at normal incidence
  ((1 - n)/(1 + n))^2 = 0.04 for n = 1.5: a draw just below takes the mirror direction with
  unit weight, just above the Lambert lobe with the albedo; an index <= 1 means 1.5.
* Fibonacci lattice: point i of N is z = 1 - (2i + 1)/N at azimuth (i + 1/2) times the golden
  angle pi (3 - sqrt 5), built here from that definition. A point on a dot takes the dot
  colour, a point farther than the dot radius from every lattice point the albedo.
* Golf ball (150 dimples of angular radius 0.24, depth 0.35 of the radius): at a dimple
  centre the surface is the base sphere; at angle a < 0.24 from the centre it is pushed out
  to r (1 + 0.35 (1 - cos(a / 0.24 * pi / 2))).
"""
import ctypes as C
import math

import numpy as np
import pytest

import rayon_amd as ra
from oracle import lib as olib
from test_oracle import mat, probe

D3 = C.c_double * 3


def fib(i, n):
    k = i + 0.5
    z = 1.0 - 2.0 * k / n
    th = k * math.pi * (3.0 - math.sqrt(5.0))
    r = math.sqrt(max(0.0, 1.0 - z * z))
    return np.array([math.cos(th) * r, math.sin(th) * r, z])


def rotate_towards(u, axis_hint, ang):
    """Unit vector at angle `ang` from unit vector u, in the plane of u and axis_hint."""
    t = axis_hint - np.dot(axis_hint, u) * u
    t /= np.linalg.norm(t)
    return math.cos(ang) * u + math.sin(ang) * t


# ---------------------------------------------------------------------------- thin film
def thin_film(n0, n1, thickness):
    return mat(ra.MAT.THIN_FILM, (1, 1, 1), refractive_index=n0, film_ior=n1, film_thickness=thickness)


@pytest.mark.parametrize("kind", ["quarter", "half"])
def test_thin_film_interference_extremes(kind):
    n0, n1, lam = 1.0, 1.5, 550.0
    R = ((n0 - n1) / (n0 + n1)) ** 2
    t = lam / (4 * n1) if kind == "quarter" else lam / (2 * n1)
    m = thin_film(n0, n1, t)
    d, n = (0.0, 0.0, -1.0), (0.0, 0.0, 1.0)  # normal incidence
    # transmission branch (draw 0.9999): weight (1 - R_c) / (1 - avg) per channel
    ok_t, od_t, at_t, _, _ = probe(m, d, n, rnd=(0.9999,))
    # reflection branch (draw 0): weight R_c / avg
    ok_r, od_r, at_r, _, _ = probe(m, d, n, rnd=(0.0,))
    assert ok_t and ok_r
    np.testing.assert_allclose(od_t, d, atol=1e-12)  # straight through
    np.testing.assert_allclose(od_r, (0, 0, 1), atol=1e-12)  # mirror
    # the channel weights of the two branches determine R_c and avg: R_c = avg * at_r[c] and
    # 1 - R_c = (1 - avg) * at_t[c], so avg = (1 - at_t[c]) / (at_r[c] - at_t[c]) for every
    # channel, and avg is the mean of the R_c
    avg = (1.0 - at_t) / (at_r - at_t)
    np.testing.assert_allclose(avg, avg[0], rtol=1e-9)
    Rc = avg * at_r
    assert Rc.mean() == pytest.approx(avg[0], rel=1e-9)
    expected_g = 0.0 if kind == "quarter" else 4 * R / (1 + R) ** 2
    assert Rc[1] == pytest.approx(expected_g, abs=1e-12)
    # the other channels follow the same Airy sum at their wavelengths (650, 450 nm), with the
    # thickness as the material stores it (float32)
    t32 = float(np.float32(t))
    for c, lam_c in ((0, 650.0), (2, 450.0)):
        cd = math.cos(4 * math.pi * n1 * t32 / lam_c)
        assert Rc[c] == pytest.approx(2 * R * (1 + cd) / (1 + R * R + 2 * R * cd), rel=1e-9)


# ---------------------------------------------------------------------------- clear coat
@pytest.mark.parametrize("ior", [1.5, 1.0, 0.0])  # <= 1 stands for 1.5
def test_clear_coat_lobe_choice_at_normal_incidence(ior):
    albedo = (0.3, 0.5, 0.7)
    m = mat(ra.MAT.CLEAR_COAT, albedo, refractive_index=ior, roughness=0.0)
    F = ((1 - 1.5) / (1 + 1.5)) ** 2  # 0.04
    d, n = (0.0, 0.0, -1.0), (0.0, 0.0, 1.0)
    ok, od, at, _, _ = probe(m, d, n, rnd=(F - 1e-4, 0.3, 0.6))
    assert ok
    np.testing.assert_allclose(od, (0, 0, 1), atol=1e-12)
    np.testing.assert_allclose(at, (1, 1, 1), atol=1e-12)
    ok, od, at, _, _ = probe(m, d, n, rnd=(F + 1e-4, 0.3, 0.6))
    assert ok
    np.testing.assert_allclose(at, albedo, rtol=1e-6)
    assert od[2] > 0 and abs(np.linalg.norm(od) - 1) < 1e-9  # a cosine-lobe direction


# ---------------------------------------------------------------------------- Fibonacci dots
def surface_color(m, p, c=(0.0, 0.0, 0.0)):
    out = D3()
    olib().oracle_surface_color(C.byref(m), D3(*p), D3(*c), out)
    return np.array(out)


def test_fibonacci_dots_follow_the_golden_angle_lattice():
    N, rad = 30, 0.2
    base, dot = (0.9, 0.1, 0.1), (0.02, 0.03, 0.04)
    m = mat(ra.MAT.LAMBERTIAN, base, pattern=ra.PATTERN.FIBONACCI_DOTS, pattern_color=dot,
            pattern_param1=float(N), pattern_param2=rad)
    pts = np.array([fib(i, N) for i in range(N)])
    c = np.array([0.5, -1.0, 2.0])
    for j in (0, 7, N - 1):
        np.testing.assert_allclose(surface_color(m, c + 3.0 * pts[j], c), dot, atol=1e-7)
        inside = rotate_towards(pts[j], np.array([0.3, 0.9, 0.1]), 0.9 * rad)
        np.testing.assert_allclose(surface_color(m, c + 2.0 * inside, c), dot, atol=1e-7)
    # a direction farther than the dot radius from every lattice point
    rng = np.random.default_rng(3)
    found = 0
    for _ in range(2000):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        ang = np.arccos(np.clip(pts @ v, -1, 1)).min()
        if ang > 1.05 * rad:
            np.testing.assert_allclose(surface_color(m, c + v, c), base, atol=1e-7)
            found += 1
        elif ang < 0.95 * rad:
            np.testing.assert_allclose(surface_color(m, c + v, c), dot, atol=1e-7)
    assert found > 100


# ---------------------------------------------------------------------------- golf ball
def golf_hit(c, rad, o, d):
    t, p, n = C.c_double(), D3(), D3()
    ok = olib().oracle_golf_hit(D3(*c), C.c_double(rad), D3(*o), D3(*d), C.byref(t), p, n)
    return bool(ok), t.value, np.array(p), np.array(n)


@pytest.mark.parametrize("frac", [0.0, 0.5, 0.75])
def test_golf_ball_dimple_profile(frac):
    c, rad = np.array([1.0, 2.0, -3.0]), 0.8
    pts = np.array([fib(i, 150) for i in range(150)])
    j = 41
    w = rotate_towards(pts[j], np.array([0.2, -0.5, 0.8]), frac * 0.24)
    assert np.arccos(np.clip(pts @ w, -1, 1)).argmin() == j  # still nearest to dimple j
    ok, t, p, n = golf_hit(c, rad, c + 5.0 * w, -w)  # radial ray
    assert ok
    assert t == pytest.approx(5.0 - rad, abs=1e-9)  # the base sphere's distance
    p = p - 1e-3 * n  # the hit point is offset along the shading normal
    expected = rad * (1.0 + 0.35 * (1.0 - math.cos(frac * math.pi / 2)))
    # the shader's golden angle is 2.39996323 (pi (3 - sqrt 5) to 8 decimals): the dimple
    # centres differ from the exact lattice by ~1e-8 rad, a ~1e-9 change of the radius
    assert np.linalg.norm(p - c) == pytest.approx(expected, abs=1e-7)
    np.testing.assert_allclose((p - c) / np.linalg.norm(p - c), w, atol=1e-9)
