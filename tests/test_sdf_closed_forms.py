"""SDF fields against closed forms: the oracle's sphere tracing, primitive formulas and
smooth union (oracle/oracle.cpp:isect_sdf, sdf_prim, smooth_union, after the reference's
sdf.hpp:105-143, :410-414, :518-563), and the GPU's marching on the same fields.

Sphere tracing stops at the first point whose field value is below eps, approaching the
surface from outside, so the hit distance lies in [t* - 2 eps, t*] for the exact surface
distance t*. Each case is chosen so that a plausible mistake moves t* by much more:
* sphere: |o - c| - r, head-on and off-centre, and a miss just beside it;
* rounded box: a face (half extent, not half extent minus rounding) and a corner, which is
  the rounding sphere around (h - r) along the diagonal;
* torus in the XZ plane: the tube's top at x = R, the outer equator R + r, a miss through
  the hole;
* smooth union: two coincident spheres give the surface at r + k/4 (the k h (1 - h) term),
  three give r + 25k/64 (the left fold), and two far-apart spheres keep their own surfaces.
"""
import ctypes as C
import math

import numpy as np
import pytest

import rayon_amd as ra
from oracle import OracleScene
from test_oracle import make_desc, mat

EPS = 1e-4
S, B, T = ra.SDF.SPHERE, ra.SDF.ROUNDED_BOX, ra.SDF.TORUS


def sdf_desc(prims, k=0.1):
    m = [mat(ra.MAT.LAMBERTIAN, (0.6, 0.6, 0.6))]
    desc, keep = make_desc(m, [])
    P = (ra.SdfPrim * len(prims))()
    for i, (t, c, p) in enumerate(prims):
        P[i].type = t
        P[i].center[:] = c
        P[i].params[:] = list(p) + [0.0] * (4 - len(p))
    desc.sdf_prims, desc.num_sdf_prims = P, len(prims)
    desc.sdf_smooth_k, desc.sdf_material_id, desc.sdf_max_steps, desc.sdf_eps = k, 0, 256, EPS
    keep.append(P)
    return desc, keep


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


_d3 = unit([1, 1, 1])
_q = np.array([0.4, 0.2, 0.3])  # centre of the rounding sphere of the box corner (h - r)
_r2 = math.sqrt(0.7 ** 2 - 0.3 ** 2)

# (name, prims, k, origin, unit direction, expected t or None for a miss)
CASES = [
    ("sphere_head_on", [(S, (1, 2, 3), (0.7,))], 0.1, (1, 2, -2), (0, 0, 1), 5 - 0.7),
    ("sphere_off_centre", [(S, (1, 2, 3), (0.7,))], 0.1, (1.3, 2, -2), (0, 0, 1), 5 - _r2),
    ("sphere_miss", [(S, (1, 2, 3), (0.7,))], 0.1, (1.8, 2, -2), (0, 0, 1), None),
    ("box_face", [(B, (0, 0, 0), (0.5, 0.3, 0.4, 0.1))], 0.1, (-3, 0, 0), (1, 0, 0), 3 - 0.5),
    ("box_corner", [(B, (0, 0, 0), (0.5, 0.3, 0.4, 0.1))], 0.1, tuple(_q + 5 * _d3), tuple(-_d3), 5 - 0.1),
    ("torus_top", [(T, (0, 0, 0), (0.6, 0.2))], 0.1, (0.6, 3, 0), (0, -1, 0), 3 - 0.2),
    ("torus_equator", [(T, (0, 0, 0), (0.6, 0.2))], 0.1, (3, 0, 0), (-1, 0, 0), 3 - 0.8),
    ("torus_hole", [(T, (0, 0, 0), (0.6, 0.2))], 0.1, (0, 3, 0), (0, -1, 0), None),
    ("union_two_coincident", [(S, (0, 0, 0), (0.5,))] * 2, 0.1, (0, 0, -5), (0, 0, 1), 5 - (0.5 + 0.1 / 4)),
    ("union_three_coincident", [(S, (0, 0, 0), (0.5,))] * 3, 0.1, (0, 0, -5), (0, 0, 1),
     5 - (0.5 + 25 * 0.1 / 64)),
    ("union_far_apart", [(S, (-2, 0, 0), (0.5,)), (S, (2, 0, 0), (0.5,))], 0.1, (-2, 0, -5), (0, 0, 1), 4.5),
]


def check(t_star, prim, t, tol):
    if t_star is None:
        assert prim < 0, f"expected a miss, got a hit at t={t}"
    else:
        assert prim >= 0, f"expected a hit at {t_star}"
        assert t_star - 2 * EPS - tol <= t <= t_star + tol, (t, t_star)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_sdf_closed_forms(case):
    name, prims, k, o, d, t_star = case
    desc, keep = sdf_desc(prims, k)
    s = OracleScene(C.pointer(desc))
    p, t = s.trace_closest(np.array([o], np.float32), np.array([d], np.float32), t_min=0.001)
    # FP64 field; the only deviation from t* is the stopping rule (and the float32 ray input)
    check(t_star, int(p[0]), float(t[0]), 1e-6)
    s.close()
    del keep


@pytest.mark.gpu
def test_gpu_sdf_closed_forms():
    ctx = ra.Context(0)
    try:
        for name, prims, k, o, d, t_star in CASES:
            print(f"[sdf closed form] {name}", flush=True)
            desc, keep = sdf_desc(prims, k)
            sc = ctx.scene(C.pointer(desc))
            p, t = sc.trace_closest(np.array([o], np.float32), np.array([d], np.float32), t_min=0.001)
            # FP32 marching: a few ulps of the ~5-unit distances on top of the stopping rule
            check(t_star, int(p[0]), float(t[0]), 2e-5)
            sc.close()
            del keep
    finally:
        ctx.close()
