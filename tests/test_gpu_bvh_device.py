"""Device BVH build (RAYON_BVH_DEVICE) against the host build.

The device tree must satisfy the same invariants as the host tree (rayon_scene_check:
every slot in exactly one leaf of 1..4, every child box containing what lies below it,
depth within the traversal stack, and the reported depth equal to the walked one), and,
because closest-hit ties resolve by original primitive index whatever the traversal order,
render images identical bit for bit to the host tree's on the same job.
"""
import ctypes as C

import numpy as np
import pytest

import rayon_amd as ra
from test_oracle import make_desc, mat, sphere, tri

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = ra.Context(0)
    yield c
    c.close()


def render(ctx, sc, cam, w, h, spp, batched=False, flags=0):
    j = ctx.job(sc, ra.job_params(w, h, cam, spp=spp, max_depth=8, seed=5, flags=flags))
    if batched:
        ctx.render_batched([j], spp)
    else:
        j.render_plain(spp)
    img = j.resolve_linear()
    j.close()
    return img


GENS = {"book": (ra.GEN.BOOK_SPHERES, 3, 0, 0), "cornell": (ra.GEN.CORNELL_ICOSPHERE, 7, 5, ra.FLAG.NEE),
        "sdf": (ra.GEN.SDF_TERRAIN, 11, 0, 0)}


@pytest.mark.parametrize("name", sorted(GENS))
def test_device_tree_valid_and_images_identical(ctx, name):
    kind, seed, detail, flags = GENS[name]
    print(f"[bvh_device] {name}", flush=True)
    g = ra.Generated.generate(kind, seed, detail=detail)
    host = ctx.scene(g.desc_ptr())
    dev = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE)
    hi, di = host.info(), dev.info()
    assert hi.builder == ra.BVH.BUILDER_HOST and di.builder == ra.BVH.BUILDER_DEVICE
    assert di.num_prims == hi.num_prims
    assert host.check(g.desc_ptr()) == hi.max_depth
    assert dev.check(g.desc_ptr()) == di.max_depth
    assert 0 < di.max_depth <= 32
    lf, la, up, fov = g.camera()
    w, h = 96, 64
    cam = ra.camera_frame(lf, la, up, fov, w, h)
    for batched in (False, True):
        a = render(ctx, host, cam, w, h, 4, batched, flags)
        b = render(ctx, dev, cam, w, h, 4, batched, flags)
        assert np.isfinite(a).all() and a.mean() > 0
        np.testing.assert_array_equal(a, b)
    host.close()
    dev.close()


def test_device_tree_closest_hits_identical(ctx):
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=6)
    host = ctx.scene(g.desc_ptr())
    dev = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE)
    rng = np.random.default_rng(4)
    n = 20000
    o = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32) + np.float32([278, 278, 278])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    ph, th = host.trace_closest(o, d)
    pd, td = dev.trace_closest(o, d)
    assert (ph >= 0).mean() > 0.9
    np.testing.assert_array_equal(ph, pd)
    np.testing.assert_array_equal(th, td)
    host.close()
    dev.close()


def small_scenes():
    m = [mat(ra.MAT.LAMBERTIAN, (0.7, 0.6, 0.5)), mat(ra.MAT.LIGHT, emission=(4, 4, 4))]
    rng = np.random.default_rng(9)
    out = {
        "one": [sphere((0, 0, -1), 0.5, 0)],
        "three": [sphere((0, 0, -1), 0.5, 0), sphere((0, -100.5, -1), 100, 0), sphere((1, 1, -1), 0.3, 1)],
        "five": [sphere((x, 0, -2), 0.3, x % 2) for x in range(-2, 3)],
        # equal centroids: every Morton code collides, the tree splits by position
        "coincident": [sphere((0, 0, -2), 0.2 + 0.001 * i, i % 2) for i in range(600)],
        "soup": [tri(*(rng.uniform(-1, 1, 3) + [0, 0, -3] for _ in range(3)), int(i % 2)) for i in range(3000)],
    }
    return m, out


@pytest.mark.parametrize("name", ["one", "three", "five", "coincident", "soup"])
def test_device_tree_small_and_degenerate_scenes(ctx, name):
    m, scenes = small_scenes()
    desc, keep = make_desc(m, scenes[name])
    dp = C.pointer(desc)
    host = ctx.scene(dp)
    dev = ctx.scene(dp, ra.BVH.DEVICE)
    assert dev.info().builder == ra.BVH.BUILDER_DEVICE
    assert dev.check(dp) == dev.info().max_depth
    assert host.check(dp) == host.info().max_depth
    w, h = 48, 32
    cam = ra.camera_frame((0, 0.5, 2), (0, 0, -1), (0, 1, 0), 50.0, w, h)
    np.testing.assert_array_equal(render(ctx, host, cam, w, h, 4), render(ctx, dev, cam, w, h, 4))
    host.close()
    dev.close()
    del keep


@pytest.mark.parametrize("name", sorted(GENS))
def test_device_built_quantized_tree(ctx, name):
    """RAYON_BVH_DEVICE | RAYON_BVH_QUANTIZED: the device-built tree collapsed into 8-wide
    quantized nodes passes the wide and binary invariants (rayon_scene_check) and renders the
    host binary tree's images bit for bit, on the plain kernel and the batched path."""
    kind, seed, detail, flags = GENS[name]
    print(f"[bvh_device quantized] {name}", flush=True)
    g = ra.Generated.generate(kind, seed, detail=detail)
    host = ctx.scene(g.desc_ptr())
    dq = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE | ra.BVH.QUANTIZED)
    qi = dq.info()
    assert qi.builder == ra.BVH.BUILDER_DEVICE
    assert qi.num_qnodes > 0 and 0 < qi.wide_depth < 16
    assert dq.check(g.desc_ptr()) == qi.max_depth
    lf, la, up, fov = g.camera()
    w, h = 96, 64
    cam = ra.camera_frame(lf, la, up, fov, w, h)
    for batched in (False, True):
        a = render(ctx, host, cam, w, h, 4, batched, flags)
        b = render(ctx, dq, cam, w, h, 4, batched, flags)
        np.testing.assert_array_equal(a, b)
    host.close()
    dq.close()


def test_device_built_quantized_closest_hits_identical(ctx):
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=6)
    host = ctx.scene(g.desc_ptr())
    dq = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE | ra.BVH.QUANTIZED)
    assert dq.info().num_qnodes > 0
    rng = np.random.default_rng(5)
    n = 20000
    o = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32) + np.float32([278, 278, 278])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    ph, th = host.trace_closest(o, d)
    pq, tq = dq.trace_closest(o, d)
    assert (ph >= 0).mean() > 0.9
    np.testing.assert_array_equal(ph, pq)
    np.testing.assert_array_equal(th, tq)
    host.close()
    dq.close()


def test_scene_check_covers_host_quantized_tree(ctx):
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=5)
    q = ctx.scene(g.desc_ptr(), ra.BVH.QUANTIZED)
    assert q.info().num_qnodes > 0
    assert q.check(g.desc_ptr()) == q.info().max_depth
    q.close()
