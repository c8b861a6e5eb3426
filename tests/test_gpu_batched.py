"""GPU tests of the batched wavefront path (rayon_render_batched).

Tolerances: a job rendered batched must equal the same job rendered alone in a batch
bit for bit (same random streams, fixed per-pixel accumulation order), whatever else is
in the batch and however the work is cut into rounds; it also equals the plain
megakernel's render bit for bit (the same device functions, contracted only within a
source expression, and closest hits that do not depend on traversal order); against the
FP64 oracle the relative RMSE is at most 1%.
"""
import sys

import numpy as np
import pytest

import rayon_amd as ra
from oracle import OracleScene

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Scenes:
    def __init__(self, ctx):
        self.ctx = ctx
        self.gens, self.scenes = {}, {}

    def get(self, kind, scene_seed, detail):
        key = (kind, scene_seed, detail)
        if key not in self.scenes:
            g = ra.Generated.generate(kind, scene_seed, detail=detail)
            self.gens[key] = g
            self.scenes[key] = self.ctx.scene(g.desc_ptr())
        return self.gens[key], self.scenes[key]

    def job(self, kind, scene_seed, detail, W, H, **kw):
        g, sc = self.get(kind, scene_seed, detail)
        lf, la, up, fov = g.camera()
        cam = ra.camera_frame(lf, la, up, fov, W, H)
        P = ra.job_params(W, H, cam, **kw)
        return self.ctx.job(sc, P), P, cam, g


@pytest.fixture(scope="module")
def scenes(ctx):
    s = Scenes(ctx)
    yield s
    for sc in s.scenes.values():
        sc.close()


BOOK = (ra.GEN.BOOK_SPHERES, 1984, 0)
CORN = (ra.GEN.CORNELL_ICOSPHERE, 7, 4)
SDF = (ra.GEN.SDF_TERRAIN, 11, 65)


def test_batched_matches_plain(ctx, scenes):
    for scene, kw in ((BOOK, dict(max_depth=50)), (CORN, dict(max_depth=8, flags=ra.FLAG.NEE))):
        log(f"batched vs plain {scene}")
        a, *_ = scenes.job(*scene, 96, 64, spp=32, seed=5, **kw)
        b, *_ = scenes.job(*scene, 96, 64, spp=32, seed=5, **kw)
        a.render_plain(32)
        ctx.render_batched([b], 32)
        ia, ib = a.resolve_linear(), b.resolve_linear()
        log(f"  identical {np.mean(ia == ib):.4f}")
        np.testing.assert_array_equal(ia, ib)
        sa, sb = a.stats(), b.stats()
        assert sa.primary_samples == sb.primary_samples == 96 * 64 * 32
        assert sa.segments == sb.segments and sa.shadow_segments == sb.shadow_segments
        np.testing.assert_array_equal(a.sample_counts(), b.sample_counts())
        a.close()
        b.close()


def test_alone_equals_in_mix_bit_exact(ctx, scenes):
    log("alone vs in the mix")
    alone, *_ = scenes.job(*CORN, 80, 60, spp=24, seed=9, max_depth=8, flags=ra.FLAG.NEE)
    ctx.render_batched([alone], 24)
    ref = alone.resolve_linear()
    mixed, *_ = scenes.job(*CORN, 80, 60, spp=24, seed=9, max_depth=8, flags=ra.FLAG.NEE)
    others = [scenes.job(*BOOK, 128, 72, spp=16, seed=1, max_depth=50)[0],
              scenes.job(*SDF, 64, 48, spp=40, seed=2, max_depth=8)[0],
              scenes.job(*CORN, 33, 17, spp=24, seed=3, max_depth=8, flags=ra.FLAG.NEE)[0]]
    batch = [others[0], mixed, others[1], others[2]]
    ctx.render_batched(batch, 8)   # three passes of 8 samples, mixed with jobs of other sizes
    ctx.render_batched(batch, 8)
    ctx.render_batched(batch, 8)
    got = mixed.resolve_linear()
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(mixed.sample_counts(), alone.sample_counts())
    for j in others + [alone, mixed]:
        j.close()


def test_rounds_and_graphs_do_not_change_results(ctx, scenes):
    log("capacity / graph invariance")
    scenes.get(*BOOK)[1].set_batch_path(ra.PATH.WAVEFRONT)
    base, *_ = scenes.job(*BOOK, 120, 80, spp=12, seed=4, max_depth=16)
    ctx.set_option(ra.OPT.GRAPH, 0)
    ctx.render_batched([base], 12)
    ref = base.resolve_linear()
    old = ctx.get_option(ra.OPT.WAVE_CAPACITY)
    try:
        ctx.set_option(ra.OPT.WAVE_CAPACITY, 65536)  # 120*80*12 = 115200 work items -> 2 rounds
        for graph in (0, 1):
            ctx.set_option(ra.OPT.GRAPH, graph)
            j, *_ = scenes.job(*BOOK, 120, 80, spp=12, seed=4, max_depth=16)
            ctx.render_batched([j], 6)
            ctx.render_batched([j], 6)
            np.testing.assert_array_equal(j.resolve_linear(), ref)
            j.close()
    finally:
        ctx.set_option(ra.OPT.WAVE_CAPACITY, old)
        ctx.set_option(ra.OPT.GRAPH, 1)
        scenes.get(*BOOK)[1].set_batch_path(ra.PATH.AUTO)
    base.close()


def test_graph_of_same_shape_other_scene(ctx, scenes):
    """A pass whose work sizes match an earlier pass on another kind of scene (here an SDF
    field after a sphere field) must not reuse that pass's graph: results with graphs
    equal the eager ones."""
    log("graph key covers the scene's stages")
    try:
        for sc in (BOOK, SDF):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.WAVEFRONT)
        ctx.set_option(ra.OPT.GRAPH, 1)
        b, *_ = scenes.job(*BOOK, 64, 48, spp=4, seed=31, max_depth=8)
        ctx.render_batched([b], 4)
        g, *_ = scenes.job(*SDF, 64, 48, spp=4, seed=32, max_depth=8)
        ctx.render_batched([g], 4)
        ctx.set_option(ra.OPT.GRAPH, 0)
        e, *_ = scenes.job(*SDF, 64, 48, spp=4, seed=32, max_depth=8)
        ctx.render_batched([e], 4)
        np.testing.assert_array_equal(g.resolve_linear(), e.resolve_linear())
        assert g.stats().segments == e.stats().segments
        for j in (b, g, e):
            j.close()
    finally:
        ctx.set_option(ra.OPT.GRAPH, 1)
        for sc in (BOOK, SDF):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.AUTO)


def test_megakernel_path_equals_plain_bit_exact(ctx, scenes):
    """Jobs on the megakernel path share one launch per pass and get the plain kernel's
    bits; the same jobs on the wavefront path get them too."""
    log("megakernel path")
    specs = [(BOOK, 96, 54, dict(spp=12, seed=21, max_depth=50)),
             (SDF, 64, 48, dict(spp=7, seed=22, max_depth=8)),
             (CORN, 40, 30, dict(spp=9, seed=23, max_depth=8, flags=ra.FLAG.NEE)),
             (BOOK, 17, 33, dict(spp=5, seed=24, max_depth=16))]
    try:
        for sc in (BOOK, SDF, CORN):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.MEGAKERNEL)
        plain = [scenes.job(*sc, W, H, **kw)[0] for sc, W, H, kw in specs]
        for j, (_, _, _, kw) in zip(plain, specs):
            j.render_plain(kw["spp"])
        mega = [scenes.job(*sc, W, H, **kw)[0] for sc, W, H, kw in specs]
        ctx.render_batched(mega, 4)   # two passes: 4 samples, then the rest
        ctx.render_batched(mega, 64)
        for a, b in zip(plain, mega):
            np.testing.assert_array_equal(b.resolve_linear(), a.resolve_linear())
            assert a.stats().segments == b.stats().segments
        for sc in (BOOK, SDF, CORN):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.WAVEFRONT)
        wave = [scenes.job(*sc, W, H, **kw)[0] for sc, W, H, kw in specs]
        ctx.render_batched(wave, 64)
        for a, b in zip(plain, wave):
            np.testing.assert_array_equal(b.resolve_linear(), a.resolve_linear())
            assert a.stats().segments == b.stats().segments
        for j in plain + mega + wave:
            j.close()
    finally:
        for sc in (BOOK, SDF, CORN):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.AUTO)


def test_graph_replay_repeats_eager_result(ctx, scenes):
    """The captured pass graph replayed on fresh jobs reproduces the eager pass each time,
    with eager work and synchronisations between the replays."""
    log("graph replays")
    ctx.set_option(ra.OPT.GRAPH, 0)
    e, *_ = scenes.job(*CORN, 64, 64, spp=4, seed=8, max_depth=8, flags=ra.FLAG.NEE)
    ctx.render_batched([e], 4)
    ref = e.resolve_linear()
    ctx.set_option(ra.OPT.GRAPH, 1)
    j, *_ = scenes.job(*CORN, 64, 64, spp=16, seed=8, max_depth=8, flags=ra.FLAG.NEE)
    for rep in range(4):
        j.reset()
        ctx.render_batched([j], 4)
        img = j.resolve_linear()
        e.resolve_linear()  # eager kernels and a synchronisation between replays
        ctx.sync()
        np.testing.assert_array_equal(img, ref)
    e.close()
    j.close()


def test_batched_against_oracle_sdf_terrain(ctx, scenes):
    log("sdf + terrain batched vs oracle")
    scenes.get(*SDF)[1].set_batch_path(ra.PATH.WAVEFRONT)
    j, P, cam, g = scenes.job(*SDF, 64, 40, spp=32, seed=6, max_depth=8)
    ctx.render_batched([j], 32)
    scenes.get(*SDF)[1].set_batch_path(ra.PATH.AUTO)
    img = j.resolve_linear().reshape(-1, 3).astype(np.float64)
    sums, segs = OracleScene(g.desc_ptr()).render(P, nsamples=32, camera=cam)
    ref = sums / 32
    err = float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum()))
    log(f"  rel_rmse={err:.5f}")
    assert err <= 0.01
    j.close()


def test_adaptive_stderr_batched_matches_plain(ctx, scenes):
    log("adaptive stderr")
    kw = dict(spp=64, seed=3, max_depth=8, flags=ra.FLAG.ADAPTIVE_STDERR, min_adaptive_samples=16,
              adaptive_threshold=0.05)
    a, *_ = scenes.job(*BOOK, 64, 48, **kw)
    b, *_ = scenes.job(*BOOK, 64, 48, **kw)
    for _ in range(4):
        a.render_plain(16)
        ctx.render_batched([b], 16)
    ca, cb = a.sample_counts(), b.sample_counts()
    agree = np.mean(ca == cb)
    log(f"  converged plain {(ca < 0).mean():.3f} batched {(cb < 0).mean():.3f} agree {agree:.4f}")
    assert (cb < 0).mean() > 0.05
    assert agree > 0.999
    assert b.stats().converged_pixels == int((cb < 0).sum())
    a.close()
    b.close()


def test_adaptive_on_the_wavefront_compacts_live_pixels(ctx, scenes):
    """Adaptive jobs on the wavefront sample only the pixels that have not converged (a
    pixel list per pass); their per-pixel counts and sums agree with the plain kernel's
    (at most 0.1% of pixels differ in count: a convergence decision can flip on the last
    bits of the two paths' rounding)."""
    log("adaptive on the wavefront")
    kw = dict(spp=96, seed=5, max_depth=8, flags=ra.FLAG.ADAPTIVE_STDERR, min_adaptive_samples=16,
              adaptive_threshold=0.05)
    try:
        scenes.get(*BOOK)[1].set_batch_path(ra.PATH.WAVEFRONT)
        a, *_ = scenes.job(*BOOK, 96, 64, **kw)
        b, *_ = scenes.job(*BOOK, 96, 64, **kw)
        for _ in range(6):
            a.render_plain(16)
            ctx.render_batched([b], 16)
        ca, cb = a.sample_counts(), b.sample_counts()
        log(f"  converged plain {(ca < 0).mean():.3f} batched {(cb < 0).mean():.3f} agree {np.mean(ca == cb):.4f}")
        assert (cb < 0).mean() > 0.2 and np.mean(ca == cb) > 0.999
        assert b.stats().converged_pixels == int((cb < 0).sum())
        same = ca == cb
        ia, ib = a.resolve_linear()[same], b.resolve_linear()[same]
        assert (np.abs(ia - ib) / np.maximum(np.abs(ia), 1e-3)).max() <= 1e-4
        assert b.stats().primary_samples == int(np.abs(cb).sum())
        a.close()
        b.close()
    finally:
        scenes.get(*BOOK)[1].set_batch_path(ra.PATH.AUTO)


def test_reference_stream_jobs_are_rejected(ctx, scenes):
    j, *_ = scenes.job(*BOOK, 16, 16, spp=4, flags=ra.FLAG.REF_RNG)
    with pytest.raises(ra.RayonError, match="REF_RNG"):
        ctx.render_batched([j], 4)
    j.close()


def test_autotune_keeps_results_and_jobs(ctx, scenes):
    """The autotuner times paths and shapes on clones (the jobs stay untouched), records a
    path per scene, caches its shape choice per workload, and the tuned shape renders the
    same bits as the default one."""
    log("autotune")
    specs = [(BOOK, 128, 72, dict(spp=8, seed=1, max_depth=8)),
             (CORN, 96, 64, dict(spp=8, seed=2, max_depth=8, flags=ra.FLAG.NEE)),
             (SDF, 64, 48, dict(spp=8, seed=3, max_depth=8))]
    opt_keys = (ra.OPT.WAVE_CAPACITY, ra.OPT.WAVE_CHUNK, ra.OPT.WAVE_WG_PER_CU, ra.OPT.SHADE_CHUNK)
    default = {o: ctx.get_option(o) for o in opt_keys}
    try:
        jobs = [scenes.job(*sc, W, H, **kw)[0] for sc, W, H, kw in specs]
        r = ctx.autotune(jobs, 2)
        log(f"  chose capacity {r.capacity} chunk {r.chunk} wg {r.wg_per_cu}: {r.best_ms:.2f} ms "
            f"(start {r.start_ms:.2f} ms, {r.candidates} passes timed; {r.mega_scenes} megakernel / "
            f"{r.wave_scenes} wavefront scenes)")
        assert r.candidates >= 5 and r.cached == 0 and r.best_ms <= r.start_ms
        assert r.mega_scenes + r.wave_scenes == 3
        assert all(j.stats().samples_done == 0 for j in jobs)
        assert ctx.get_option(ra.OPT.WAVE_CHUNK) == r.chunk
        again = ctx.autotune(jobs, 2)
        assert again.cached == 1 and again.chunk == r.chunk and again.capacity == r.capacity
        tuned = {o: ctx.get_option(o) for o in opt_keys}
        for o, v in default.items():
            ctx.set_option(o, v)
        refs = [scenes.job(*sc, W, H, **kw)[0] for sc, W, H, kw in specs]
        ctx.render_batched(refs, 8)
        for o, v in tuned.items():
            ctx.set_option(o, v)
        ctx.render_batched(jobs, 8)
        for j, ref in zip(jobs, refs):
            np.testing.assert_array_equal(j.resolve_linear(), ref.resolve_linear())
            j.close()
            ref.close()
    finally:
        for o, v in default.items():
            ctx.set_option(o, v)
        for sc in (BOOK, SDF, CORN):
            scenes.get(*sc)[1].set_batch_path(ra.PATH.AUTO)


@pytest.mark.parametrize("scene,kw", [(CORN, dict(max_depth=8, flags=ra.FLAG.NEE)), (SDF, dict(max_depth=8)),
                                      (BOOK, dict(max_depth=50))], ids=["cornell_nee", "sdf", "book"])
def test_quantized_wavefront_refill_equals_per_ray_traversal(ctx, scene, kw):
    """Quantized trees on the wavefront: a pass over quantized-tree jobs alone runs the
    lane-refill wide traversal; batched with a binary-tree job it runs the per-ray wide
    traversal (closest_hit_wide / any_hit_wide). Both find the same hits, so the images are
    identical bit for bit, and so are other refill thresholds."""
    g = ra.Generated.generate(scene[0], scene[1], detail=scene[2])
    q = ctx.scene(g.desc_ptr(), ra.BVH.QUANTIZED)
    gb = ra.Generated.generate(*BOOK[:2], detail=BOOK[2])
    b = ctx.scene(gb.desc_ptr())
    try:
        q.set_batch_path(ra.PATH.WAVEFRONT)
        b.set_batch_path(ra.PATH.WAVEFRONT)
        lf, la, up, fov = g.camera()
        cam = ra.camera_frame(lf, la, up, fov, 72, 40)
        P = ra.job_params(72, 40, cam, spp=12, seed=17, **kw)
        lfb, lab, upb, fovb = gb.camera()
        Pb = ra.job_params(40, 24, ra.camera_frame(lfb, lab, upb, fovb, 40, 24), spp=12, seed=3, max_depth=8)

        def run(with_binary, refill):
            ctx.set_option(ra.OPT.TRACE_REFILL, refill)
            j = ctx.job(q, P)
            other = ctx.job(b, Pb) if with_binary else None
            ctx.render_batched([j] + ([other] if other else []), 12)
            img, segs = j.resolve_linear(), j.stats().segments
            j.close()
            if other:
                other.close()
            return img, segs

        ref, segs = run(True, 0)
        log(f"quantized {scene}: per-ray traversal, {segs} segments")
        for refill, leaf_exit in ((0, -1), (1, 0), (64, 64), (16, 0)):
            ctx.set_option(ra.OPT.TRACE_LEAF_EXIT, leaf_exit)
            img, s2 = run(False, refill)
            np.testing.assert_array_equal(img, ref, err_msg=f"refill {refill}, leaf exit {leaf_exit}")
            assert s2 == segs
    finally:
        ctx.set_option(ra.OPT.TRACE_REFILL, 0)
        ctx.set_option(ra.OPT.TRACE_LEAF_EXIT, -1)
        q.close()
        b.close()


@pytest.mark.parametrize("scene,kw", [(CORN, dict(max_depth=8, flags=ra.FLAG.NEE)), (SDF, dict(max_depth=8)),
                                      (BOOK, dict(max_depth=50))], ids=["cornell_nee", "sdf", "book"])
def test_binary_wavefront_leaf_exit_and_refill_keep_images(ctx, scene, kw):
    """Binary trees on the wavefront: when a trace wave stops descending to test its lanes'
    postponed leaves (RAYON_OPT_TRACE_LEAF_EXIT) and how many idle lanes it refills at once
    (RAYON_OPT_TRACE_REFILL) change only the order of the node visits and primitive tests,
    so every setting gives the same image bit for bit, equal to the per-ray traversal (a
    pass batched with a quantized-tree job)."""
    g = ra.Generated.generate(scene[0], scene[1], detail=scene[2])
    b = ctx.scene(g.desc_ptr())
    gq = ra.Generated.generate(*BOOK[:2], detail=BOOK[2])
    q = ctx.scene(gq.desc_ptr(), ra.BVH.QUANTIZED)
    try:
        b.set_batch_path(ra.PATH.WAVEFRONT)
        q.set_batch_path(ra.PATH.WAVEFRONT)
        lf, la, up, fov = g.camera()
        P = ra.job_params(72, 40, ra.camera_frame(lf, la, up, fov, 72, 40), spp=12, seed=23, **kw)
        lfq, laq, upq, fovq = gq.camera()
        Pq = ra.job_params(40, 24, ra.camera_frame(lfq, laq, upq, fovq, 40, 24), spp=12, seed=3, max_depth=8)

        def run(with_quantized):
            j = ctx.job(b, P)
            other = ctx.job(q, Pq) if with_quantized else None
            ctx.render_batched([j] + ([other] if other else []), 12)
            img, segs = j.resolve_linear(), j.stats().segments
            j.close()
            if other:
                other.close()
            return img, segs

        ref, segs = run(True)
        for refill, leaf_exit in ((0, -1), (1, 0), (64, 64), (8, 64), (32, 0)):
            ctx.set_option(ra.OPT.TRACE_REFILL, refill)
            ctx.set_option(ra.OPT.TRACE_LEAF_EXIT, leaf_exit)
            img, s2 = run(False)
            np.testing.assert_array_equal(img, ref, err_msg=f"refill {refill}, leaf exit {leaf_exit}")
            assert s2 == segs
    finally:
        ctx.set_option(ra.OPT.TRACE_REFILL, 0)
        ctx.set_option(ra.OPT.TRACE_LEAF_EXIT, -1)
        b.close()
        q.close()


def test_several_sdf_fields_in_one_pass_equal_alone(ctx):
    """Marching and shading waves keep one SDF field in LDS and reload it when their rays
    move to another field: jobs of three different fields batched in one pass render as
    they do alone, bit for bit."""
    gens = [ra.Generated.generate(SDF[0], seed, detail=SDF[2]) for seed in (11, 12, 13)]
    scs = [ctx.scene(g.desc_ptr()) for g in gens]
    try:
        for sc in scs:
            sc.set_batch_path(ra.PATH.WAVEFRONT)

        def params(g, seed):
            lf, la, up, fov = g.camera()
            return ra.job_params(48, 32, ra.camera_frame(lf, la, up, fov, 48, 32), spp=16, seed=seed, max_depth=8)

        alone = []
        for k, (g, sc) in enumerate(zip(gens, scs)):
            j = ctx.job(sc, params(g, 100 + k))
            ctx.render_batched([j], 16)
            alone.append((j.resolve_linear(), j.stats().segments))
            j.close()
        jobs = [ctx.job(sc, params(g, 100 + k)) for k, (g, sc) in enumerate(zip(gens, scs))]
        ctx.render_batched(jobs, 16)
        for j, (ref, segs) in zip(jobs, alone):
            np.testing.assert_array_equal(j.resolve_linear(), ref)
            assert j.stats().segments == segs
            j.close()
    finally:
        for sc in scs:
            sc.close()
