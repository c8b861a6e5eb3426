"""GPU tests of the continuous-batching render server (rayon_server_*).

Tolerances: a job served among others must equal the same job rendered alone with
rayon_render_batched bit for bit (the random streams depend only on seed, pixel and
sample index, and each pixel accumulates its samples in sample order, whatever the
pass sizes the server's controller picks); the first frame equals a one-sample render
of the job by the plain kernel resolved to RGBA8 with gamma 2, byte for byte. A job
admitted at arrival keeps its first frame's sample as its sample 0 (the plain kernel and
the batched path give the same bits per sample), so its result and its segment count
still equal the alone render's exactly.
"""
import sys

import numpy as np
import pytest

import rayon_amd as ra

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def world():
    """A context of its own (the server turns graph capture off on its context)."""
    ctx = ra.Context(0)
    gens, scenes = {}, {}
    for key, (kind, seed, detail) in {"corn": (ra.GEN.CORNELL_ICOSPHERE, 7, 4), "book": (ra.GEN.BOOK_SPHERES, 1984, 0),
                                      "sdf": (ra.GEN.SDF_TERRAIN, 11, 65)}.items():
        gens[key] = ra.Generated.generate(kind, seed, detail=detail)
        scenes[key] = ctx.scene(gens[key].desc_ptr())
    yield ctx, gens, scenes
    for s in scenes.values():
        s.close()
    ctx.close()


def params(gens, key, W, H, **kw):
    lf, la, up, fov = gens[key].camera()
    return ra.job_params(W, H, ra.camera_frame(lf, la, up, fov, W, H), **kw)


JOBS = [("corn", 80, 60, dict(spp=24, seed=9, max_depth=8, flags=ra.FLAG.NEE)),
        ("book", 128, 72, dict(spp=16, seed=1, max_depth=50)),
        ("sdf", 64, 48, dict(spp=40, seed=2, max_depth=8)),
        ("corn", 33, 17, dict(spp=7, seed=3, max_depth=8, flags=ra.FLAG.NEE)),
        ("book", 96, 54, dict(spp=33, seed=5, max_depth=16)),
        ("corn", 120, 90, dict(spp=12, seed=4, max_depth=8))]


def alone(ctx, scenes, gens, key, W, H, kw, samples=None):
    P = params(gens, key, W, H, **kw)
    j = ctx.job(scenes[key], P)
    ctx.render_batched([j], samples or P.spp)
    return j


def test_served_jobs_equal_alone_renders(world):
    ctx, gens, scenes = world
    refs, firsts, segs = [], [], []
    for key, W, H, kw in JOBS:
        j = alone(ctx, scenes, gens, key, W, H, kw)
        refs.append(j.resolve_linear())
        segs.append(j.stats().segments)
        j.close()
        f = ctx.job(scenes[key], params(gens, key, W, H, **dict(kw, spp=1)))
        f.render_plain(1)
        firsts.append(f.resolve_u8(4, 2.0))
        f.close()
    log("alone renders done")
    with ctx.server(target_pass_ms=5.0, max_pass_samples=8) as srv:
        ids = [srv.submit(scenes[key], params(gens, key, W, H, **kw)) for key, W, H, kw in JOBS]
        srv.drain(120.0)
        st = srv.stats()
        log(f"passes {st.passes} first passes {st.first_passes} completed {st.completed} pass_samples "
            f"{st.pass_samples}")
        assert st.completed == len(JOBS) and st.failed == 0 and st.queued == 0 and st.active == 0
        # every job was admitted at arrival (the budget holds them all): each kept its first frame
        assert st.first_kept == len(JOBS)
        assert st.first_kept_paths == sum(W * H for _, W, H, _ in JOBS)
        for i, jid, (key, W, H, kw), ref, first, sg in zip(range(len(ids)), ids, JOBS, refs, firsts, segs):
            inf = srv.info(jid)
            assert inf.state == ra.JOB.DONE
            assert inf.first_kept == 1
            assert inf.samples_done == kw["spp"]
            assert inf.segments == sg, f"job {i} {key}: {inf.segments} segments served, {sg} alone"
            assert 0 <= inf.queue_ms <= inf.ttff_ms <= inf.ttc_ms
            np.testing.assert_array_equal(srv.result(jid), ref, err_msg=f"job {i} {key}")
            np.testing.assert_array_equal(srv.first_frame(jid), first, err_msg=f"first frame {i} {key}")
        for jid in ids:
            srv.release(jid)
        assert srv.stats().used_bytes == 0
        with pytest.raises(ra.RayonError):
            srv.result(ids[0])


def test_admission_by_budget(world):
    """Jobs that do not all fit the budget run in admission waves and still finish with the
    same results; a job larger than the whole budget is refused at submission. Queued jobs
    get their first frames before admission (an eighth of the budget is held for them)."""
    ctx, gens, scenes = world
    key, W, H, kw = JOBS[0]
    ref = alone(ctx, scenes, gens, key, W, H, kw)
    img = ref.resolve_linear()
    ref.close()
    f = ctx.job(scenes[key], params(gens, key, W, H, **dict(kw, spp=1)))
    f.render_plain(1)
    first = f.resolve_u8(4, 2.0)
    f.close()
    scene_bytes = scenes[key].info().device_bytes
    with ctx.server(hbm_budget_bytes=scene_bytes + 2 * W * H * 48, max_pass_pixels=W * H) as srv:
        ids = [srv.submit(scenes[key], params(gens, key, W, H, **kw)) for _ in range(5)]
        with pytest.raises(ra.RayonError, match="BUDGET"):
            srv.submit(scenes[key], params(gens, key, 640, 480, spp=4))
        srv.drain(120.0)
        assert srv.stats().completed == 5
        infos = [srv.info(i) for i in ids]
        log(f"queue ms {[round(x.queue_ms, 2) for x in infos]}, ttc ms {[round(x.ttc_ms, 2) for x in infos]}")
        log(f"ttff ms {[round(x.ttff_ms, 2) for x in infos]}")
        # the jobs do not all fit at once: the last waits until an earlier one has finished,
        # and its first frame came before its admission
        assert infos[4].queue_ms >= min(infos[0].ttc_ms, infos[1].ttc_ms) - 5.0
        assert infos[4].ttff_ms < infos[4].queue_ms
        # the first job kept its first frame's sample; the last one's came before it had
        # buffers, so its sample 0 was rendered again by a refinement pass
        assert infos[0].first_kept == 1 and infos[4].first_kept == 0
        for i in ids:
            np.testing.assert_array_equal(srv.result(i), img)
            np.testing.assert_array_equal(srv.first_frame(i), first)
            srv.release(i)


def test_cancel_and_adaptive(world):
    ctx, gens, scenes = world
    with ctx.server() as srv:
        big = srv.submit(scenes["book"], params(gens, "book", 320, 180, spp=1 << 20, max_depth=50))
        ad = srv.submit(scenes["corn"], params(gens, "corn", 64, 48, spp=4096, max_depth=8,
                                               flags=ra.FLAG.NEE | ra.FLAG.ADAPTIVE_STDERR,
                                               min_adaptive_samples=16, adaptive_threshold=0.25))
        srv.cancel(big)
        srv.drain(120.0)
        assert srv.info(big).state == ra.JOB.CANCELLED
        inf = srv.info(ad)
        log(f"adaptive job: {inf.samples_done} samples, ttc {inf.ttc_ms:.1f} ms")
        assert inf.state == ra.JOB.DONE and inf.samples_done < 4096
        assert np.isfinite(srv.result(ad)).all()
        srv.release(big)
        srv.release(ad)
        assert srv.stats().used_bytes == 0


def test_first_frames_of_many_jobs(world):
    """More new jobs than one first-frame launch holds (8), over binary and quantized trees:
    every first frame equals the job's one-sample plain render resolved to RGBA8, and every
    result equals the job rendered alone."""
    ctx, gens, scenes = world
    quant = ctx.scene(gens["corn"].desc_ptr(), ra.BVH.QUANTIZED)
    try:
        sc = dict(scenes, cornq=quant)
        gens_q = dict(gens, cornq=gens["corn"])
        jobs = []
        for i in range(20):
            key = ["corn", "book", "sdf", "cornq"][i % 4]
            kw = dict(spp=3 + i % 5, seed=100 + i, max_depth=8 if key != "book" else 20)
            if key.startswith("corn"):
                kw["flags"] = ra.FLAG.NEE
            jobs.append((key, 24 + 8 * (i % 3), 16 + 4 * (i % 4), kw))
        refs, firsts = [], []
        for key, W, H, kw in jobs:
            j = alone(ctx, sc, gens_q, key, W, H, kw)
            refs.append(j.resolve_linear())
            j.close()
            f = ctx.job(sc[key], params(gens_q, key, W, H, **dict(kw, spp=1)))
            f.render_plain(1)
            firsts.append(f.resolve_u8(4, 2.0))
            f.close()
        log("alone renders done")
        with ctx.server(target_pass_ms=5.0, max_pass_samples=2) as srv:
            ids = [srv.submit(sc[key], params(gens_q, key, W, H, **kw)) for key, W, H, kw in jobs]
            srv.drain(120.0)
            st = srv.stats()
            log(f"first passes {st.first_passes} for {len(ids)} jobs, passes {st.passes}")
            assert st.completed == len(jobs) and st.failed == 0
            for i, jid in enumerate(ids):
                np.testing.assert_array_equal(srv.first_frame(jid), firsts[i], err_msg=f"first frame {i}")
                np.testing.assert_array_equal(srv.result(jid), refs[i], err_msg=f"job {i}")
                srv.release(jid)
    finally:
        quant.close()
