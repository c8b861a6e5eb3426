"""The batching and serving workloads at their stated size, with property checks.

* Batching: the full 256-job draw of the batching workload (resolutions 256^2..1920x1080,
  the three scene generators with per-job seeds, spp U{16..256}, depth 8, quantized BVH
  nodes, autotuned launch shape), rendered in batched passes of 32 samples over all
  unfinished jobs, as bench.py --workload batch256 does. Checks: 8 sampled jobs equal their
  alone renders bit for bit (each pixel folds its samples in sample order whatever else
  shares a pass) and the plain kernel's render, also bit for bit; the smallest job of each generator is within 1% relative RMSE of the
  FP64 oracle on a pixel subset (every 8th row and column, same random streams).
* Serving: Poisson arrivals at 200 jobs/s for 10 s of the same job mix into the server;
  up to 8 sampled completed jobs equal their alone renders bit for bit (image and segment
  count), every arrival got a first frame, jobs admitted at arrival kept theirs as their
  first samples, and the first frames' p99 time is reported.
"""
import sys
import time

import numpy as np
import pytest

import rayon_amd as ra
from bench import GEN_NAMES, create_scenes, draw_jobs, job_params_for

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_batching_workload_full_draw():
    specs = draw_jobs(256, 3)
    ctx = ra.Context(0)
    t0 = time.perf_counter()
    made = create_scenes(ra, ctx, [(sp["gen"], sp["scene_seed"]) for sp in specs], "quantized")
    scenes = [made[(sp["gen"], sp["scene_seed"])][1] for sp in specs]
    gens = [made[(sp["gen"], sp["scene_seed"])][0] for sp in specs]
    params = [job_params_for(ra, g, sp) for g, sp in zip(gens, specs)]
    assert all(sc.info().num_qnodes > 0 for sc in scenes)
    log(f"256 scenes in {time.perf_counter() - t0:.0f} s; "
        f"{sum(sp['W'] * sp['H'] * sp['spp'] for sp in specs) / 1e9:.1f} G samples")
    for sc in scenes:
        sc.set_batch_path(ra.PATH.WAVEFRONT)
    jobs = [ctx.job(sc, P) for sc, (P, _) in zip(scenes, params)]
    tune = ctx.autotune(jobs, 2)
    log(f"autotune: chunk {tune.chunk}, shade chunk {tune.shade_chunk}, wg/CU {tune.wg_per_cu}, "
        f"capacity {tune.capacity}, {tune.candidates} shapes")
    left, done, npass = list(range(256)), [0] * 256, 0
    t0 = time.perf_counter()
    while left:
        ctx.render_batched([jobs[i] for i in left], 32)
        npass += 1
        for i in left:
            done[i] = min(specs[i]["spp"], done[i] + 32)
        left = [i for i in left if done[i] < specs[i]["spp"]]
    ctx.sync()
    log(f"batched: {npass} passes, {time.perf_counter() - t0:.1f} s")
    assert all(j.stats().samples_done == sp["spp"] for j, sp in zip(jobs, specs))
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(256, 8, replace=False).tolist())
    for i in sample:
        P = params[i][0]
        a = ctx.job(scenes[i], P)
        ctx.render_batched([a], P.spp)
        mix, alone = jobs[i].resolve_linear(), a.resolve_linear()
        assert np.array_equal(mix, alone), f"job {i} differs from its alone render"
        p = ctx.job(scenes[i], P)
        p.render_plain(P.spp)
        plain = p.resolve_linear()
        log(f"job {i} ({GEN_NAMES[specs[i]['gen']]} {P.width}x{P.height} {P.spp} spp): alone == mix, "
            f"plain identical {np.mean(plain == mix):.6f}")
        assert np.array_equal(plain, mix), f"job {i}: the plain kernel's render differs"
        a.close()
        p.close()
    from oracle import OracleScene
    for gen in range(3):
        i = min((k for k in range(256) if specs[k]["gen"] == gen),
                key=lambda k: specs[k]["W"] * specs[k]["H"] * specs[k]["spp"])
        sp, (P, cam) = specs[i], params[i]
        ys, xs = np.meshgrid(np.arange(4, sp["H"], 8), np.arange(4, sp["W"], 8), indexing="ij")
        pix = (ys * sp["W"] + xs).reshape(-1).astype(np.int64)
        sums, _ = OracleScene(gens[i].desc_ptr()).render(P, pixels=pix, nsamples=sp["spp"], camera=cam)
        ref = sums / sp["spp"]
        img = jobs[i].resolve_linear().reshape(-1, 3).astype(np.float64)[pix]
        err = float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum()))
        log(f"oracle, {GEN_NAMES[gen]} job {i} ({len(pix)} pixels, {sp['spp']} spp): rel RMSE {err:.5f}")
        assert err <= 0.01
    for j in jobs:
        j.close()
    for sc in scenes:
        sc.close()
    ctx.close()


def test_serving_poisson_200_per_s():
    ctx = ra.Context(0)
    pool = create_scenes(ra, ctx, [(gen, 100 + s) for gen in range(3) for s in range(4)], "binary")
    rng = np.random.default_rng(242)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / 200.0)
        if t >= 10.0:
            break
        arrivals.append(t)
    specs = draw_jobs(len(arrivals), 207)
    plist = []
    for sp in specs:
        key = (sp["gen"], 100 + int(rng.integers(4)))
        g, sc = pool[key]
        plist.append((key, job_params_for(ra, g, sp)[0]))
    srv = ctx.server()
    ids = []
    t0 = time.perf_counter()
    for (key, P), ta in zip(plist, arrivals):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(pool[key][1], P))
    log(f"{len(ids)} arrivals in {time.perf_counter() - t0:.1f} s")
    try:
        srv.drain(5.0)
    except ra.RayonError:
        pass
    infos = [srv.info(i) for i in ids]
    done = [k for k, inf in enumerate(infos) if inf.state == ra.JOB.DONE]
    for i, inf in zip(ids, infos):
        if inf.state in (ra.JOB.QUEUED, ra.JOB.RUNNING):
            srv.cancel(i)
    srv.drain(60.0)
    ttff = np.array([srv.info(i).ttff_ms for i in ids])
    st = srv.stats()
    log(f"completed {len(done)} of {len(ids)}; first frames {int((ttff >= 0).sum())} ({st.first_kept} kept as "
        f"their jobs' first samples), TTFF p50 {np.percentile(ttff, 50):.1f} ms p99 {np.percentile(ttff, 99):.1f} ms")
    assert (ttff >= 0).all()
    assert st.first_kept > 0
    assert len(done) >= 8
    sample = sorted(np.random.default_rng(3).choice(done, 8, replace=False).tolist())
    for k in sample:
        key, P = plist[k]
        served = srv.result(ids[k])
        a = ctx.job(pool[key][1], P)
        ctx.render_batched([a], P.spp)
        assert np.array_equal(served, a.resolve_linear()), f"served job {k} differs from its alone render"
        assert srv.info(ids[k]).segments == a.stats().segments
        a.close()
    log(f"served == alone for jobs {sample} (first frames kept: {[srv.info(ids[k]).first_kept for k in sample]})")
    srv.close()
    for g, sc in pool.values():
        sc.close()
    ctx.close()
