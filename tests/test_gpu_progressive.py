"""GPU tests of the headless progressive renderer (rayon_progressive_*).

Tolerances: with the depth schedule off, a job rendered progressively in controller-sized
batches equals the same job rendered in one batched pass bit for bit (sample order per
pixel is fixed); the display image equals the job's RGBA8 resolve byte for byte; the
convergence percentage equals the share of negative per-pixel counts exactly.
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
def cornell(ctx):
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=4)
    sc = ctx.scene(g.desc_ptr())
    yield g, sc
    sc.close()


def params(g, W, H, **kw):
    lf, la, up, fov = g.camera()
    return ra.job_params(W, H, ra.camera_frame(lf, la, up, fov, W, H), **kw)


def test_progressive_equals_one_pass(ctx, cornell):
    g, sc = cornell
    P = params(g, 96, 64, spp=100, seed=3, max_depth=8, flags=ra.FLAG.NEE)
    ref = ctx.job(sc, P)
    ctx.render_batched([ref], 100)
    pr = ctx.progressive(sc, P, target_fps=500.0, max_batch=24, first_batch=4)
    frames = []
    while True:
        f = pr.step()
        if f.done and f.batch == 0:
            break
        frames.append((f.batch, f.next_batch, round(f.frame_ms, 2)))
        if f.done:
            break
    log(f"batches {frames}")
    assert sum(b for b, _, _ in frames) == 100 and len(frames) >= 5
    assert all(4 <= nb <= 24 for _, nb, _ in frames)
    np.testing.assert_array_equal(pr.job.resolve_linear(), ref.resolve_linear())
    img, n = pr.display()
    assert n == 100
    np.testing.assert_array_equal(img, pr.job.resolve_u8(4, 2.0))
    pr.close()
    ref.close()


def test_progressive_depth_schedule_and_reset(ctx, cornell):
    g, sc = cornell
    P = params(g, 64, 48, spp=48, seed=4, max_depth=8,
               flags=ra.FLAG.NEE | ra.FLAG.ADAPTIVE_STDERR, min_adaptive_samples=16, adaptive_threshold=0.2)
    pr = ctx.progressive(sc, P, target_fps=1000.0, max_batch=8, first_batch=4, adaptive_depth=1)
    depths = []
    while True:
        f = pr.step()
        if f.batch:
            depths.append((f.samples, f.depth))
        if f.done:
            break
    log(f"depths {depths}")
    assert depths[0][1] == 4 and all(d <= 8 for _, d in depths)
    counts = pr.job.sample_counts()
    assert abs(f.convergence - 100.0 * (counts < 0).mean()) < 1e-9
    # camera change: everything starts again
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame([lf[0] + 20, lf[1], lf[2]], la, up, fov, 64, 48)
    pr.reset(cam)
    assert pr.job.stats().samples_done == 0 and (pr.job.sample_counts() == 0).all()
    f = pr.step()
    assert f.samples == 4 and f.depth == 4
    pr.close()


def reference_depth(samples, max_depth):
    """calculateProgressiveMaxDepth (renderer_cuda_progressive_host.hpp:853-876), not moving."""
    for lim, d in ((4, 4), (16, 5), (32, 6), (64, 7), (128, 8), (512, 16), (1024, 24)):
        if samples <= lim:
            return d
    return min(512, max_depth)


def test_progressive_depth_schedule_matches_oracle(ctx, cornell):
    """The progressive renderer with the depth schedule on, against the FP64 oracle rendering
    the same batches at the same depths (sample ranges and counter streams identical). Each
    batch's depth follows calculateProgressiveMaxDepth of the samples after the batch, as
    renderBatch computes it (:884-910). Tolerance: relative RMSE <= 1%."""
    from oracle import OracleScene
    g, sc = cornell
    W, H, spp = 48, 32, 160
    P = params(g, W, H, spp=spp, seed=6, max_depth=12, flags=ra.FLAG.NEE,
               background_intensity=g.background_intensity())
    pr = ctx.progressive(sc, P, target_fps=2000.0, max_batch=32, first_batch=4, adaptive_depth=1)
    steps, before = [], 0
    while True:
        f = pr.step()
        if f.batch:
            steps.append((before, f.batch, f.depth))
            before = f.samples
        if f.done:
            break
    log(f"(sample0, batch, depth): {steps}")
    assert before == spp
    for s0, b, d in steps:
        assert d == reference_depth(s0 + b, 12), (s0, b, d)
    assert len({d for _, _, d in steps}) >= 3  # the schedule really changed depth
    img = pr.job.resolve_linear().reshape(-1, 3).astype(np.float64)
    osc = OracleScene(g.desc_ptr())
    sums = np.zeros((W * H, 3))
    for s0, b, d in steps:
        Pd = params(g, W, H, spp=spp, seed=6, max_depth=d, flags=ra.FLAG.NEE,
                    background_intensity=g.background_intensity())
        part, _ = osc.render(Pd, sample0=s0, nsamples=b)
        sums += part
    ref = sums / spp
    err = float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum()))
    log(f"progressive depth schedule vs oracle: rel RMSE {err:.5f}")
    assert err <= 0.01
    pr.close()
