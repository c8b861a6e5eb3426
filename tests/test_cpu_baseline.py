"""The benchmark's CPU baseline (oracle/libcpubase.so): the path's restatement built in FP32
with the GPU path's tree rule (binned SAH) and the reference CPU renderer's schedule (one
contiguous chunk of rows per thread, renderer_cpu_parallel.hpp:56-63). It is timed, not
trusted: these checks only show it renders the same images as the FP64 checker on the same
random streams (relative RMSE within 3% at 16 spp, where FP32 path decisions occasionally
part from FP64 ones) and the same closest hits on random rays up to near-ties, whatever the
thread count."""
import numpy as np

import rayon_amd as ra
from oracle import OracleScene, cpu_baseline_lib


def test_cpu_baseline_renders_like_the_oracle():
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=3)
    W, H, spp = 40, 30, 16
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    P = ra.job_params(W, H, cam, spp=spp, max_depth=8, seed=2, flags=ra.FLAG.NEE,
                      background_intensity=g.background_intensity())
    ref, rsegs = OracleScene(g.desc_ptr()).render(P, nsamples=spp, camera=cam)
    base = OracleScene(g.desc_ptr(), cpu_baseline_lib())
    imgs = [base.render(P, nsamples=spp, camera=cam, threads=t)[0] for t in (1, 3, 8)]
    for img in imgs[1:]:
        np.testing.assert_array_equal(img, imgs[0])  # the row chunks change nothing
    err = float(np.sqrt(((imgs[0] - ref) ** 2).sum() / (ref ** 2).sum()))
    assert err <= 0.03, err


def test_cpu_baseline_closest_hits_match_the_oracle():
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=4)
    rng = np.random.default_rng(4)
    n = 20000
    o = rng.uniform(0, 555, (n, 3)).astype(np.float32)
    d = rng.normal(0, 1, (n, 3)).astype(np.float32)
    op, ot = OracleScene(g.desc_ptr()).trace_closest(o, d)
    bp, bt = OracleScene(g.desc_ptr(), cpu_baseline_lib()).trace_closest(o, d)
    same = bp == op
    assert same.mean() >= 0.999
    ok = same & (op >= 0)
    np.testing.assert_allclose(bt[ok], ot[ok], rtol=1e-4)
