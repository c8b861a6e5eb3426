"""Converged images against the FP64 oracle: each configured scene generator rendered at
4096 samples per pixel on the batched path (the served/batched product path, path chosen
as in production) and by the oracle at 4096 samples with the same random-number streams.

Tolerance: relative RMSE sqrt(sum|I-O|^2 / sum|O|^2) over linear RGB <= 1 %, reported per
result. The frames are small (64x36) so the oracle's share stays at seconds on the host
cores; the scenes are the full configured ones (the Cornell box holds the 1.3 M-triangle
displaced icosphere, the SDF scene 64 primitives and the 200 k-triangle terrain).
"""
import os
import sys

import numpy as np
import pytest

import rayon_amd as ra
from oracle import OracleScene

pytestmark = pytest.mark.gpu

SPP = 4096
W, H = 64, 36

CASES = [
    ("book_spheres", ra.GEN.BOOK_SPHERES, 1984, dict(max_depth=50)),
    ("cornell_icosphere_nee", ra.GEN.CORNELL_ICOSPHERE, 7, dict(max_depth=8, flags=ra.FLAG.NEE)),
    ("sdf_terrain", ra.GEN.SDF_TERRAIN, 11, dict(max_depth=8)),
]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = ra.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name,kind,seed,kw", CASES, ids=[c[0] for c in CASES])
def test_converged_4096spp_against_oracle(ctx, name, kind, seed, kw):
    log(f"[converged] {name} {W}x{H} {SPP} spp")
    g = ra.Generated.generate(kind, seed)
    sc = ctx.scene(g.desc_ptr())
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame(lf, la, up, fov, W, H)
    P = ra.job_params(W, H, cam, spp=SPP, seed=31, **kw)
    job = ctx.job(sc, P)
    ctx.render_batched([job], SPP)
    img = job.resolve_linear().reshape(-1, 3).astype(np.float64)
    st = job.stats()
    assert st.samples_done == SPP
    osc = OracleScene(g.desc_ptr())
    sums, segs = osc.render(P, nsamples=SPP, camera=cam, threads=min(16, os.cpu_count() or 1))
    ref = sums / SPP
    err = float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum()))
    log(f"  {name}: rel RMSE {err:.6f} at {SPP} spp (path {sc.batch_path}, gpu segments {st.segments}, "
        f"oracle segments {segs})")
    assert err <= 0.01
    assert abs(st.segments - segs) / segs < 0.01
    osc.close()
    job.close()
    sc.close()
