"""The allocation-free resolve (rayon_job_resolve_u8_noalloc) against rayon_job_resolve_u8.

Both are called through the C ABI on the same jobs: two calls in a row, the second on a
smaller image and with a different channel count, must give the same bytes as the
allocating entry point (bit-exact, as for all byte output). An image beyond the 64 MiB
staging array is refused with RAYON_ERR_INVALID and a message, before any launch.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import rayon_amd as ra

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _resolve(fn, ctx, job, w, h, ch):
    out = np.full((h, w, ch), 0xA5, np.uint8)
    st = fn(ctx.handle, job._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), ch, 2.0)
    return st, out


def test_noalloc_resolve_equals_allocating_resolve():
    L = ra.lib()
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=2)
    lf, la, up, fov = g.camera()
    with ra.Context(0) as ctx:
        sc = ctx.scene(g.desc_ptr())
        for (w, h, ch, spp) in ((96, 64, 4, 8), (40, 24, 3, 4)):  # the second call is smaller
            log(f"no-alloc resolve {w}x{h} x{ch}, {spp} spp")
            cam = ra.camera_frame(lf, la, up, fov, w, h)
            job = ctx.job(sc, ra.job_params(w, h, cam, spp=spp, max_depth=4, seed=9, flags=ra.FLAG.NEE))
            job.render_plain(spp)
            st_a, ref = _resolve(L.rayon_job_resolve_u8, ctx, job, w, h, ch)
            st_b, got = _resolve(L.rayon_job_resolve_u8_noalloc, ctx, job, w, h, ch)
            assert st_a == 0 and st_b == 0, (st_a, st_b, L.rayon_last_error())
            assert np.array_equal(got, ref)
            assert ref.max() > 0  # a rendered image, not an untouched buffer
            job.close()
        sc.close()


def test_noalloc_resolve_refuses_images_beyond_the_array():
    L = ra.lib()
    g = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 1, detail=0)
    lf, la, up, fov = g.camera()
    w, h = 4200, 4100  # 4 channels: 68.9 MB > 64 MiB
    with ra.Context(0) as ctx:
        sc = ctx.scene(g.desc_ptr())
        cam = ra.camera_frame(lf, la, up, fov, w, h)
        job = ctx.job(sc, ra.job_params(w, h, cam, spp=1, max_depth=2, seed=1))
        out = np.zeros((h, w, 4), np.uint8)
        st = L.rayon_job_resolve_u8_noalloc(ctx.handle, job._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), 4, 2.0)
        assert st == 1  # RAYON_ERR_INVALID
        assert b"staging array" in L.rayon_last_error()
        assert not out.any()
        job.close()
        sc.close()
