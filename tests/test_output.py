"""Output files of a render: PNG image and .txt/.json statistics.

Where the reference's sources are present, oracle/build_ref.sh compiles its output code
(utils::FileUtils in utils.hpp with the stb writer it vendors) into oracle/_ref/ref_output_dump:
the statistics files must then match the reference's byte for byte, formatDuration must give
the same text, and a PNG written by stb and one written by rayon_write_png must decode to the
same pixels. The PNG decoder below (standard library only) is itself checked by decoding the
stb files back to their input. Without the reference, the round trips and the rule-derived
duration strings still run.
"""
import json
import os
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

import rayon_amd as ra

ROOT = Path(__file__).resolve().parents[1]
REF_TOOL = ROOT / "oracle" / "_ref" / "ref_output_dump"
needs_ref = pytest.mark.skipif(not REF_TOOL.exists(), reason="the reference's output code (oracle/_ref) is not built")


def decode_png(path):
    """8-bit, non-interlaced PNG -> H x W x C array; checks the signature and every CRC."""
    data = Path(path).read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) == crc, kind
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, comp, filt, interlace = hdr
    assert depth == 8 and comp == 0 and filt == 0 and interlace == 0
    ch = {0: 1, 4: 2, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = w * ch
    assert len(raw) == h * (stride + 1)
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - ch] if i >= ch else 0
            b = prev[i]
            c = prev[i - ch] if i >= ch else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) >> 1
            else:
                pp = a + b - c
                pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 255
        out[y] = cur
        prev = cur
    return out.astype(np.uint8).reshape(h, w, ch)


def gradient(w, h):
    """fillGradientImage (utils.hpp:53-67): R = 255*y/W, G = 255*x/H, B = 100."""
    y, x = np.mgrid[0:h, 0:w]
    return np.stack([(255.0 * y / w).astype(np.uint8), (255.0 * x / h).astype(np.uint8),
                     np.full((h, w), 100, np.uint8)], -1)


@pytest.mark.parametrize("ch", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", [(1, 1), (7, 3), (33, 17)])
def test_png_round_trip(tmp_path, ch, shape):
    rng = np.random.default_rng(ch * 100 + shape[0])
    h, w = shape
    img = rng.integers(0, 256, (h, w, ch), dtype=np.uint8)
    img[h // 2:] = img[h // 2:] // 7 * 7  # smooth rows exercise the other filters
    p = tmp_path / "sub" / "dir" / f"img{ch}.png"  # parent directories are created
    ra.write_png(str(p), img)
    np.testing.assert_array_equal(decode_png(p), img)


def test_png_errors(tmp_path):
    with pytest.raises(ValueError):
        ra.write_png(str(tmp_path / "x.png"), np.zeros((2, 2, 2, 2), np.uint8))
    with pytest.raises(ra.RayonError, match="INVALID"):
        ra.write_png(str(tmp_path / "x.png"), np.zeros((2, 2, 5), np.uint8))
    (tmp_path / "file").write_text("not a directory")
    with pytest.raises(ra.RayonError, match="IO"):
        ra.write_png(str(tmp_path / "file" / "x.png"), np.zeros((2, 2, 3), np.uint8))


@needs_ref
def test_png_pixels_match_reference_writer(tmp_path):
    img = gradient(97, 61)
    raw = tmp_path / "img.raw"
    raw.write_bytes(img.tobytes())
    ref_png = tmp_path / "ref" / "img.png"
    subprocess.run([str(REF_TOOL), "png", str(raw), "97", "61", str(ref_png)], check=True,
                   stdout=subprocess.DEVNULL, timeout=60)
    np.testing.assert_array_equal(decode_png(ref_png), img)  # the decoder reads stb's file back
    ours = tmp_path / "ours" / "img.png"
    ra.write_png(str(ours), img)
    np.testing.assert_array_equal(decode_png(ours), decode_png(ref_png))


DURATIONS_NS = [0, 999_999, 1_000_000, 999_999_999, 1_000_000_000, 1_234_567_890, 9_999_999_999,
                10_000_000_000, 59_999_999_999, 60_000_000_000, 61_500_000_000, 3_725_000_000_000]


def rule_duration(ns):
    """formatDuration (utils.hpp:90-116) as stated: truncating ms/s/min, then the four forms."""
    ms, s, m = ns // 1_000_000, ns // 1_000_000_000, ns // 60_000_000_000
    if m > 0:
        return f"{m}m {s % 60}s"
    if s >= 10:
        return f"{s}s"
    if s >= 1:
        return f"{ms / 1000.0:.2f}s"
    return f"{ms}ms"


@pytest.mark.parametrize("ns", DURATIONS_NS)
def test_format_duration(ns):
    got = ra.format_duration(ns)
    assert got == rule_duration(ns)
    if REF_TOOL.exists():
        ref = subprocess.run([str(REF_TOOL), "dur", str(ns)], check=True, capture_output=True, text=True,
                             timeout=60).stdout.strip()
        assert got == ref


STATS_CASES = [
    dict(spp=64, w=1920, h=1080, rays=408_820_912, size=4567, ns=128_960_000),
    dict(spp=1, w=7, h=3, rays=0, size=0, ns=0),
    dict(spp=256, w=1024, h=1024, rays=12_345_678_901, size=123_456, ns=75_000_000_001),
    dict(spp=16, w=320, h=200, rays=3, size=10, ns=2),  # 1.5e9 rays/s: rounding of the rate
]


@pytest.mark.parametrize("case", range(len(STATS_CASES)))
def test_render_stats_files(tmp_path, case):
    c = STATS_CASES[case]
    max_depth = 16
    if REF_TOOL.exists():
        (tmp_path / "ref").mkdir()
        out = subprocess.run([str(REF_TOOL), "stats", str(tmp_path / "ref" / "out.png"), str(c["spp"]), str(c["w"]),
                              str(c["h"]), str(c["rays"]), str(c["size"]), str(c["ns"])],
                             check=True, capture_output=True, text=True, timeout=60).stdout
        max_depth = int(out.strip().splitlines()[-1])  # the reference's camera depth constant
    (tmp_path / "ours").mkdir()
    ra.write_render_stats(str(tmp_path / "ours" / "out.png"), samples_per_pixel=c["spp"], width=c["w"], height=c["h"],
                          max_depth=max_depth, rays_traced=c["rays"], render_time_ns=c["ns"],
                          image_size_bytes=c["size"])
    txt = (tmp_path / "ours" / "out.txt").read_text()
    js = (tmp_path / "ours" / "out.json").read_text()
    if REF_TOOL.exists():
        assert txt == (tmp_path / "ref" / "out.txt").read_text()
        assert js == (tmp_path / "ref" / "out.json").read_text()
    d = json.loads(js)
    assert list(d) == ["image", "samples_per_pixel", "resolution", "max_depth", "rays_traced", "image_size_bytes",
                       "rays_per_second", "render_time_ms", "render_time_pretty"]
    assert d["image"] == "out.png" and d["resolution"] == {"width": c["w"], "height": c["h"]}
    rate = round(c["rays"] / (c["ns"] * 1e-9)) if c["ns"] > 0 else 0
    assert d["rays_per_second"] == rate and d["render_time_ms"] == c["ns"] // 1_000_000
    assert d["render_time_pretty"] == rule_duration(c["ns"])
    assert txt.splitlines()[2] == f"resolution: {c['w']} x {c['h']}"


def test_render_stats_file_size_and_errors(tmp_path):
    img = tmp_path / "a.b.png"
    ra.write_png(str(img), gradient(16, 9))
    ra.write_render_stats(str(img), samples_per_pixel=4, width=16, height=9, max_depth=8, rays_traced=100,
                          render_time_ns=5_000_000)
    d = json.loads((tmp_path / "a.b.json").read_text())  # only the last extension is replaced
    assert d["image_size_bytes"] == os.path.getsize(img)
    with pytest.raises(ra.RayonError, match="IO"):  # the statistics writer creates no directory
        ra.write_render_stats(str(tmp_path / "absent" / "x.png"), samples_per_pixel=1, width=1, height=1,
                              max_depth=1, rays_traced=1, render_time_ns=1)


@pytest.mark.gpu
def test_job_png_is_the_resolved_image(tmp_path):
    ctx = ra.Context(0)
    g = ra.Generated.generate(ra.GEN.BOOK_SPHERES, 5)
    sc = ctx.scene(g.desc_ptr())
    lf, la, up, fov = g.camera()
    w, h = 64, 40
    j = ctx.job(sc, ra.job_params(w, h, ra.camera_frame(lf, la, up, fov, w, h), spp=8, max_depth=8, seed=1))
    j.render_plain(8)
    p = tmp_path / "job.png"
    j.write_png(str(p), gamma=2.0)
    np.testing.assert_array_equal(decode_png(p), j.resolve_u8(3, 2.0))
    j.close()
    sc.close()
    ctx.close()
