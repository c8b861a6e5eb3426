#!/usr/bin/env python3
"""Throughput benchmark of the MI355X path-tracing hot path.

Workload (default): the triangle-mesh scene — Cornell box (10 triangles + quad light of
radiance 15) around a value-noise-displaced icosphere of 1,310,720 triangles with a GGX
material, 1920x1080, 64 spp, max depth 8, next-event estimation, FP32.

One step = one full 64-spp render of one 1920x1080 frame (a fresh job with its own seed,
inputs already resident in HBM). The timed region covers K such renders, bracketed by
device synchronisation. Mrays/s counts every traced segment (extension + shadow rays).
Both paths are measured: the plain per-pixel megakernel (one launch per frame) and the
batched wavefront path (one hipGraph-replayed pass per frame); `value` is the batched
one unless --path plain.

Printed: one JSON line with the metric, the roofline of the dominant kernel (compulsory
bytes per launch / its launch time from HIP events on the library's stream; the bytes the
traversal requests and the HBM traffic rocprofv3 measured for this workload beside them),
the CPU baseline on a bounded pixel subset of the same job, which also gives the relative
RMSE of the GPU image on that subset with identical random-number streams, and the metric's
latency half: p50/p99 job completion, batched vs plain, over the 256-job batching workload,
and p50/p99 time-to-first-frame and time-to-converge of the serving workload at 200 and 5
jobs/s, both with every job's spp divided by 8 so the run stays within a few minutes.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s at ≤1% RMSE vs FP64 oracle; p50/p99 job latency, batched vs plain"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per fetch of each device record (layout.h)
NODE_B, PRIM_B, MAT_B, LIGHT_B = 64, 48, 96, 64  # LIGHT_B: emitter record (primitive, emission, area)
WNODE_B = 80  # one 8-wide quantized node
REF_NODE_B, REF_PRIM_B = 64, 112  # the reference's BVHNode and Geometry records
PIXEL_RMW_B = 2 * (16 + 16 + 4)  # acc + comp + count, read and written once per launch


class HipEvents:
    """hipEvent timing on an arbitrary stream through libamdhip64."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        for n in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime",
                  "hipEventDestroy"):
            getattr(self.h, n).restype = C.c_int
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--detail", type=int, default=8, help="icosphere subdivision level (8 = 1,310,720 tris)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--path", choices=["plain", "batched"], default="batched",
                    help="which path's throughput is `value` (both are measured)")
    ap.add_argument("--cpu-stride", type=int, default=2, help="CPU baseline pixel subset: every Nth row/col")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--capacity", type=int, default=0, help="path slots per wavefront round (0: library default)")
    ap.add_argument("--chunk", type=int, default=0, help="queue entries (x64) a trace wave claims (0: default)")
    ap.add_argument("--shade-chunk", type=int, default=0, help="queue entries (x64) a shade wave claims (0: default)")
    ap.add_argument("--refill", type=int, default=0, help="idle lanes at which a trace wave takes new rays (0: default)")
    ap.add_argument("--workload", choices=["cornell_1080p", "batch256", "serve", "book_640", "sdf_720p"],
                    default="cornell_1080p",
                    help="cornell_1080p: one 1080p 64-spp mesh frame per step (default); batch256: 256 mixed "
                         "jobs, plain one-job-per-launch vs batched passes; serve: Poisson arrivals into the "
                         "server; book_640: the sphere field, 640x360, 32 spp, depth 50, CPU baseline at full "
                         "size; sdf_720p: SDF + terrain, 1280x720, adaptive 16..1024 spp to 0.5% std error")
    ap.add_argument("--jobs", type=int, default=256, help="batch256: number of jobs")
    ap.add_argument("--layout", choices=["binary", "quantized"], default="binary",
                    help="BVH node layout of the scenes")
    ap.add_argument("--pass-spp", type=int, default=32, help="batch256: samples per batched pass")
    ap.add_argument("--batch-path", choices=["auto", "wavefront", "megakernel"], default="auto",
                    help="batch256: batched path of every scene (auto: measured per scene by the autotuner)")
    ap.add_argument("--lam", type=str, default="5,50,200", help="serve: arrival rates (jobs/s), comma separated")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--grace", type=float, default=5.0, help="serve: seconds to finish after arrivals stop")
    ap.add_argument("--pool", type=int, default=4, help="serve: resident scenes per generator")
    ap.add_argument("--target-pass-ms", type=float, default=0.0,
                    help="serve: refinement pass duration the server aims for (0: library default)")
    ap.add_argument("--spp-div", type=int, default=1,
                    help="batch256 / serve: divide every job's spp (and the adaptive cap) by this")
    ap.add_argument("--no-latency", action="store_true",
                    help="cornell_1080p: skip the shortened batching and serving runs that give job latency")
    ap.add_argument("--fixed-spp", type=int, default=0,
                    help="sdf_720p: render this many samples per pixel without the adaptive rule (0: adaptive)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device-to-device 4 times (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args(argv)

    if a.copy_probe:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_per_dispatch": x.numel() * 4, "dispatches": 4}), flush=True)
        return 0

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; jobs shard across GPUs with no exchange, and one large "
              "job by sample range with one sum of the per-pixel accumulators, but neither is built "
              "here.", file=sys.stderr)
        return 2

    import rayon_amd as ra

    if a.workload == "batch256":
        return bench_batch256(a, ra)
    if a.workload == "serve":
        return bench_serve(a, ra)
    if a.workload in ("book_640", "sdf_720p"):
        return bench_scene(a, ra)

    log("generating scene")
    g = ra.Generated.generate(ra.GEN.CORNELL_ICOSPHERE, 7, detail=a.detail)
    ntris = g.desc.meshes[0].num_triangles + 10  # the mesh and the walls (plus a sphere and the light)
    ctx = ra.Context(0)
    if a.capacity:
        ctx.set_option(ra.OPT.WAVE_CAPACITY, a.capacity)
    if a.chunk:
        ctx.set_option(ra.OPT.WAVE_CHUNK, a.chunk)
    if a.shade_chunk:
        ctx.set_option(ra.OPT.SHADE_CHUNK, a.shade_chunk)
    if a.refill:
        ctx.set_option(ra.OPT.TRACE_REFILL, a.refill)
    t0 = time.perf_counter()
    sc = ctx.scene(g.desc_ptr(), ra.BVH.QUANTIZED if a.layout == "quantized" else ra.BVH.DEFAULT)
    create_ms = (time.perf_counter() - t0) * 1e3
    node_b = WNODE_B if a.layout == "quantized" else NODE_B  # bytes per visited node
    info = sc.info()
    log(f"scene: {info.num_prims} prims, {info.num_nodes} pair nodes, depth {info.max_depth}, "
        f"build {info.build_ms:.0f} ms, upload {info.upload_ms:.0f} ms, {info.device_bytes / 2**20:.1f} MiB")
    # the same scene with the tree built on the device (RAYON_BVH_DEVICE): creation time
    # here, one frame's throughput after the main measurement
    ctx.sync()
    t0 = time.perf_counter()
    dsc = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE)
    dev_create_ms = (time.perf_counter() - t0) * 1e3
    dinfo = dsc.info()
    log(f"device-built tree: {dinfo.num_nodes} pair nodes, depth {dinfo.max_depth}, build {dinfo.build_ms:.1f} ms, "
        f"upload {dinfo.upload_ms:.1f} ms, create {dev_create_ms:.0f} ms (host build: create {create_ms:.0f} ms)")
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame(lf, la, up, fov, a.width, a.height)
    flags = ra.FLAG.NEE
    W, H, spp = a.width, a.height, a.spp

    def new_job(seed, extra=0):
        return ctx.job(sc, ra.job_params(W, H, cam, spp=spp, max_depth=a.depth, seed=seed, flags=flags | extra,
                                         background_intensity=g.background_intensity()))

    def timed_steps(path):
        """K fresh 64-spp renders; wall time over the region, HIP-event time per launch."""
        jobs = [new_job(1 + i) for i in range(a.warmup + a.steps)]
        ev = HipEvents()
        stream = ctx.stream

        def run(j):
            if path == "plain":
                j.render_plain(spp)
            else:
                ctx.render_batched([j], spp)

        for j in jobs[:a.warmup]:
            log(f"warmup step ({path})")
            run(j)
        ctx.sync()
        timed = jobs[a.warmup:]
        evs = [(ev.create(), ev.create()) for _ in timed]
        ctx.sync()
        t0 = time.perf_counter()
        for j, (e0, e1) in zip(timed, evs):
            ev.record(e0, stream)
            run(j)
            ev.record(e1, stream)
        ctx.sync()
        wall = time.perf_counter() - t0
        ms = [ev.elapsed_ms(e0, e1) for e0, e1 in evs]
        stats = [j.stats() for j in timed]
        keep = timed[0]
        for j in jobs:
            if j is not keep:
                j.close()
        return wall, ms, stats, keep

    # ---- plain megakernel: counting pass (untimed) for algorithmic bytes, then K renders
    log("plain: traversal-count pass")
    cj = new_job(1000, ra.FLAG.COUNT_TRAVERSAL)
    cj.render_plain(spp)
    cs = cj.stats()
    cj.close()
    plain_bps = (node_b * cs.node_fetches + PRIM_B * cs.prim_fetches + MAT_B * cs.material_fetches
                 + LIGHT_B * cs.light_fetches) / cs.segments
    p_wall, p_ms, p_stats, p_job = timed_steps("plain")
    p_segs = sum(s.segments for s in p_stats)
    plain_mrays = p_segs / p_wall / 1e6
    p_avg_ms = float(np.mean(p_ms))
    plain_achieved = (p_segs / len(p_stats) * plain_bps + W * H * PIXEL_RMW_B) / (p_avg_ms * 1e-3) / 1e9
    log(f"plain: {plain_mrays:.0f} Mrays/s, launch {p_avg_ms:.1f} ms, {plain_bps:.0f} B/segment")
    p_job.close()

    # ---- batched wavefront (graph-replayed passes)
    b_wall, b_ms, b_stats, b_job = timed_steps("batched")
    b_segs = sum(s.segments for s in b_stats)
    batched_mrays = b_segs / b_wall / 1e6
    log(f"batched: {batched_mrays:.0f} Mrays/s, pass {np.mean(b_ms):.1f} ms")
    # per-stage HIP-event times and fetch counts from two extra (untimed) passes
    ctx.set_option(ra.OPT.STAGE_TIMING, 1)
    ctx.wave_stats(reset=True)
    tj = new_job(2000)
    ctx.render_batched([tj], spp)
    st_t = ctx.wave_stats(reset=True)
    tj.close()
    ctx.set_option(ra.OPT.STAGE_TIMING, 0)
    ctx.set_option(ra.OPT.COUNT_TRAVERSAL, 1)
    cj = new_job(2001)
    ctx.render_batched([cj], spp)
    st_c = ctx.wave_stats(reset=True)
    cjs = cj.stats()
    cj.close()
    ctx.set_option(ra.OPT.COUNT_TRAVERSAL, 0)
    ext_rays = cjs.segments - cjs.shadow_segments
    sh_rays = cjs.shadow_segments
    # trace kernel: extension rays (nodes + primitives + ray in 48 B + hit out 16 B) and
    # shadow rays (nodes + primitives + ray/contribution in 48 B + radiance RMW 32 B)
    ext_bytes = (node_b * (st_c.extend_node_fetches + st_c.shadow_node_fetches)
                 + PRIM_B * (st_c.extend_prim_fetches + st_c.shadow_prim_fetches) + 64 * ext_rays + 80 * sh_rays)
    ext_ms = st_t.stage_ms[1]
    ext_launches = st_t.launches[1]
    ext_achieved = ext_bytes / (ext_ms * 1e-3) / 1e9
    # the same visits priced at the reference's record sizes (64 B per node, 112 B per
    # primitive record): a fixed yardstick independent of the layouts used here
    ref_bytes = (REF_NODE_B * (st_c.extend_node_fetches + st_c.shadow_node_fetches)
                 + REF_PRIM_B * (st_c.extend_prim_fetches + st_c.shadow_prim_fetches) + 64 * ext_rays + 80 * sh_rays)
    ref_achieved = ref_bytes / (ext_ms * 1e-3) / 1e9
    stage_names = ["generate", "trace", "shade", "megakernel", "accumulate", "bookkeeping", "sdf_march", "unused"]
    stage_ms = {n: round(st_t.stage_ms[k], 2) for k, n in enumerate(stage_names) if n != "unused"}
    log(f"stages (ms per frame): {stage_ms}; trace {ext_achieved:.0f} GB/s algorithmic")

    # measured HBM bytes per traced ray (rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by a copy
    # kernel's counts) for this exact workload, from the committed profile of the same command
    traffic, traffic_src = None, None
    prof = Path(__file__).resolve().parent / "profiles"
    tf = prof / "k_wf_trace_cornell_icosphere_1080p_64spp.traffic.json"
    default_cfg = (W, H, spp, a.depth, a.detail, a.layout) == (1920, 1080, 64, 8, 8, "binary")
    if tf.exists() and default_cfg:
        tr = json.loads(tf.read_text())
        traffic = round(tr["hbm_bytes_per_segment"] * (ext_rays + sh_rays) / max(ext_launches, 1))
        traffic_src = f"profiles/{tf.name} (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, of this workload)"
    valu = None
    vf = prof / "valu_cornell_icosphere_1080p_64spp.json"
    if vf.exists() and default_cfg:
        valu = json.loads(vf.read_text())
    # compulsory (algorithmic) bytes of one trace launch: every ray's state in and out once
    # (extension: ray 48 B in, hit 16 B out; shadow: ray and contribution 48 B in, radiance
    # read-modify-write 32 B) and the scene's node and primitive arrays read at most once
    scene_b = int(info.device_bytes)
    state_b = 64 * ext_rays + 80 * sh_rays
    comp_per_launch = (state_b + scene_b * ext_launches) / max(ext_launches, 1)
    avg_launch_s = ext_ms / max(ext_launches, 1) * 1e-3
    comp_achieved = comp_per_launch / avg_launch_s / 1e9

    # one batched frame over the device-built tree (same job, identical image)
    dj = ctx.job(dsc, ra.job_params(W, H, cam, spp=spp, max_depth=a.depth, seed=1, flags=flags,
                                    background_intensity=g.background_intensity()))
    ctx.render_batched([dj], spp)  # warm-up frame: the timed one replays its captured pass
    dj.close()
    dj = ctx.job(dsc, ra.job_params(W, H, cam, spp=spp, max_depth=a.depth, seed=1, flags=flags,
                                    background_intensity=g.background_intensity()))
    ctx.sync()
    t0 = time.perf_counter()
    ctx.render_batched([dj], spp)
    ctx.sync()
    d_wall = time.perf_counter() - t0
    dev_tree = {"build_ms": round(dinfo.build_ms, 2), "upload_ms": round(dinfo.upload_ms, 1),
                "scene_create_ms": round(dev_create_ms, 1), "pair_nodes": int(dinfo.num_nodes),
                "depth": int(dinfo.max_depth), "builder": "device" if dinfo.builder == ra.BVH.BUILDER_DEVICE else "host",
                "batched_mrays_per_s": round(dj.stats().segments / d_wall / 1e6, 2)}
    log(f"device-built tree: {dev_tree['batched_mrays_per_s']} Mrays/s batched")
    dj.close()
    dsc.close()
    # the device-built tree collapsed into 8-wide quantized nodes: creation time and one frame
    ctx.sync()
    t0 = time.perf_counter()
    qsc = ctx.scene(g.desc_ptr(), ra.BVH.DEVICE | ra.BVH.QUANTIZED)
    q_create_ms = (time.perf_counter() - t0) * 1e3
    qinfo = qsc.info()
    for k in range(2):  # a warm-up frame, then the timed one (replaying its captured pass)
        qj = ctx.job(qsc, ra.job_params(W, H, cam, spp=spp, max_depth=a.depth, seed=1, flags=flags,
                                        background_intensity=g.background_intensity()))
        ctx.sync()
        t0 = time.perf_counter()
        ctx.render_batched([qj], spp)
        ctx.sync()
        q_wall = time.perf_counter() - t0
        q_segs = qj.stats().segments
        qj.close()
    dev_tree["quantized"] = {"scene_create_ms": round(q_create_ms, 1), "build_ms": round(qinfo.build_ms, 2),
                             "wide_nodes": int(qinfo.num_qnodes), "wide_depth": int(qinfo.wide_depth),
                             "builder": "device" if qinfo.builder == ra.BVH.BUILDER_DEVICE else "host",
                             "batched_mrays_per_s": round(q_segs / q_wall / 1e6, 2),
                             "note": "device-built binary tree read back, collapsed into 8-wide quantized nodes on "
                                     "the host, records permuted on the device; build_ms covers all of it"}
    log(f"device-built quantized tree: create {q_create_ms:.0f} ms, {dev_tree['quantized']['batched_mrays_per_s']} "
        f"Mrays/s batched")
    qsc.close()
    use_batched = a.path == "batched"
    value = batched_mrays if use_batched else plain_mrays
    seg_per_step = (b_segs / len(b_stats)) if use_batched else (p_segs / len(p_stats))
    prim_samples = sum(s.primary_samples for s in (b_stats if use_batched else p_stats))
    wall = b_wall if use_batched else p_wall
    if use_batched:
        roof = {"bound": "hbm", "achieved": round(comp_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(comp_achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_wf_trace",
                "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches_per_frame": int(ext_launches),
                "alg_bytes_per_launch": round(comp_per_launch),
                "alg_bytes_note": "compulsory bytes: ray state in/out (64 B per extension ray, 80 B per shadow "
                                  "ray) + the scene's arrays once per launch",
                "traffic_source": traffic_src,
                "traffic_frac": (None if traffic is None else round(traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS, 4)),
                "traffic_over_alg": None if traffic is None else round(traffic / comp_per_launch, 2),
                "requested_bytes_per_launch": round(ext_bytes / max(ext_launches, 1)),
                "requested_GBs": round(ext_achieved, 1),
                "requested_note": "every node (64 B) and primitive (48 B) record the traversal fetches, plus ray "
                                  "state; most are served by L2 and the 256 MB MALL, so this rate exceeds HBM peak",
                "requested_bytes_per_ray": round(ext_bytes / max(ext_rays + sh_rays, 1), 1),
                "visits_per_ray": {
                    "extension": {"nodes": round(st_c.extend_node_fetches / max(ext_rays, 1), 2),
                                  "primitives": round(st_c.extend_prim_fetches / max(ext_rays, 1), 2),
                                  "rays": int(ext_rays)},
                    "shadow": {"nodes": round(st_c.shadow_node_fetches / max(sh_rays, 1), 2),
                               "primitives": round(st_c.shadow_prim_fetches / max(sh_rays, 1), 2),
                               "rays": int(sh_rays)}},
                "requested_GBs_ref_layout": round(ref_achieved, 1),
                "valu": valu}
    else:
        plain_comp = W * H * PIXEL_RMW_B + int(info.device_bytes)
        roof = {"bound": "hbm", "achieved": round(plain_comp / (p_avg_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(plain_comp / (p_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                "traffic": None, "kernel": "k_plain", "avg_launch_ms": round(p_avg_ms, 3),
                "alg_bytes_per_launch": plain_comp,
                "alg_bytes_note": "compulsory bytes: accumulation read-modify-write per pixel + the scene once",
                "requested_GBs": round(plain_achieved, 1), "requested_bytes_per_segment": round(plain_bps, 1)}

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene generated in-process)",
        "config": {
            "workload": f"cornell_icosphere: Cornell box + displaced icosphere ({ntris} triangles) + sphere, "
                        f"{W}x{H}, {spp} spp, max depth {a.depth}, NEE, GGX conductor and dielectric",
            "path": "batched wavefront (hipGraph pass)" if use_batched else "plain megakernel",
            "width": W, "height": H, "spp": spp, "max_depth": a.depth, "triangles": int(ntris),
            "layout": a.layout,
        },
        "plain_mrays_per_s": round(plain_mrays, 2),
        "batched_mrays_per_s": round(batched_mrays, 2),
        "primary_samples_per_s": round(prim_samples / wall, 1),
        "segments_per_step": int(seg_per_step),
        "stage_ms_per_frame": stage_ms,
        "bvh": {"build_ms": round(info.build_ms, 1), "pair_nodes": int(info.num_nodes),
                "depth": int(info.max_depth), "device_MiB": round(info.device_bytes / 2**20, 1),
                "scene_create_ms": round(create_ms, 1),
                "device_build": dev_tree},
        "roofline": roof,
        "roofline_plain": {"kernel": "k_plain", "requested_GBs": round(plain_achieved, 1),
                           "avg_launch_ms": round(p_avg_ms, 3)},
    }
    timed = [b_job if use_batched else None]

    if not a.no_cpu_baseline:
        from oracle import OracleScene, cpu_baseline_lib
        job0 = b_job
        img = job0.resolve_linear().reshape(-1, 3).astype(np.float64)
        ys, xs = np.meshgrid(np.arange(a.cpu_stride // 2, H, a.cpu_stride),
                             np.arange(a.cpu_stride // 2, W, a.cpu_stride), indexing="ij")
        pix = (ys * W + xs).reshape(-1).astype(np.int64)
        cores = min(16, os.cpu_count() or 1)  # the host CPU share of one GPU
        # CPU baseline: the FP32 build of the path's restatement with the GPU's tree rule
        # (binned SAH) and the reference CPU renderer's row chunks, on the same pixels and streams
        log("cpu baseline: FP32 CPU renderer on a pixel subset")
        t_b = time.perf_counter()
        bsc = OracleScene(g.desc_ptr(), cpu_baseline_lib())
        t_bbuild = time.perf_counter() - t_b
        t_r = time.perf_counter()
        bsums, bsegs = bsc.render(job0.params, pixels=pix, nsamples=spp, threads=cores, camera=cam)
        t_cpu = time.perf_counter() - t_r
        bsc.close()
        # checker: the FP64 oracle on the same pixels and streams
        log("oracle on the same pixel subset")
        osc = OracleScene(g.desc_ptr())
        sums, osegs = osc.render(job0.params, pixels=pix, nsamples=spp, threads=cores, camera=cam)
        ref = sums / spp
        err = float(np.sqrt(((img[pix] - ref) ** 2).sum() / (ref ** 2).sum()))
        berr = float(np.sqrt(((bsums / spp - ref) ** 2).sum() / (ref ** 2).sum()))
        result["cpu_baseline"] = {
            "value": round(bsegs / t_cpu / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"FP32 multi-threaded CPU renderer (the path's restatement built in FP32, binned-SAH tree, "
                      f"one contiguous chunk of rows per thread as renderer_cpu_parallel.hpp:56-63), every "
                      f"{a.cpu_stride}th row and column of the same job ({len(pix)} pixels x {spp} spp, {bsegs} "
                      f"segments, {t_cpu:.1f} s; tree build {t_bbuild:.1f} s excluded)",
            "nproc": os.cpu_count(), "cpu": cpu_model(),
            "primary_samples_per_s": round(len(pix) * spp / t_cpu, 1),
            "rel_rmse_vs_oracle": round(berr, 5),
        }
        result["rel_rmse_vs_oracle"] = round(err, 5)
        result["rmse_note"] = f"same random streams, {spp} spp, {len(pix)}-pixel subset, linear RGB"
        log(f"cpu baseline {bsegs / t_cpu / 1e6:.2f} Mrays/s on {cores} threads; GPU rel RMSE vs oracle {err:.5f}")
        # converged image: the same scene and camera at 64x36, 4096 spp, batched path vs oracle
        cw, chh, cspp = 64, 36, 4096
        ccam = ra.camera_frame(lf, la, up, fov, cw, chh)
        cP = ra.job_params(cw, chh, ccam, spp=cspp, max_depth=job0.params.max_depth, flags=job0.params.flags, seed=31)
        cj = ctx.job(sc, cP)
        ctx.render_batched([cj], cspp)
        cimg = cj.resolve_linear().reshape(-1, 3).astype(np.float64)
        cj.close()
        t_c = time.perf_counter()
        csums, _ = osc.render(cP, nsamples=cspp, threads=cores, camera=ccam)
        cref = csums / cspp
        cerr = float(np.sqrt(((cimg - cref) ** 2).sum() / (cref ** 2).sum()))
        result["rel_rmse_4096spp"] = {"value": round(cerr, 6), "pixels": cw * chh, "spp": cspp,
                                      "note": "same scene and camera at 64x36, batched path vs FP64 oracle, "
                                              "same random streams, linear RGB"}
        log(f"converged {cspp} spp: rel RMSE {cerr:.6f} (oracle {time.perf_counter() - t_c:.1f} s)")

    b_job.close()
    sc.close()
    ctx.close()

    if not a.no_latency:
        # the metric's latency half, batched vs plain and served, on shortened runs of the
        # batching and serving workloads: the same job mixes, scenes and rates with every
        # job's spp (and the adaptive cap) divided by 8, so the default run stays short
        div = 8
        lctx = ra.Context(0)
        log("job latency: batching workload, 256 jobs, spp / 8, quantized nodes")
        rb = run_batch(a, ra, lctx, draw_jobs(256, 3, div), "quantized", 32 // div, "wavefront", oracle=False)
        lctx.close()
        lctx = ra.Context(0)
        pool = create_scenes(ra, lctx, [(gen, 100 + s) for gen in range(3) for s in range(4)], "binary")
        pool = {(gen, s - 100): v for (gen, s), v in pool.items()}
        serve = []
        for lam in (200.0, 5.0):
            log(f"job latency: serving, {lam:g} jobs/s for 10 s, spp / 8")
            serve.append(run_serve(a, ra, lctx, pool, lam, 10.0, 5.0, div))
        for g2, sc2 in pool.values():
            sc2.close()
        lctx.close()
        keep_b = ("jobs", "megapixels", "samples", "pass_spp", "layout", "batched_mrays_per_s", "plain_mrays_per_s",
                  "batched_over_plain", "latency_ms", "latency_note", "autotune", "batched_vs_plain_max_rel_diff")
        keep_s = ("lambda", "duration_s", "jobs", "completed_fraction", "first_frame_fraction", "ttff_p50_ms",
                  "ttff_p99_ms", "ttc_p50_ms", "ttc_p99_ms", "ttc_note", "ttc_p50_ms_completed",
                  "ttc_p99_ms_completed", "mrays_per_s", "offered_msamples_per_s")
        result["job_latency"] = {
            "note": "shortened runs: the batching (256 jobs) and serving (Poisson, 10 s per rate) workloads with "
                    "every job's spp divided by 8; the full-length runs are bench.py --workload batch256 / serve",
            "batched_vs_plain": {k: rb[k] for k in keep_b},
            "served": [{k: r[k] for k in keep_s} for r in serve],
        }
    print(json.dumps(result), flush=True)
    return 0


# ---------------------------------------------------------------------------------------
# Multi-job workloads
# ---------------------------------------------------------------------------------------

RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
GEN_NAMES = ["book_spheres", "cornell_icosphere", "sdf_terrain"]


def draw_jobs(n, seed, spp_div=1):
    """Job mix of the batching workload: resolution uniform over RESOLUTIONS, generator
    uniform over the three scene generators, per-job scene seed, spp uniform in 16..256
    (divided by spp_div, at least 2, in the shortened runs)."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n):
        W, H = RESOLUTIONS[int(rng.integers(len(RESOLUTIONS)))]
        spp = int(rng.integers(16, 257))
        out.append(dict(index=j, gen=int(rng.integers(3)), scene_seed=j + 1, W=W, H=H,
                        spp=max(2, spp // spp_div), seed=1000 + j))
    return out


def percentile(xs, q):
    xs = [x for x in xs]
    if not xs:
        return None
    return float(np.percentile(np.array(xs, dtype=np.float64), q))


def make_scene(ra, ctx, gen, scene_seed, layout):
    kind = [ra.GEN.BOOK_SPHERES, ra.GEN.CORNELL_ICOSPHERE, ra.GEN.SDF_TERRAIN][gen]
    g = ra.Generated.generate(kind, scene_seed)
    sc = ctx.scene(g.desc_ptr(), ra.BVH.QUANTIZED if layout == "quantized" else ra.BVH.DEFAULT)
    return g, sc


def job_params_for(ra, g, spec, depth=8, **kw):
    lf, la, up, fov = g.camera()
    cam = ra.camera_frame(lf, la, up, fov, spec["W"], spec["H"])
    flags = ra.FLAG.NEE if spec["gen"] == 1 else 0
    flags |= kw.pop("extra_flags", 0)
    return ra.job_params(spec["W"], spec["H"], cam, spp=spec["spp"], max_depth=depth, seed=spec["seed"],
                         flags=flags, background_intensity=g.background_intensity(), **kw), cam


def create_scenes(ra, ctx, keys, layout):
    """Scenes for (generator, seed) keys, generated and built on host threads in parallel
    (every call releases the interpreter lock); returns {key: (generated, scene)}."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        made = list(ex.map(lambda k: make_scene(ra, ctx, k[0], k[1], layout), keys))
    return dict(zip(keys, made))


def run_batch(a, ra, ctx, specs, layout, pass_spp, batch_path="auto", oracle=True):
    """The batching workload over `specs`: batched (autotune on clones, then passes of pass_spp
    samples over all unfinished jobs) and plain (one launch per job, submission order). All
    jobs are submitted at t = 0; a job's latency is its completion time. Returns a dict."""
    ev = HipEvents()
    t0 = time.perf_counter()
    made = create_scenes(ra, ctx, [(sp["gen"], sp["scene_seed"]) for sp in specs], layout)
    gens = [made[(sp["gen"], sp["scene_seed"])][0] for sp in specs]
    scenes = [made[(sp["gen"], sp["scene_seed"])][1] for sp in specs]
    params = [job_params_for(ra, g, sp) for g, sp in zip(gens, specs)]
    dev_bytes = sum(sc.info().device_bytes for sc in scenes)
    px = sum(sp["W"] * sp["H"] for sp in specs)
    samples = sum(sp["W"] * sp["H"] * sp["spp"] for sp in specs)
    log(f"{len(specs)} jobs, {px / 1e6:.1f} Mpixels, {samples / 1e9:.2f} G samples; scenes {dev_bytes / 2**30:.1f} GiB "
        f"created in {time.perf_counter() - t0:.0f} s")
    stream = ctx.stream

    def completion(evs, e_start, done_at):
        return [ev.elapsed_ms(e_start, evs[i]) for i in done_at]

    # ---- batched: autotune on clones, then progressive passes over all unfinished jobs
    if batch_path != "auto":
        for sc in scenes:
            sc.set_batch_path(ra.PATH.WAVEFRONT if batch_path == "wavefront" else ra.PATH.MEGAKERNEL)
    jobs = [ctx.job(scenes[k], params[k][0]) for k in range(len(specs))]
    t_tune = time.perf_counter()
    tune = ctx.autotune(jobs, 2)
    log(f"autotune: capacity {tune.capacity} chunk {tune.chunk} wg/CU {tune.wg_per_cu} "
        f"({tune.best_ms:.1f} ms vs {tune.start_ms:.1f} ms, {tune.candidates} shapes, "
        f"{time.perf_counter() - t_tune:.0f} s)")
    ctx.sync()
    e_start = ev.create()
    pass_events, done_at = [], [None] * len(jobs)
    ev.record(e_start, stream)
    tb = time.perf_counter()
    left = list(range(len(jobs)))
    done_spp = [0] * len(jobs)
    while left:
        ctx.render_batched([jobs[i] for i in left], pass_spp)
        e = ev.create()
        ev.record(e, stream)
        pass_events.append(e)
        nxt = []
        for i in left:
            done_spp[i] = min(specs[i]["spp"], done_spp[i] + pass_spp)
            if done_spp[i] >= specs[i]["spp"]:
                done_at[i] = len(pass_events) - 1
            else:
                nxt.append(i)
        left = nxt
        log(f"batched pass {len(pass_events)}: {len(left)} jobs left")
    ctx.sync()
    b_wall = time.perf_counter() - tb
    b_lat = completion(pass_events, e_start, done_at)
    b_stats = [j.stats() for j in jobs]
    b_segs = sum(s.segments for s in b_stats)
    check = [0, 1, 2] + list(range(3, len(jobs), max(1, len(jobs) // 8)))
    b_imgs = {i: jobs[i].resolve_linear() for i in check}
    for j in jobs:
        j.close()
    log(f"batched: {b_segs / b_wall / 1e6:.0f} Mrays/s, {b_wall:.1f} s")

    # ---- plain: one launch per job, jobs in submission order
    jobs = [ctx.job(scenes[k], params[k][0]) for k in range(len(specs))]
    ctx.sync()
    e_start = ev.create()
    job_events = []
    ev.record(e_start, stream)
    tp = time.perf_counter()
    for k, j in enumerate(jobs):
        j.render_plain(specs[k]["spp"])
        e = ev.create()
        ev.record(e, stream)
        job_events.append(e)
        if (k + 1) % 16 == 0:
            ctx.sync()  # keeps the log moving; each job is one long launch anyway
            log(f"plain: {k + 1} jobs done, {time.perf_counter() - tp:.0f} s")
    ctx.sync()
    p_wall = time.perf_counter() - tp
    p_lat = completion(job_events, e_start, list(range(len(jobs))))
    p_stats = [j.stats() for j in jobs]
    p_segs = sum(s.segments for s in p_stats)
    worst = 0.0
    for i in check:
        ia, ib = jobs[i].resolve_linear(), b_imgs[i]
        worst = max(worst, float((np.abs(ia - ib) / np.maximum(np.abs(ia), 1e-3)).max()))
    log(f"plain: {p_segs / p_wall / 1e6:.0f} Mrays/s, {p_wall:.1f} s; batched vs plain max rel diff {worst:.2e}")

    # ---- FP64 oracle on a pixel subset of one job per generator (same streams)
    rmse = {}
    if oracle:
        from oracle import OracleScene
        for gen in range(3):
            cand = [i for i in range(len(specs)) if specs[i]["gen"] == gen]
            if not cand:
                continue
            i = min(cand, key=lambda i: specs[i]["W"] * specs[i]["H"] * specs[i]["spp"])
            sp, (P, cam) = specs[i], params[i]
            st = 8
            ys, xs = np.meshgrid(np.arange(st // 2, sp["H"], st), np.arange(st // 2, sp["W"], st), indexing="ij")
            pix = (ys * sp["W"] + xs).reshape(-1).astype(np.int64)
            osc = OracleScene(gens[i].desc_ptr())
            sums, _ = osc.render(P, pixels=pix, nsamples=sp["spp"], threads=min(16, os.cpu_count() or 1),
                                 camera=cam)
            ref = sums / sp["spp"]
            img = jobs[i].resolve_linear().reshape(-1, 3).astype(np.float64)[pix]
            rmse[GEN_NAMES[gen]] = round(float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum())), 5)
            log(f"oracle {GEN_NAMES[gen]} job {i}: rel RMSE {rmse[GEN_NAMES[gen]]}")
    for j in jobs:
        j.close()
    for sc in scenes:
        sc.close()
    b_mrays, p_mrays = b_segs / b_wall / 1e6, p_segs / p_wall / 1e6
    return {
        "jobs": len(specs), "megapixels": round(px / 1e6, 2), "samples": samples, "pass_spp": pass_spp,
        "layout": layout, "batched_mrays_per_s": round(b_mrays, 2), "plain_mrays_per_s": round(p_mrays, 2),
        "batched_over_plain": round(b_mrays / p_mrays, 3), "batched_s": round(b_wall, 2), "plain_s": round(p_wall, 2),
        "latency_ms": {"batched_p50": round(percentile(b_lat, 50), 1), "batched_p99": round(percentile(b_lat, 99), 1),
                       "plain_p50": round(percentile(p_lat, 50), 1), "plain_p99": round(percentile(p_lat, 99), 1)},
        "latency_note": "all jobs submitted at t = 0; latency = completion time of each job",
        "segments": int(b_segs), "plain_segments": int(p_segs),
        "autotune": {"capacity": tune.capacity, "chunk": tune.chunk, "shade_chunk": tune.shade_chunk,
                     "refill": tune.refill, "wg_per_cu": tune.wg_per_cu, "mega_scenes": tune.mega_scenes,
                     "wave_scenes": tune.wave_scenes, "best_ms": round(tune.best_ms, 2),
                     "start_ms": round(tune.start_ms, 2), "shapes": tune.candidates},
        "batched_vs_plain_max_rel_diff": worst,
        "rel_rmse_vs_oracle": rmse,
    }


def bench_batch256(a, ra) -> int:
    """256 jobs of uneven size: the plain kernel one job per launch, then the batched
    wavefront (all jobs in each hipGraph-replayed pass of --pass-spp samples, launch shape
    from the autotuner). Same jobs, same seeds, same results; reports Mrays/s of both and
    p50/p99 job completion latency (all jobs submitted at t = 0)."""
    specs = draw_jobs(a.jobs, 3, a.spp_div)
    ctx = ra.Context(0)
    r = run_batch(a, ra, ctx, specs, a.layout, a.pass_spp, a.batch_path, oracle=not a.no_cpu_baseline)
    result = {
        "metric": METRIC, "value": r["batched_mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": round(r["batched_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scenes generated in-process)",
        "config": {"workload": f"batch256: {len(specs)} jobs, resolutions uniform over 256^2/512^2/1024^2/1920x1080, "
                               f"generators uniform over book spheres / Cornell icosphere (1.31 M tris, NEE) / "
                               f"SDF+terrain, per-job scene seeds, spp U{{16..256}}"
                               + (f" / {a.spp_div}" if a.spp_div > 1 else "") + f", depth 8, {a.layout} BVH",
                   "jobs": len(specs), "megapixels": r["megapixels"], "samples": r["samples"],
                   "pass_spp": a.pass_spp, "layout": a.layout},
    }
    result.update({k: v for k, v in r.items() if k not in ("jobs", "megapixels", "samples", "pass_spp", "layout")})
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


def bench_scene(a, ra) -> int:
    """The sphere field (640x360, 32 spp, depth 50) or the SDF + terrain scene (1280x720,
    adaptive: 16 spp, then passes of 64 until every pixel's relative std error is below 0.5 %
    or it has 1024 samples). Plain kernel and batched path timed on the same job; the CPU baseline is
    the FP64 oracle on the host cores, over every pixel of the sphere field and over a pixel
    subset (every 4th row and column, 16 spp) of the SDF scene, where the relative RMSE is
    taken against a 16-spp GPU render of the same pixels and streams."""
    book = a.workload == "book_640"
    gen = 0 if book else 2
    kind = ra.GEN.BOOK_SPHERES if book else ra.GEN.SDF_TERRAIN
    g = ra.Generated.generate(kind, 1984 if book else 11)
    ctx = ra.Context(0)
    sc = ctx.scene(g.desc_ptr(), ra.BVH.QUANTIZED if a.layout == "quantized" else ra.BVH.DEFAULT)
    W, H = (640, 360) if book else (1280, 720)
    fixed = 0 if book else a.fixed_spp
    spec = dict(gen=gen, W=W, H=H, spp=32 if book else (fixed or 1024), seed=1234)
    extra = {} if (book or fixed) else dict(extra_flags=ra.FLAG.ADAPTIVE_STDERR, min_adaptive_samples=16,
                                            adaptive_threshold=0.005)
    depth = 50 if book else 8
    P, cam = job_params_for(ra, g, spec, depth=depth, **dict(extra))

    def render(batched):
        j = ctx.job(sc, P)
        ctx.sync()
        t0 = time.perf_counter()
        if book or fixed:
            if batched:
                ctx.render_batched([j], spec["spp"])
            else:
                j.render_plain(spec["spp"])
        else:  # the first 16 samples, then passes of 64 until the job is done
            n = 16
            while not j.stats().done:
                if batched:
                    ctx.render_batched([j], n)
                else:
                    j.render_plain(n)
                n = 64
        ctx.sync()
        wall = time.perf_counter() - t0
        st = j.stats()
        return j, wall, st

    res = {}
    keep = None  # the last batched frame, for the image comparison
    for name, batched in (("plain", False), ("batched", True)):
        j, _, _ = render(batched)  # warm-up
        j.close()
        walls, segs = [], 0
        for _ in range(max(1, a.steps)):
            j, wall, st = render(batched)
            walls.append(wall)
            segs += st.segments
            last = st
            if keep is not None:
                keep.close()
                keep = None
            if batched:
                keep = j
            else:
                j.close()
        res[name] = {"mrays_per_s": round(segs / sum(walls) / 1e6, 2), "ms_per_frame": round(1e3 * np.mean(walls), 2),
                     "segments_per_frame": int(last.segments), "samples": int(last.samples_done),
                     "converged_pixels": int(last.converged_pixels)}
        log(f"{name}: {res[name]}")
    out = {"metric": METRIC, "value": res["batched"]["mrays_per_s"], "unit": "Mrays/s", "n_gpus": 1,
           "steps": max(1, a.steps), "warmup": 1, "ms_per_step": res["batched"]["ms_per_frame"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (procedural scene generated in-process)",
           "config": {"workload": ("book spheres (~480 spheres), 640x360, 32 spp, depth 50" if book else
                                   "SDF (64 primitives, smooth union k=0.1) + 200k-triangle terrain, 1280x720, "
                                   + (f"{fixed} spp" if fixed else
                                      "adaptive 16..1024 spp to 0.5% relative std error (passes of 16, then 64)")
                                   + ", depth 8"),
                      "layout": a.layout, "scene_path": int(sc.batch_path)},
           "plain_mrays_per_s": res["plain"]["mrays_per_s"], "batched_mrays_per_s": res["batched"]["mrays_per_s"],
           "batched_over_plain": round(res["batched"]["mrays_per_s"] / max(res["plain"]["mrays_per_s"], 1e-9), 3),
           "frames": res}
    if not a.no_cpu_baseline:
        from oracle import OracleScene
        cores = min(16, os.cpu_count() or 1)
        osc = OracleScene(g.desc_ptr())
        if book:
            pix, nspp, Pc = None, 32, P
            img = keep.resolve_linear().reshape(-1, 3).astype(np.float64)
        else:
            ys, xs = np.meshgrid(np.arange(2, H, 4), np.arange(2, W, 4), indexing="ij")
            pix = (ys * W + xs).reshape(-1).astype(np.int64)
            nspp = 16
            Pc, _ = job_params_for(ra, g, dict(spec, spp=16), depth=depth)
            fj = ctx.job(sc, Pc)
            ctx.render_batched([fj], 16)
            img = fj.resolve_linear().reshape(-1, 3).astype(np.float64)[pix]
            fj.close()
        from oracle import cpu_baseline_lib
        bsc = OracleScene(g.desc_ptr(), cpu_baseline_lib())
        t0 = time.perf_counter()
        bsums, bsegs = bsc.render(Pc, pixels=pix, nsamples=nspp, threads=cores, camera=cam)
        t_cpu = time.perf_counter() - t0
        bsc.close()
        sums, osegs = osc.render(Pc, pixels=pix, nsamples=nspp, threads=cores, camera=cam)
        ref = sums / nspp
        err = float(np.sqrt(((img - ref) ** 2).sum() / (ref ** 2).sum()))
        berr = float(np.sqrt(((bsums / nspp - ref) ** 2).sum() / (ref ** 2).sum()))
        npx = W * H if pix is None else len(pix)
        out["cpu_baseline"] = {"value": round(bsegs / t_cpu / 1e6, 3), "unit": "Mrays/s", "cores": cores,
                               "kind": "port",
                               "sample": (f"FP32 multi-threaded CPU renderer (binned-SAH tree, row chunks per "
                                          f"thread), {'every pixel' if book else 'every 4th row and column'} "
                                          f"({npx} pixels x {nspp} spp, {bsegs} segments, {t_cpu:.1f} s)"),
                               "nproc": os.cpu_count(), "cpu": cpu_model(), "rel_rmse_vs_oracle": round(berr, 6)}
        out["rel_rmse_vs_oracle"] = round(err, 6)
        out["rmse_note"] = f"same random streams, {nspp} spp, {npx} pixels, linear RGB"
        log(f"cpu baseline {bsegs / t_cpu / 1e6:.2f} Mrays/s on {cores} threads; GPU rel RMSE {err:.6f}")
        osc.close()
    keep.close()
    sc.close()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def run_serve(a, ra, ctx, pool, lam, duration, grace, spp_div=1, target_pass_ms=0.0):
    """Poisson arrivals (rate lam, `duration` s, seed 42 + lam) into a render server. Jobs are
    drawn like the batching workload from the scene pool; half fixed spp U{16..256}, half a
    noise target U(0.5%, 2%) (relative std error, 16..1024 spp); spp divided by spp_div in the
    shortened runs. After the arrivals, `grace` s to finish, then the rest is cancelled.
    Time-to-first-frame and time-to-converge are reported over all arrivals; a job that did
    not finish counts with its time up to the end of the run (censored), and the fraction of
    completed jobs is given beside it."""
    npool = max(s for _, s in pool) + 1
    rng = np.random.default_rng(42 + int(lam))
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / lam)
        if t >= duration:
            break
        arrivals.append(t)
    specs = draw_jobs(len(arrivals), 7 + int(lam), spp_div)
    plist = []
    for sp in specs:
        g, sc = pool[(sp["gen"], int(rng.integers(npool)))]
        if rng.random() < 0.5:
            P, _ = job_params_for(ra, g, sp)
        else:
            sp = dict(sp, spp=max(2, 1024 // spp_div))
            P, _ = job_params_for(ra, g, sp, extra_flags=ra.FLAG.ADAPTIVE_STDERR,
                                  min_adaptive_samples=max(2, 16 // spp_div),
                                  adaptive_threshold=float(rng.uniform(0.005, 0.02)))
        plist.append((sc, P))
    srv = ctx.server(**({"target_pass_ms": target_pass_ms} if target_pass_ms > 0 else {}))
    ids = []
    t0 = time.perf_counter()
    next_log = 5.0
    for (sc, P), ta in zip(plist, arrivals):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(sc, P))
        if ta >= next_log:
            st = srv.stats()
            log(f"lam {lam:g}: t={ta:.0f}s submitted {len(ids)} active {st.active} queued {st.queued} "
                f"completed {st.completed}")
            next_log += 5.0
    try:
        srv.drain(grace)
    except ra.RayonError:
        pass
    wall = time.perf_counter() - t0
    infos = [srv.info(i) for i in ids]
    for i, inf in zip(ids, infos):
        if inf.state in (ra.JOB.QUEUED, ra.JOB.RUNNING):
            srv.cancel(i)
    srv.drain(60.0)
    infos = [srv.info(i) for i in ids]
    segs = sum(inf.segments for inf in infos)
    st = srv.stats()
    srv.close()
    done = [inf.state == ra.JOB.DONE for inf in infos]
    censor = [(wall - ta) * 1e3 for ta in arrivals]  # arrival to the end of the run
    ttff = [inf.ttff_ms if inf.ttff_ms >= 0 else c for inf, c in zip(infos, censor)]
    ttc = [inf.ttc_ms if (d and inf.ttc_ms >= 0) else c for inf, d, c in zip(infos, done, censor)]
    ttc_done = [inf.ttc_ms for inf, d in zip(infos, done) if d and inf.ttc_ms >= 0]
    mrays = segs / wall / 1e6
    r = {"lambda": lam, "duration_s": duration, "jobs": len(ids), "completed": int(sum(done)),
         "completed_fraction": round(sum(done) / max(len(ids), 1), 4),
         "first_frame_fraction": round(sum(inf.ttff_ms >= 0 for inf in infos) / max(len(ids), 1), 4),
         "ttff_p50_ms": percentile(ttff, 50), "ttff_p99_ms": percentile(ttff, 99),
         "ttc_p50_ms": percentile(ttc, 50), "ttc_p99_ms": percentile(ttc, 99),
         "ttc_note": "over all arrivals; unfinished jobs count with arrival-to-end-of-run time (censored)",
         "ttc_p50_ms_completed": percentile(ttc_done, 50), "ttc_p99_ms_completed": percentile(ttc_done, 99),
         "mrays_per_s": round(mrays, 2), "wall_s": round(wall, 1), "passes": int(st.passes),
         "first_passes": int(st.first_passes),
         "first_frames_kept": int(st.first_kept),
         "first_frames_kept_fraction": round(st.first_kept / max(len(ids), 1), 4),
         "first_kept_note": "first frames of jobs admitted at arrival are their jobs' first samples; a queued "
                            "job's first frame is rendered before it has buffers and its sample 0 again later",
         "refine_s": round(st.pass_ms / 1e3, 2), "first_pass_s": round(st.first_ms / 1e3, 2),
         "stats_read_s": round(st.check_ms / 1e3, 2),
         "refine_msamples_per_s": round(st.pass_paths / max(st.pass_ms, 1e-9) / 1e3, 1),
         "segments_by_generator": {GEN_NAMES[gi]: int(sum(inf.segments for inf, sp in zip(infos, specs)
                                                         if sp["gen"] == gi)) for gi in range(3)}}
    # offered load: the sample budget (pixels x spp cap) of the arrivals per second of the run
    r["offered_msamples_per_s"] = round(sum(P.width * P.height * P.spp for _, P in plist) / max(wall, 1e-9) / 1e6, 1)
    for k in ("ttff_p50_ms", "ttff_p99_ms", "ttc_p50_ms", "ttc_p99_ms", "ttc_p50_ms_completed",
              "ttc_p99_ms_completed"):
        r[k] = None if r[k] is None or not np.isfinite(r[k]) else round(r[k], 1)
    log(json.dumps(r))
    return r


def bench_serve(a, ra) -> int:
    """Poisson arrivals into the render server at each rate of --lam (see run_serve)."""
    ctx = ra.Context(0)
    pool = create_scenes(ra, ctx, [(gen, 100 + s) for gen in range(3) for s in range(a.pool)], a.layout)
    pool = {(gen, s - 100): v for (gen, s), v in pool.items()}
    log(f"scene pool: {len(pool)} scenes")
    out = {"metric": METRIC, "unit": "Mrays/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scenes generated in-process)",
           "config": {"workload": f"serve: Poisson arrivals, jobs as batch256 (pool of {a.pool} scenes per "
                                  f"generator), 50% fixed spp U{{16..256}}, 50% noise target U(0.5%,2%)"
                                  + (f", spp / {a.spp_div}" if a.spp_div > 1 else "")
                                  + f", {a.duration:.0f} s per rate, HBM budget 200 GiB", "layout": a.layout},
           "rates": []}
    best = 0.0
    for lam in [float(x) for x in a.lam.split(",") if x]:
        r = run_serve(a, ra, ctx, pool, lam, a.duration, a.grace, a.spp_div, a.target_pass_ms)
        best = max(best, r["mrays_per_s"])
        out["rates"].append(r)
    out["value"] = round(best, 2)
    out["steps"], out["warmup"], out["ms_per_step"] = len(out["rates"]), 0, None
    for g, sc in pool.values():
        sc.close()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
