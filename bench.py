#!/usr/bin/env python
"""Benchmark: stereo foveated frames/s with p50/p99 frame latency, batched
(one shared preprocess for both eyes, _C_stereo.render) against the plain
one-view-per-launch path (two _C.render calls), on one MI355X.

Workload (the stereo foveated configuration): 3M Gaussians from the seeded
generator in a 20 m cube (log-scales N(-4.5,0.6), uniform quaternions,
opacity sigmoid(N(0,1.5)), SH degree 3), two eyes of 2064x2208 at 100 deg
horizontal fov with 63 mm IPD, K=8, around a seeded gaze within 15 deg of the
axis, bilinear foveal upsampling. Three foveation levels (--levels 3, the
default): full resolution within 15 deg of the gaze, half resolution to
35 deg, quarter beyond -- radii fx*tan(angle) in pixels (232 and 606 px at
this fov), with a blend ring of a third of the full-resolution radius as in
the reference's 375/125 px; --levels 2 uses the reference's two levels
(375/125 px). Synthetic data; inputs are resident in HBM before timing.

Prints one JSON line. `roofline` is the dominant kernel's algorithmic rate
from HIP events on the render stream, over frames in which each kernel runs
alone (the stage profile; in the timed frames the two eyes' branches overlap);
`cpu_baseline` is the scalar oracle on the host's cores (one thread) rendering
one whole stereo frame of the same workload.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def progress(msg: str) -> None:
    """One line per finished phase on stderr (the JSON line goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


# per visible Gaussian and view the preprocess writes the 64-B record row,
# the 48-B support-ellipse record and 8 B of tile bounds; the blend reads
# both records and the 4-B index per listed instance
VIS_WRITE_BYTES = 64 + 48 + 8
INSTANCE_READ_BYTES = 4 + 64 + 48


def stage_bytes(counts: list, n_gauss: int, pixels: int) -> dict:
    """Algorithmic HBM bytes per stage for one view (per-unit figures of the
    measurement plan; k = key bytes, P = radix passes, I = instances)."""
    N, I, k, P = n_gauss, counts[4], counts[7], counts[15]
    n_small, n_vis, n_valid = counts[2], max(counts[8], 0), max(counts[5], 0)
    return {
        "preprocess": N * (236 + 4) + n_vis * VIS_WRITE_BYTES,
        "instances": 8 * N + 8 * n_vis + I * (k + 4),
        "sort": I * k + 2 * P * I * (k + 4),
        "ranges": I * k + 8 * n_small,
        "blend": n_valid * INSTANCE_READ_BYTES + 12 * pixels,
    }


def foveation_radii(levels: int, width: int, fov_deg: float) -> tuple[int, int, int]:
    """(foveation radius, blend width, quarter radius) in pixels."""
    if levels == 2:
        return 375, 125, 0
    fx = 0.5 * width / math.tan(math.radians(0.5 * fov_deg))
    fr = int(round(fx * math.tan(math.radians(15.0))))
    qr = int(round(fx * math.tan(math.radians(35.0))))
    return fr, int(round(fr / 3.0)), qr


def blend_evaluations(handle, view: int, n_small: int) -> int:
    """Sum over the frame's partitioned tiles of instances x samples (64 for
    FOVEA/BLENDED 8x8 tiles and PERIPHERY 16x16 tiles at stride 2, 16 for
    QUARTER tiles at stride 4): the per-pixel evaluations the blend's
    algorithm does."""
    rng = np.frombuffer(handle.debug_read(view, "ranges", 8 * n_small), dtype=np.uint32).reshape(-1, 2)
    part = np.frombuffer(handle.debug_read(view, "partition", 24), dtype=np.uint32)
    quart = np.frombuffer(handle.debug_read(view, "partition_quarter", 8), dtype=np.uint32)
    n_act = int(max(part.max(), quart.max()))
    tmp = np.frombuffer(handle.debug_read(view, "tile_map_part", 4 * n_act), dtype=np.uint32)
    inst = (rng[:, 1].astype(np.int64) - rng[:, 0])
    evals = 0
    for a, b, samples in ((part[0], part[1], 64), (part[2], part[3], 64), (part[4], part[5], 64),
                          (quart[0], quart[1], 16)):
        if b > a:
            evals += samples * int(inst[tmp[a:b]].sum())
    return evals


def cpu_baseline(scene: dict, cams: tuple, gaze, radii: tuple, K: int, frames: int) -> dict:
    """`frames` whole stereo frames on the scalar oracle: every stage of both
    eyes, every active tile blended. Per-view set-up (the visibility and
    fovea tables, which callers pass in) is outside the timed region."""
    import oracle
    from foveatedhtgscudabackend_amd import tables as T
    fr, bw, qr = radii
    views = []
    for cam in cams:
        words = T.visibility_mask_np(cam.width, cam.height)
        tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height),
                T.fovea_area_table_np(T.fovea_radius_tiles_for(cam.width, cam.height, fr, bw)))
        views.append(oracle.view_from_camera(cam, gaze, tabs, K=K, foveation_radius_px=fr, blend_width_px=bw,
                                             quarter_radius_px=qr))
    stage_ms: dict = {}
    t0 = time.perf_counter()
    for _ in range(frames):
        for view in views:
            o = oracle.run(scene, view, want=("counts",))
            for k, v in o["stage_ms"].items():
                stage_ms[k] = stage_ms.get(k, 0.0) + v / frames
    wall = time.perf_counter() - t0
    return {"frame_ms": wall * 1e3 / frames, "wall_s": wall, "stage_ms": stage_ms}


def blend_source_sha() -> str:
    """First 16 hex digits of the SHA-256 of the blend kernel's own sources."""
    h = hashlib.sha256()
    for f in ("k_blend.hip", "fhtgs_cull.h"):
        with open(os.path.join(ROOT, "foveatedhtgscudabackend_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def traffic_record(levels: int) -> dict | None:
    """The blend's HBM bytes per dispatch from the committed PMC profile of
    the stereo workload (profiles/stereo_3M_2064x2208_<L>level_pmc_traffic.md),
    with the blend source it was measured on."""
    path = os.path.join(ROOT, "profiles", f"stereo_3M_2064x2208_{levels}level_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_record:"):
            kv = dict(tok.split("=", 1) for tok in line.split()[1:])
            if kv.get("kernel") == "blend_kernel" and int(kv.get("levels", 0)) == levels:
                return {"bytes": int(kv["bytes_per_dispatch"]), "blend_source": kv["blend_source"],
                        "profile": os.path.relpath(path, ROOT)}
    return None


def plain_config(rast, scene_t, settings, steps: int, warmup: int, handle, _lib, gaussians: int) -> dict:
    """The plain one-view call (exact frames, the reference's per-view path)
    on one configuration: ms per frame, per-stage ms from HIP events and each
    stage's algorithmic bytes as a fraction of the HBM peak."""
    import torch
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_EXACT)
    for _ in range(warmup):
        rast.render(*scene_t, settings, False, False, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        rast.render(*scene_t, settings, False, False, False)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    handle.set_option(_lib.OPT_STAGE_TIMING, 1)
    stages = []
    for _ in range(5):
        rast.render(*scene_t, settings, False, False, False)
        torch.cuda.synchronize()
        stages.append(handle.stage_times())
    handle.set_option(_lib.OPT_STAGE_TIMING, 0)
    st = {k: float(np.mean([x[k] for x in stages])) for k in stages[0]}
    c = handle.counts(0)
    b = stage_bytes(c, gaussians, settings.width * settings.height)
    return {"ms_per_frame": round(ms, 4), "fps": round(1e3 / ms, 1), "instances": c[4], "visible": c[8],
            "stage_ms": {k: round(v, 4) for k, v in st.items()},
            "stage_hbm_frac": {k: round(b[k] / (st[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                               for k in b if st.get(k, 0) > 0}}


PLAIN_CONFIGS = {
    # the plain CPU configuration and the plain GPU drop-in configuration, no foveation
    "cfg1": dict(n=100_000, cube=10.0, width=800, height=800, fov=60.0, K=4),
    "cfg2": dict(n=1_000_000, cube=20.0, width=1920, height=1080, fov=60.0, K=8),
}


def config_scene(name: str, seed: int):
    from foveatedhtgscudabackend_amd import scene as S
    c = PLAIN_CONFIGS[name]
    sc = S.make_scene(c["n"], seed=seed, cube=c["cube"])
    cam = S.orbit_camera(c["width"], c["height"], c["fov"], seed=seed + 1000)
    return sc, cam, S.seeded_gaze(cam, seed + 2000)


def cpu_config_ms(name: str, seed: int, repeats: int) -> dict:
    """The scalar oracle (one thread) on a plain configuration: median
    end-to-end ms of `repeats` runs and the stage ms of the median run."""
    import oracle
    from foveatedhtgscudabackend_amd import tables as T
    c = PLAIN_CONFIGS[name]
    sc, cam, gaze = config_scene(name, seed)
    words = T.visibility_mask_np(cam.width, cam.height)
    tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height),
            T.fovea_area_table_np(T.fovea_radius_tiles_for(cam.width, cam.height, -1, 125)))
    view = oracle.view_from_camera(cam, gaze, tabs, K=c["K"], foveation_radius_px=-1)
    runs = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        o = oracle.run(sc, view, want=("counts",))
        runs.append(((time.perf_counter() - t0) * 1e3, o["stage_ms"]))
    runs.sort(key=lambda r: r[0])
    ms, stage = runs[len(runs) // 2]
    return {"ms_per_frame": round(ms, 1), "fps": round(1e3 / ms, 4), "runs": repeats, "cores": 1,
            "stage_ms": {k: round(v, 1) for k, v in stage.items()},
            "workload": f"{c['n'] / 1e6:g}M Gaussians ({c['cube']:g} m cube), {c['width']}x{c['height']}, "
                        f"{c['fov']:g} deg, K={c['K']}, no foveation"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--latency-frames", type=int, default=500,
                    help="frames of the separate loop the p50/p99 latencies come from")
    ap.add_argument("--gaussians", type=int, default=3_000_000)
    ap.add_argument("--width", type=int, default=2064)
    ap.add_argument("--height", type=int, default=2208)
    ap.add_argument("--fov", type=float, default=100.0)
    ap.add_argument("--cube", type=float, default=20.0)
    ap.add_argument("--K", type=int, default=8)
    ap.add_argument("--levels", type=int, choices=(2, 3), default=3,
                    help="foveation: 3 = full/half/quarter resolution at 15/35 deg, 2 = the reference's 375/125 px")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--view-branches", type=int, default=None,
                    help="concurrent view branches of a frame (library default 2; 1 = views one after another)")
    ap.add_argument("--work-tile", type=int, choices=(1, 2, 4), default=None,
                    help="blend work tile as waves per workgroup: 1 = 8x8 (default), 2 = 16x8, 4 = 16x16")
    ap.add_argument("--xcd-run", type=int, default=None,
                    help="blend tiles per XCD run (library default when omitted)")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-view-per-launch comparison")
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-frames", type=int, default=1, help="stereo frames the CPU baseline renders")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the plain 100k/800x800 and 1M/1920x1080 configurations (GPU and CPU)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched sub-object of the stereo line")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sub-object of the stereo line")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run 3 plain device copies of 256 MiB (calibrates FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--workload", choices=("stereo", "batched", "serving"), default="stereo",
                    help="stereo: one stereo foveated frame (default line); batched: 24 mixed view "
                         "requests over 4 resident scenes; serving: 8 headsets at 90 Hz over 8 scenes")
    ap.add_argument("--duration", type=float, default=2.0, help="serving: seconds of requests")
    ap.add_argument("--serve-batch", type=int, default=1, help="serving: most headsets per batch")
    ap.add_argument("--serve-workers", type=int, default=2, help="serving: batches in flight")
    ap.add_argument("--headsets", type=int, default=8, help="serving: headsets (at most 8, one scene each)")
    ap.add_argument("--serve-graph", type=int, default=1, help="serving: replay captured graphs (1) or not (0)")
    ap.add_argument("--load", type=float, nargs="+", default=None,
                    help="serving: offered-load multipliers of 90 Hz per headset (capacity sweep)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the renderer shards by view/viewer with no collectives, "
              "and this build runs on a single MI355X only (--gpus 1).", file=sys.stderr)
        return 2
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)

    import torch
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer, _lib
    from foveatedhtgscudabackend_amd import scene as S

    dev = "cuda"
    t0 = time.perf_counter()
    sc = S.make_scene(args.gaussians, seed=args.seed, cube=args.cube)
    left, right = S.stereo_cameras(args.width, args.height, args.fov, seed=args.seed + 1000)
    gaze = S.seeded_gaze(left, args.seed + 2000)
    gen_s = time.perf_counter() - t0
    scene_t = S.scene_to_torch(sc, dev)
    radii = foveation_radii(args.levels, args.width, args.fov)
    fr, bw, qr = radii
    sl = S.settings_for(left, gaze, device=dev, K=args.K, foveation_radius_px=fr, blend_width_px=bw)
    sr = S.settings_for(right, gaze, device=dev, K=args.K, foveation_radius_px=fr, blend_width_px=bw)
    rast = FoveatedHTGSRasterizer(fr, bw, qr)
    handle = _lib.handle_for(torch.cuda.current_device())
    # the timed and latency frames run without the stage-timing events
    # (a profiling aid, off by default); stage_profile turns them on
    handle.set_option(_lib.OPT_STAGE_TIMING, 0)
    if args.xcd_run is not None:
        handle.set_option(_lib.OPT_BLEND_XCD_RUN, args.xcd_run)
    if args.work_tile is not None:
        handle.set_option(_lib.OPT_BLEND_WORK_TILE, args.work_tile)
    work_tile = handle.tuned_work_tile(args.width, args.height)  # the one the timed frames use
    if args.view_branches is not None:
        handle.set_option(_lib.OPT_VIEW_BRANCHES, args.view_branches)
    torch.cuda.synchronize()

    def stereo():
        return rast.render_stereo(*scene_t, sl, sr, False, False, False)

    def plain():
        a = rast.render(*scene_t, sl, False, False, False)
        b = rast.render(*scene_t, sr, False, False, False)
        return a, b

    def run(fn, steps, warmup, mode):
        """`steps` frames back to back, bracketed by device synchronisation."""
        handle.set_option(_lib.OPT_FRAME_MODE, mode)
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t_start

    def latencies(fn, frames, mode):
        """Per-frame latency: call to completion, one frame at a time."""
        handle.set_option(_lib.OPT_FRAME_MODE, mode)
        fn()
        torch.cuda.synchronize()
        lat = []
        for _ in range(frames):
            s0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - s0) * 1e3)
        return lat

    def stage_profile(fn, frames):
        """Per-stage HIP-event times of the batched frame's kernels: the
        device-count mode launches exactly what the graph replays, one event
        pair around each stage."""
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_DEVICE)
        handle.set_option(_lib.OPT_STAGE_TIMING, 1)
        fn()
        torch.cuda.synchronize()
        stages = []
        for _ in range(frames):
            fn()
            torch.cuda.synchronize()
            stages.append(handle.stage_times())
        handle.set_option(_lib.OPT_STAGE_TIMING, 0)
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
        return stages

    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # batched: both eyes in one call with one shared preprocess, no
    # mid-frame wait, replayed from a captured hipGraph
    progress(f"scene of {args.gaussians} Gaussians generated in {gen_s:.1f} s")
    total = run(stereo, args.steps, args.warmup, _lib.FRAME_GRAPH)
    fps = args.steps / total
    progress(f"stereo: {fps:.1f} frames/s over {args.steps} steps")
    lat = latencies(stereo, args.latency_frames, _lib.FRAME_GRAPH)
    stages = stage_profile(stereo, max(5, min(args.steps, 20)))
    progress(f"stereo latencies ({args.latency_frames} frames) and stage profile done")
    # slots 0 and 1 now hold the left and right eye of the stereo frame: read
    # its counts and blend evaluations before the plain leg reuses slot 0
    counts_eye = [handle.counts(0), handle.counts(1)]
    evals = [blend_evaluations(handle, v, counts_eye[v][2]) for v in range(2)]
    plain_fps = None
    plain_p50 = plain_p99 = None
    if not args.no_plain:
        # plain: one view per call, each waiting mid-frame for its counts
        # (the reference's per-view path)
        ptotal = run(plain, args.steps, args.warmup, _lib.FRAME_EXACT)
        plain_fps = args.steps / ptotal
        plat = latencies(plain, args.latency_frames, _lib.FRAME_EXACT)
        plain_p50, plain_p99 = percentile(plat, 50), percentile(plat, 99)
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
        progress(f"plain: {plain_fps:.1f} stereo frames/s")

    stage_ms = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    pixels = args.width * args.height
    bytes_eye = [stage_bytes(c, args.gaussians, pixels) for c in counts_eye]
    # the shared preprocess reads the scene once for both eyes
    pre_bytes = args.gaussians * 236 + sum(args.gaussians * 4 + max(c[8], 0) * VIS_WRITE_BYTES for c in counts_eye)
    stage_bytes_frame = {k: sum(b[k] for b in bytes_eye) for k in bytes_eye[0]}
    stage_bytes_frame["preprocess"] = pre_bytes
    stage_gbs = {k: (stage_bytes_frame[k] / (stage_ms[k] * 1e-3) / 1e9 if stage_ms.get(k, 0) > 0 else None)
                 for k in stage_bytes_frame}
    # dominant kernel by time
    kernel_stages = ("preprocess", "instances", "sort", "ranges", "blend", "upsample")
    dom = max(kernel_stages, key=lambda k: stage_ms[k])
    if dom == "blend":
        # VALU-bound: ~28 FP32 operations per (pixel, Gaussian) evaluation to
        # the cutoff test, for every sample of every tile a Gaussian is listed in
        e = sum(evals)
        achieved = 28.0 * e / (stage_ms["blend"] * 1e-3) / 1e12
        default_workload = (args.gaussians, args.width, args.height, args.K, args.fov) == (3_000_000, 2064, 2208, 8, 100.0)
        rec = traffic_record(args.levels) if default_workload else None
        fresh = rec is not None and rec["blend_source"] == blend_source_sha()
        roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS,
                "traffic": rec["bytes"] if fresh else None,
                "traffic_source": (f"{rec['profile']} (blend source {rec['blend_source']}"
                                   f"{'' if fresh else ', stale: the blend source has changed since'})") if rec else None,
                "kernel": "blend_kernel", "evaluations": e, "evaluations_per_eye": evals,
                "note": "FP32 vector ALU bound (no MFMA, no packed FP32); achieved = 28 FP32 ops x sum over tiles "
                        "of listed instances x samples, left + right eye, / blend time of both eyes with each blend "
                        "alone (stage profile, HIP events on the render stream, incl. the two work-order kernels); "
                        "traffic = calibrated FETCH_SIZE + WRITE_SIZE per dispatch (one eye) of the PMC profile "
                        "named, null when that profile predates the current blend source"}
    else:
        ach = stage_gbs[dom] if stage_gbs.get(dom) else 0.0
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": dom}

    del scene_t
    torch.cuda.empty_cache()
    configs = None
    if not args.no_configs:
        configs = {}
        for name, c in PLAIN_CONFIGS.items():
            csc, ccam, cgaze = config_scene(name, args.seed)
            ct = S.scene_to_torch(csc, dev)
            cst = S.settings_for(ccam, cgaze, device=dev, K=c["K"], foveation_radius_px=-1)
            configs[name] = plain_config(FoveatedHTGSRasterizer(-1, 125, 0), ct, cst, 20, 3, handle, _lib, c["n"])
            configs[name]["workload"] = (f"plain one-view call: {c['n'] / 1e6:g}M Gaussians ({c['cube']:g} m cube), "
                                         f"{c['width']}x{c['height']}, {c['fov']:g} deg, K={c['K']}, no foveation")
            del ct
            progress(f"plain {name}: {configs[name]['ms_per_frame']} ms per frame")
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
        torch.cuda.empty_cache()
    batched = None if args.no_batched else batched_summary(args)
    if batched:
        progress(f"batched: {batched['views_per_s']} views/s ({batched['batched_vs_plain']}x plain)")
    serving = None if args.no_serving else serving_summary(args)
    if serving:
        progress(f"serving capacity at p99 <= 11.1 ms: {serving['capacity_fps_p99_11ms']}")

    cpu = None
    if args.cpu_baseline:
        cb = cpu_baseline(sc, (left, right), gaze, radii, args.K, args.cpu_frames)
        cpu = {"value": 1000.0 / cb["frame_ms"], "unit": "stereo frames/s", "cores": 1, "kind": "port",
               "sample": (f"scalar oracle, {args.cpu_frames} whole stereo frame{'s' if args.cpu_frames != 1 else ''} "
                          f"of the same workload "
                          f"({args.gaussians} Gaussians, 2 eyes x {args.width}x{args.height}, K={args.K}, same "
                          f"foveation, every active tile blended); wall {cb['wall_s']:.1f} s; stage_ms per frame, "
                          f"both eyes"),
               "stage_ms": {k: round(v, 2) for k, v in cb["stage_ms"].items()},
               "host": host_cpu()}
        progress(f"CPU baseline: {cb['frame_ms'] / 1e3:.1f} s per stereo frame")
        if not args.no_configs:
            # the plain configurations as the measurement plan times them: median of 3, and once
            cpu["cfg1"] = cpu_config_ms("cfg1", args.seed, 3)
            cpu["cfg2"] = cpu_config_ms("cfg2", args.seed, 1)
            progress("CPU baselines of the plain configurations done")

    out = {
        "metric": "stereo foveated frames/sec + p50/p99 frame latency, batched vs plain, 1 GPU",
        "value": fps,
        "unit": "stereo frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": (f"stereo foveated frame: {args.gaussians / 1e6:g}M Gaussians ({args.cube:g} m cube), "
                                f"2 eyes x {args.width}x{args.height}, {args.fov:g} deg fov, 63 mm IPD, K={args.K}, "
                                + (f"3 levels: full res < {fr} px (15 deg, blend ring {bw} px), half res < {qr} px "
                                   f"(35 deg), quarter beyond" if qr else f"2 levels: foveation {fr}/{bw} px")
                                + ", shared preprocess"),
                   "gaussians": args.gaussians, "width": args.width, "height": args.height, "K": args.K,
                   "fov_deg": args.fov, "foveation_radius_px": fr, "blend_width_px": bw,
                   "quarter_radius_px": qr, "parallelism": "single GPU",
                   "blend_work_tile": {1: "8x8", 2: "16x8", 4: "16x16"}[work_tile]},
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "latency_frames": args.latency_frames,
        "plain_fps": plain_fps,
        "plain_p50_ms": plain_p50,
        "plain_p99_ms": plain_p99,
        "batched_vs_plain": (fps / plain_fps) if plain_fps else None,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_GBps": {k: (round(v, 1) if v else None) for k, v in stage_gbs.items()},
        "instances_per_eye": [c[4] for c in counts_eye],
        "visible_per_eye": [c[8] for c in counts_eye],
        "active_tiles_per_eye": [c[3] for c in counts_eye],
        "blend_evaluations_per_eye": evals,
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_configs": configs,
        "batched": batched,
        "serving": serving,
        "scene_gen_s": round(gen_s, 2),
    }
    print(json.dumps(out))
    return 0


def host_cpu() -> str:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible, 1 thread used"


def batched_summary(args) -> dict:
    """The batched workload (24 view requests over 4 resident scenes,
    autotuned, one graph per step) against one view per call, compactly."""
    import torch
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer, _lib
    from foveatedhtgscudabackend_amd.batch import autotune, render_batch
    scenes = batched_scenes(args.seed, args.cube)
    views = batched_views(args.seed, args.K, args.levels)
    handle = _lib.handle_for(torch.cuda.current_device())
    t0 = time.perf_counter()
    tuned = autotune(scenes, views, frames=2)
    tune_s = time.perf_counter() - t0
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    for _ in range(3):
        render_batch(scenes, views)
    torch.cuda.synchronize()
    steps = 10
    t0 = time.perf_counter()
    for _ in range(steps):
        render_batch(scenes, views)
    torch.cuda.synchronize()
    vps = steps * len(views) / (time.perf_counter() - t0)
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_EXACT)
    rast = {}

    def plain():
        for bv in views:
            key = (bv.foveation_radius_px, bv.blend_width_px, bv.quarter_radius_px)
            r = rast.setdefault(key, FoveatedHTGSRasterizer(*key))
            r.benchmark(*scenes[bv.scene], bv.settings, False, False, False)
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        plain()
    torch.cuda.synchronize()
    plain_vps = 3 * len(views) / (time.perf_counter() - t0)
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    del scenes
    torch.cuda.empty_cache()
    return {"workload": "24 view requests (640x480..2064x2208, 60..100 deg, half stereo pairs, half foveated), "
                        f"{len(views)} views over 4 resident scenes of 0.5/1.5/3/6 M Gaussians, autotuned, "
                        "one captured graph per step",
            "views_per_s": round(vps, 1), "plain_views_per_s": round(plain_vps, 1),
            "batched_vs_plain": round(vps / plain_vps, 3), "autotune": tuned, "autotune_s": round(tune_s, 2)}


P99_BUDGET_MS = 1000.0 / 90.0


def serving_summary(args) -> dict:
    """Capacity of the serving loop: 8 headsets (stereo 2064x2208 at 100
    deg, the three levels) over the 8 resident scenes (24 M Gaussians), at
    offered loads of `--load` x 90 Hz per headset, served with continuous
    batching on `--serve-workers` workers (batches of one headset, and of up
    to two) and on the plain per-request path. Capacity = the highest served stereo
    frames/s among the loads whose p99 latency is within 11.1 ms (90 Hz)."""
    import torch
    from foveatedhtgscudabackend_amd import serving as V
    width, height, fov = 2064, 2208, 100.0
    scenes = serving_scenes(args.seed, args.cube)
    radii = foveation_radii(args.levels, width, fov)
    loads = args.load or (1.0, 1.5, 2.0)
    duration = 1.0
    srv = V.Server(scenes, workers=args.serve_workers, max_batch_headsets=2, foveation=radii, keep_images=False)
    admitted = [h for h in range(8) if srv.admit(h, h, width, height)]
    sweep, warmed = [], False
    for m in loads:
        headsets = [V.HeadsetSim(h, h, seed=args.seed + 700 + h, rate_hz=90.0 * m) for h in admitted]
        reqs = V.make_requests(headsets, duration, width, height, fov, foveation=radii)
        row = {"load": m, "offered_per_s": round(len(reqs) / duration, 1)}
        if not warmed:
            srv.reset(2)
            srv.warm_up(reqs)  # singletons and pairs, on both workers
            warmed = True
        for b in (1, 2):
            for r in reqs:
                r.dropped, r.images, r.t_start, r.t_done = False, None, 0.0, 0.0
            srv.reset(b)
            st = srv.run(reqs)
            row[f"batch{b}"] = {"fps": round(st["frames_per_s"], 1), "p50_ms": round(st["p50_ms"], 3),
                                "p99_ms": round(st["p99_ms"], 3), "dropped": st["dropped"]}
        for r in reqs:
            r.dropped, r.images, r.t_start, r.t_done = False, None, 0.0, 0.0
        pl = V.run_plain(scenes, srv.admitted, reqs, foveation=radii, keep_images=False)
        row["plain"] = {"fps": round(pl["frames_per_s"], 1), "p50_ms": round(pl["p50_ms"], 3),
                        "p99_ms": round(pl["p99_ms"], 3), "dropped": pl["dropped"]}
        sweep.append(row)
        progress(f"serving load {m}x: {row}")
    cap = {}
    for key in ("batch1", "batch2", "plain"):
        ok = [row[key]["fps"] for row in sweep if row[key]["p99_ms"] <= P99_BUDGET_MS]
        cap[key] = max(ok) if ok else None
    best = max((cap[k] for k in ("batch1", "batch2") if cap[k]), default=None)
    # the saturated rate of each path (highest served rate at any latency):
    # what batching two headsets per call adds when one stereo frame already
    # fills the GPU, against what it costs in latency at the p99 bound
    sat = {key: max(row[key]["fps"] for row in sweep) for key in ("batch1", "batch2", "plain")}
    del srv, scenes
    torch.cuda.empty_cache()
    return {"workload": f"{len(admitted)} headsets x stereo {width}x{height} at {fov:g} deg, three levels, over 8 "
                        f"resident scenes of 1..6 M Gaussians (24 M), {duration:g} s of requests per load, 2 workers, "
                        "latest frame wins",
            "capacity_fps_p99_11ms": cap,
            "served_vs_plain_at_capacity": round(best / cap["plain"], 3) if best and cap["plain"] else None,
            "saturated_fps": sat,
            "batch2_vs_batch1_saturated": round(sat["batch2"] / sat["batch1"], 3) if sat["batch1"] else None,
            "sweep": sweep}


def _common(args, metric, value, unit, ms, config, extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator, inputs resident in HBM)",
           "config": config}
    out.update(extra)
    print(json.dumps(out))
    return 0


BATCH_SCENES = (500_000, 1_500_000, 3_000_000, 6_000_000)
SERVING_SCENES = (1_000_000, 1_500_000, 2_000_000, 2_500_000, 3_000_000, 3_500_000, 4_500_000, 6_000_000)


def batched_scenes(seed: int = 0, cube: float = 20.0) -> list:
    """The batched workload's 4 resident scenes (0.5/1.5/3/6 M Gaussians) on the GPU."""
    from foveatedhtgscudabackend_amd import scene as S
    return [S.scene_to_torch(S.make_scene(n, seed=seed + 10 + k, cube=cube), "cuda")
            for k, n in enumerate(BATCH_SCENES)]


def batched_views(seed: int = 0, K: int = 8, levels: int = 3) -> list:
    """24 view requests of uneven size (640x480 .. 2064x2208, 60..100 deg),
    half stereo pairs (two BatchViews), half foveated, over the 4 scenes."""
    from foveatedhtgscudabackend_amd import scene as S
    from foveatedhtgscudabackend_amd.batch import BatchView
    rng = np.random.default_rng(seed + 4)
    views = []
    for r in range(24):
        w = int(rng.integers(640, 2065)) // 16 * 16
        h = int(rng.integers(480, 2209)) // 16 * 16
        sc = int(rng.integers(0, 4))
        fov = float(rng.uniform(60.0, 100.0))
        foveated = rng.random() < 0.5
        fr, bw, qr = foveation_radii(levels, w, fov) if foveated else (-1, 125, 0)
        if rng.random() < 0.5:
            cams = S.stereo_cameras(w, h, fov, seed=seed + 3000 + r)
        else:
            cams = (S.orbit_camera(w, h, fov, seed=seed + 3000 + r),)
        gaze = S.seeded_gaze(cams[0], seed + 5000 + r)
        for cam in cams:
            st = S.settings_for(cam, gaze, K=K, foveation_radius_px=fr, blend_width_px=bw)
            views.append(BatchView(sc, st, foveation_radius_px=fr, blend_width_px=bw, quarter_radius_px=qr,
                                   clear_output=False))
    return views


def serving_scenes(seed: int = 0, cube: float = 20.0) -> list:
    """The serving workload's 8 resident scenes of 1..6 M Gaussians (24 M) on the GPU."""
    from foveatedhtgscudabackend_amd import scene as S
    return [S.scene_to_torch(S.make_scene(n, seed=seed + 20 + k, cube=cube), "cuda")
            for k, n in enumerate(SERVING_SCENES)]


def bench_batched(args) -> int:
    """24 view requests of uneven size (640x480 .. 2064x2208), half of them
    stereo pairs, half foveated, over 4 resident scenes (0.5/1.5/3/6 M
    Gaussians), as one autotuned batch replayed from a captured graph,
    against the plain path (one view per call, exact frames)."""
    import torch
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer, _lib
    from foveatedhtgscudabackend_amd.batch import autotune, render_batch
    scenes = batched_scenes(args.seed, args.cube)
    views = batched_views(args.seed, args.K, args.levels)
    handle = _lib.handle_for(torch.cuda.current_device())
    t0 = time.perf_counter()
    tuned = autotune(scenes, views, frames=2)
    tune_s = time.perf_counter() - t0
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    for _ in range(args.warmup):
        render_batch(scenes, views)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        render_batch(scenes, views)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - s0) * 1e3)
    total = time.perf_counter() - t0
    n_views = len(views)
    plain_vps = None
    if not args.no_plain:
        rast = {}
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_EXACT)

        def plain():
            for bv in views:
                key = (bv.foveation_radius_px, bv.blend_width_px, bv.quarter_radius_px)
                r = rast.setdefault(key, FoveatedHTGSRasterizer(*key))
                r.benchmark(*scenes[bv.scene], bv.settings, False, False, False)
        plain()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(max(3, args.steps // 4)):
            plain()
            torch.cuda.synchronize()
        plain_vps = max(3, args.steps // 4) * n_views / (time.perf_counter() - t1)
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    vps = args.steps * n_views / total
    fov_desc = ("three levels at 15/35 deg" if args.levels == 3 else "375/125 px")
    cfg = {"workload": f"batched: 24 view requests (640x480..2064x2208, 60..100 deg, half stereo pairs, half foveated "
                       f"({fov_desc}), K={args.K}) over 4 resident scenes of 0.5/1.5/3/6 M Gaussians, one autotuned "
                       "batch per step, graph replay", "views_per_step": n_views,
           "scenes_m": [n / 1e6 for n in BATCH_SCENES], "levels": args.levels, "parallelism": "single GPU"}
    return _common(args, "batched views/sec, 1 GPU (vs one view per call)", vps, "views/s", 1e3 * total / args.steps,
                   cfg, {"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99), "plain_views_per_s": plain_vps,
                         "batched_vs_plain": (vps / plain_vps) if plain_vps else None, "autotune": tuned,
                         "autotune_s": round(tune_s, 2), "roofline": None, "cpu_baseline": None})


def bench_serving(args) -> int:
    """8 headsets requesting stereo foveated 2064x2208 frames at 90 Hz over 8
    resident scenes of 1..6 M Gaussians (24 M), continuous batching on two
    workers, against the plain per-request path on the same arrivals."""
    import torch
    from foveatedhtgscudabackend_amd import scene as S
    from foveatedhtgscudabackend_amd import serving as V
    sizes = SERVING_SCENES
    scenes = serving_scenes(args.seed, args.cube)
    headsets = [V.HeadsetSim(h, h, seed=args.seed + 700 + h) for h in range(max(1, min(8, args.headsets)))]
    radii = foveation_radii(args.levels, args.width, args.fov)
    reqs = V.make_requests(headsets, args.duration, args.width, args.height, args.fov, foveation=radii)
    srv = V.Server(scenes, workers=args.serve_workers, max_batch_headsets=args.serve_batch,
                   graph=bool(args.serve_graph), foveation=radii, keep_images=False,
                   view_branches=args.view_branches)
    admitted = [hs.hid for hs in headsets if srv.admit(hs.hid, hs.scene, args.width, args.height)]
    reqs = [r for r in reqs if r.headset in srv.admitted]
    # warm-up before the clock: per worker and headset, one exact frame sizes
    # the slots and one more captures the graph
    srv.warm_up(reqs)
    stats = srv.run(reqs)
    plain = None
    if not args.no_plain:
        for r in reqs:
            r.dropped, r.images = False, None
        plain = V.run_plain(scenes, srv.admitted, reqs, foveation=radii, keep_images=False)
    fr, bw, qr = radii
    lv = f"3 levels {fr}/{qr} px (15/35 deg), blend ring {bw} px" if qr else f"2 levels {fr}/{bw} px"
    cfg = {"workload": f"serving: {len(headsets)} headsets x stereo {args.width}x{args.height} @ 90 Hz (jitter 1 ms), "
                       f"{lv}, gaze random walk with "
                       f"saccades, 8 resident scenes 1..6 M ({sum(sizes) / 1e6:.0f} M Gaussians), "
                       f"{args.serve_workers} workers, batches of up to {args.serve_batch} headsets, latest frame wins",
           "duration_s": args.duration, "admitted": admitted, "graph": bool(args.serve_graph),
           "rejected": srv.rejected, "parallelism": "single GPU"}
    return _common(args, "served stereo foveated frames/sec + p50/p99 latency, 1 GPU", stats["frames_per_s"],
                   "stereo frames/s", stats["wall_s"] * 1e3 / max(stats["served"], 1), cfg,
                   {"p50_ms": stats["p50_ms"], "p99_ms": stats["p99_ms"], "served": stats["served"],
                    "dropped": stats["dropped"], "requests": stats["requests"], "plain": plain,
                    "served_vs_plain": (stats["frames_per_s"] / plain["frames_per_s"]) if plain else None,
                    "roofline": None, "cpu_baseline": None})


if __name__ == "__main__":
    sys.exit(main())
