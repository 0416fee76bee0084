#!/usr/bin/env python
"""Benchmark: stereo foveated frames/s with p50/p99 frame latency, batched
(one shared preprocess for both eyes, _C_stereo.render) against the plain
one-view-per-launch path (two _C.render calls), on one MI355X.

Workload (the stereo foveated configuration): 3M Gaussians from the seeded
generator in a 20 m cube (log-scales N(-4.5,0.6), uniform quaternions,
opacity sigmoid(N(0,1.5)), SH degree 3), two eyes of 2064x2208 at 100 deg
horizontal fov with 63 mm IPD, K=8, reference foveation (375 px radius,
125 px blend ring) around a seeded gaze within 15 deg of the axis, bilinear
foveal upsampling. Synthetic data; inputs are resident in HBM before timing.

Prints one JSON line. `roofline` is the dominant kernel's algorithmic rate
from HIP events on the render stream; `cpu_baseline` is the scalar oracle on
the box's host (one thread) over a bounded sample of the same frame.
"""
from __future__ import annotations

import argparse
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


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def stage_bytes(counts: list, n_gauss: int, pixels: int) -> dict:
    """Algorithmic HBM bytes per stage for one view (per-unit figures of the
    measurement plan; k = key bytes, P = radix passes, I = instances)."""
    N, I, k, P = n_gauss, counts[4], counts[7], counts[15]
    n_small, n_vis, n_valid = counts[2], max(counts[8], 0), max(counts[5], 0)
    return {
        "preprocess": N * (236 + 4) + n_vis * 72,
        "instances": 8 * N + 8 * n_vis + I * (k + 4),
        "sort": I * k + 2 * P * I * (k + 4),
        "ranges": I * k + 8 * n_small,
        "blend": n_valid * (4 + 64) + 12 * pixels,
    }


def cpu_baseline(scene: dict, cam, gaze, tile_stride: int) -> dict:
    """One eye of the frame on the scalar oracle, blending every
    `tile_stride`-th active tile; the blend time is scaled by the stride."""
    import oracle
    from foveatedhtgscudabackend_amd import tables as T
    words = T.visibility_mask_np(cam.width, cam.height)
    tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height),
            T.fovea_area_table_np(T.fovea_radius_tiles_for(cam.width, cam.height, 375, 125)))
    view = oracle.view_from_camera(cam, gaze, tabs, K=8, blend_tile_stride=tile_stride)
    t0 = time.perf_counter()
    o = oracle.run(scene, view, want=("counts",))
    wall = time.perf_counter() - t0
    st = o["stage_ms"]
    eye_ms = sum(v for kk, v in st.items() if kk not in ("blend", "upsample")) + (st["blend"] + st["upsample"]) * tile_stride
    return {"eye_ms": eye_ms, "wall_s": wall, "stage_ms": st}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--gaussians", type=int, default=3_000_000)
    ap.add_argument("--width", type=int, default=2064)
    ap.add_argument("--height", type=int, default=2208)
    ap.add_argument("--fov", type=float, default=100.0)
    ap.add_argument("--cube", type=float, default=20.0)
    ap.add_argument("--K", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--xcd-run", type=int, default=None,
                    help="blend tiles per XCD run (library default when omitted)")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-view-per-launch comparison")
    ap.add_argument("--cpu-baseline", dest="cpu_baseline", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--cpu-tile-stride", type=int, default=4)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run 3 plain device copies of 256 MiB (calibrates FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the renderer shards by view/viewer with no collectives, "
              "and this build runs on a single MI355X only (--gpus 1).", file=sys.stderr)
        return 2

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
    sl = S.settings_for(left, gaze, device=dev, K=args.K)
    sr = S.settings_for(right, gaze, device=dev, K=args.K)
    rast = FoveatedHTGSRasterizer()
    handle = _lib.handle_for(torch.cuda.current_device())
    handle.set_option(_lib.OPT_STAGE_TIMING, 1)
    if args.xcd_run is not None:
        handle.set_option(_lib.OPT_BLEND_XCD_RUN, args.xcd_run)
    torch.cuda.synchronize()

    def stereo():
        return rast.render_stereo(*scene_t, sl, sr, False, False, False)

    def plain():
        a = rast.render(*scene_t, sl, False, False, False)
        b = rast.render(*scene_t, sr, False, False, False)
        return a, b

    def run(fn, steps, warmup, mode):
        handle.set_option(_lib.OPT_FRAME_MODE, mode)
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        lat = []
        t_start = time.perf_counter()
        for _ in range(steps):
            s0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - s0) * 1e3)
        torch.cuda.synchronize()
        total = time.perf_counter() - t_start
        return total, lat

    def stage_profile(fn, frames):
        """Per-stage HIP-event times of the batched frame's kernels: the
        device-count mode launches exactly what the graph replays, one event
        pair around each stage."""
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_DEVICE)
        fn()
        torch.cuda.synchronize()
        stages = []
        for _ in range(frames):
            fn()
            torch.cuda.synchronize()
            stages.append(handle.stage_times())
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
    total, lat = run(stereo, args.steps, args.warmup, _lib.FRAME_GRAPH)
    counts_eye = [handle.counts(0), handle.counts(1)]
    fps = args.steps / total
    stages = stage_profile(stereo, max(5, min(args.steps, 20)))
    plain_fps = None
    plain_p50 = plain_p99 = None
    if not args.no_plain:
        # plain: one view per call, each waiting mid-frame for its counts
        # (the reference's per-view path)
        ptotal, plat = run(plain, args.steps, args.warmup, _lib.FRAME_EXACT)
        plain_fps = args.steps / ptotal
        plain_p50, plain_p99 = percentile(plat, 50), percentile(plat, 99)
        handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)

    stage_ms = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    pixels = args.width * args.height
    bytes_eye = [stage_bytes(c, args.gaussians, pixels) for c in counts_eye]
    # the shared preprocess reads the scene once for both eyes
    pre_bytes = args.gaussians * 236 + sum(args.gaussians * 4 + max(c[8], 0) * 72 for c in counts_eye)
    stage_bytes_frame = {k: sum(b[k] for b in bytes_eye) for k in bytes_eye[0]}
    stage_bytes_frame["preprocess"] = pre_bytes
    stage_gbs = {k: (stage_bytes_frame[k] / (stage_ms[k] * 1e-3) / 1e9 if stage_ms.get(k, 0) > 0 else None)
                 for k in stage_bytes_frame}
    # dominant kernel by time
    kernel_stages = ("preprocess", "instances", "sort", "ranges", "blend", "upsample")
    dom = max(kernel_stages, key=lambda k: stage_ms[k])
    if dom == "blend":
        # VALU-bound: ~28 FP32 operations per (pixel, Gaussian) evaluation to
        # the cutoff test; each valid instance is evaluated by 64 lanes
        evals = sum(64 * max(c[5], 0) for c in counts_eye)
        achieved = 28.0 * evals / (stage_ms["blend"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "blend_kernel",
                "note": "VALU fp32 (no MFMA); peak is MI355X fp32 vector/matrix 157.3 TFLOP/s"}
    else:
        ach = stage_gbs[dom] if stage_gbs.get(dom) else 0.0
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": dom}

    cpu = None
    if args.cpu_baseline:
        cb = cpu_baseline(sc, left, gaze, args.cpu_tile_stride)
        frame_ms = 2.0 * cb["eye_ms"]
        cpu = {"value": 1000.0 / frame_ms, "unit": "stereo frames/s", "cores": 1, "kind": "port",
               "sample": (f"scalar oracle, left eye of the same frame (3M Gaussians, 2064x2208, K=8), "
                          f"every {args.cpu_tile_stride}th active tile blended and the blend time scaled by "
                          f"{args.cpu_tile_stride}; frame = 2 eyes; wall {cb['wall_s']:.1f} s"),
               "stage_ms": {k: round(v, 2) for k, v in cb["stage_ms"].items()}}

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
        "config": {"workload": "stereo foveated frame: 3M Gaussians (20 m cube), 2 eyes x 2064x2208, "
                               "100 deg fov, 63 mm IPD, K=8, foveation 375/125 px, shared preprocess",
                   "gaussians": args.gaussians, "width": args.width, "height": args.height, "K": args.K,
                   "fov_deg": args.fov, "parallelism": "single GPU"},
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "plain_fps": plain_fps,
        "plain_p50_ms": plain_p50,
        "plain_p99_ms": plain_p99,
        "batched_vs_plain": (fps / plain_fps) if plain_fps else None,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_GBps": {k: (round(v, 1) if v else None) for k, v in stage_gbs.items()},
        "instances_per_eye": [c[4] for c in counts_eye],
        "visible_per_eye": [c[8] for c in counts_eye],
        "active_tiles_per_eye": [c[3] for c in counts_eye],
        "roofline": roof,
        "cpu_baseline": cpu,
        "scene_gen_s": round(gen_s, 2),
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
