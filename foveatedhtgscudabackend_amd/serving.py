"""Serving: a request loop for stereo foveated frames from many headsets.

Each simulated headset asks for a stereo frame at a fixed rate (90 Hz with a
normal jitter of the arrival time); its gaze follows a random walk with
saccades (exponential intervals, 10-40 degree jumps) and its head a smooth
random walk on a sphere around the scene. The server keeps every scene
resident, gives every admitted headset two persistent handle slots (one per
eye) so its buffers and instance capacity stay on the device, and admits a
headset only while the device has the memory its slots need.

Continuous batching: ``workers`` threads, each with its own library handle
and HIP stream, take whatever requests have arrived when they become free and
render them as one batch (``fhtgs_render_batch``, both eyes of a headset
sharing one preprocess, replayed from a captured graph when the batch shape
repeats). While one batch runs, requests keep arriving and the next free
worker takes them, so batches overlap on the device. A headset holds at most
one waiting request: a newer frame replaces an older one that has not started
(it is counted as dropped), which bounds the latency under overload.

The plain path for comparison renders each request alone with
``_C_stereo.render`` semantics (one call per request, exact frame mode, one
worker). Every served image equals the image of the same request rendered
alone (deterministic pipeline); tests check this.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _binding, _lib
from . import scene as S


# ---------------------------------------------------------------- headsets

@dataclass
class HeadsetSim:
    """Seeded motion of one headset: head on a sphere looking at the origin,
    gaze in degrees from the optical axis."""
    hid: int
    scene: int
    seed: int
    radius: float = 5.0
    rate_hz: float = 90.0
    jitter_ms: float = 1.0
    saccade_ms: float = 300.0
    rng: np.random.Generator = field(init=False)

    def __post_init__(self):
        self.rng = np.random.default_rng(self.seed)
        self.az = self.rng.uniform(0.0, 2.0 * math.pi)
        self.el = self.rng.uniform(-0.4, 0.4)
        self.vaz = 0.0
        self.vel = 0.0
        self.gaze = np.zeros(2)            # degrees (x right, y down)
        self.next_saccade = self.rng.exponential(self.saccade_ms) * 1e-3

    def arrivals(self, duration_s: float) -> list:
        """Arrival times in [0, duration): one per frame period, jittered."""
        period = 1.0 / self.rate_hz
        phase = self.rng.uniform(0.0, period)
        out = []
        k = 0
        while True:
            t = phase + k * period + self.rng.normal(0.0, self.jitter_ms * 1e-3)
            if phase + k * period >= duration_s:
                break
            out.append(max(0.0, t))
            k += 1
        return sorted(out)

    def pose_at(self, t: float, dt: float) -> tuple:
        """Advances the walk by dt and returns (eye position, gaze degrees)."""
        # smooth head: angular velocity is itself a damped random walk
        self.vaz = 0.95 * self.vaz + self.rng.normal(0.0, 0.02) * math.sqrt(max(dt, 1e-4) * 90.0)
        self.vel = 0.95 * self.vel + self.rng.normal(0.0, 0.01) * math.sqrt(max(dt, 1e-4) * 90.0)
        self.az += self.vaz * dt
        self.el = float(np.clip(self.el + self.vel * dt, -0.6, 0.6))
        # gaze: fixational drift plus saccades of 10-40 degrees
        self.gaze += self.rng.normal(0.0, 0.5, size=2) * math.sqrt(max(dt, 1e-4) * 90.0) * 0.1
        if t >= self.next_saccade:
            amp = self.rng.uniform(10.0, 40.0)
            ang = self.rng.uniform(0.0, 2.0 * math.pi)
            self.gaze = self.gaze + amp * np.array([math.cos(ang), math.sin(ang)])
            self.next_saccade = t + self.rng.exponential(self.saccade_ms) * 1e-3
        # keep the gaze inside the central 30 degrees of a 100 degree display
        n = float(np.linalg.norm(self.gaze))
        if n > 30.0:
            self.gaze *= 30.0 / n
        eye = self.radius * np.array([math.cos(self.el) * math.sin(self.az), -math.sin(self.el),
                                      -math.cos(self.el) * math.cos(self.az)])
        return eye, self.gaze.copy()


REFERENCE_FOVEATION = (_lib.DEFAULT_FOVEATION_RADIUS_PX, _lib.DEFAULT_BLEND_WIDTH_PX, 0)


def stereo_views(eye: np.ndarray, gaze_deg: np.ndarray, width: int, height: int, fov_deg: float,
                 ipd: float = 0.063, device="cuda", foveation: tuple = REFERENCE_FOVEATION) -> tuple:
    """RasterizerSettings of both eyes for a head position looking at the
    origin, with the gaze projected into the left eye (as the stereo call
    uses the left eye's gaze). ``foveation``: (radius, blend width, quarter
    radius) in pixels; the culling tables follow the radius."""
    R = S.look_at_rotation(eye, np.zeros(3))
    left = S.make_camera(width, height, fov_deg, eye - 0.5 * ipd * R[0], R)
    right = S.make_camera(width, height, fov_deg, eye + 0.5 * ipd * R[0], R)
    gx = left.cx + left.fx * math.tan(math.radians(gaze_deg[0]))
    gy = left.cy + left.fy * math.tan(math.radians(gaze_deg[1]))
    gaze = np.array([np.clip(gx, 0, width - 1), np.clip(gy, 0, height - 1)], dtype=np.float32)
    fr, bw = int(foveation[0]), int(foveation[1])
    return (S.settings_for(left, gaze, device=device, foveation_radius_px=fr, blend_width_px=bw),
            S.settings_for(right, gaze, device=device, foveation_radius_px=fr, blend_width_px=bw))


# ---------------------------------------------------------------- requests

@dataclass
class Request:
    headset: int
    frame: int
    t_arrival: float
    settings: tuple              # (left, right) RasterizerSettings
    t_start: float = 0.0
    t_done: float = 0.0
    images: tuple | None = None
    dropped: bool = False


def make_requests(headsets: list, duration_s: float, width: int, height: int, fov_deg: float,
                  device="cuda", foveation: tuple = REFERENCE_FOVEATION) -> list:
    """All requests of a run, in arrival order, with their cameras prepared
    (table building is caller-side work, done before the clock starts)."""
    reqs = []
    for hs in headsets:
        last = 0.0
        for k, t in enumerate(hs.arrivals(duration_s)):
            eye, gaze = hs.pose_at(t, t - last)
            last = t
            reqs.append(Request(hs.hid, k, t, stereo_views(eye, gaze, width, height, fov_deg, device=device,
                                                           foveation=foveation)))
    reqs.sort(key=lambda r: r.t_arrival)
    return reqs


# ---------------------------------------------------------------- server

def slot_bytes(n_gaussians: int, width: int, height: int, foveation: tuple = REFERENCE_FOVEATION) -> int:
    """Device bytes one eye's slot holds, from the library's own sizing
    (fhtgs_workspace_query): the per-Gaussian and per-tile buffers of a view
    of this size with device-built culling tables, plus instance buffers for
    about 2.5 instances per Gaussian with the headroom a device-count frame
    adds (25 % + 65536)."""
    fr, bw, qr = (int(x) for x in foveation)
    view = _lib.View(width=int(width), height=int(height), K=8, foveation_radius_px=fr,
                     blend_width_px=bw, quarter_radius_px=qr)
    inst = int(2.5 * n_gaussians * 1.25) + 65536
    fixed, inst_bytes = _lib.workspace_query(view, n_gaussians, inst)
    return fixed + inst_bytes


class Server:
    """Continuous-batching renderer for admitted headsets (see module doc)."""

    def __init__(self, scenes: list, device: int = 0, workers: int = 2, max_batch_headsets: int = 8,
                 graph: bool = True, reserve_bytes: int = 2 << 30, foveation: tuple = REFERENCE_FOVEATION,
                 keep_images: bool = True, view_branches: int | None = None):
        self.scenes = scenes            # torch scene tuples, resident
        self.device = device
        self.foveation = tuple(int(x) for x in foveation)  # (radius, blend width, quarter radius) px
        self.keep_images = keep_images  # False: served images are dropped once rendered
        self.max_batch = max_batch_headsets
        self.reserve = reserve_bytes
        self.handles = [_lib.Handle(device) for _ in range(workers)]
        for h in self.handles:
            h.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH if graph else _lib.FRAME_EXACT)
            if view_branches is not None:
                h.set_option(_lib.OPT_VIEW_BRANCHES, int(view_branches))
        self.streams = [torch.cuda.Stream(device=device) for _ in range(workers)]
        self.admitted: dict = {}       # headset id -> scene index
        self.rejected: list = []
        self.lock = threading.Condition()
        self.pending: dict = {}        # headset id -> newest waiting Request
        self.done: list = []
        self.closed = False

    def admit(self, hid: int, scene: int, width: int, height: int) -> bool:
        """Admits a headset if the device can hold its slots (in every
        worker's handle) beside what is already allocated."""
        free, _ = torch.cuda.mem_get_info(self.device)
        need = 2 * len(self.handles) * slot_bytes(int(self.scenes[scene][0].shape[0]), width, height,
                                                  self.foveation)
        if free - need < self.reserve:
            self.rejected.append(hid)
            return False
        self.admitted[hid] = scene
        return True

    def submit(self, req: Request) -> None:
        with self.lock:
            old = self.pending.get(req.headset)
            if old is not None:
                old.dropped = True
                self.done.append(old)
            self.pending[req.headset] = req
            self.lock.notify()

    def close(self) -> None:
        with self.lock:
            self.closed = True
            self.lock.notify_all()

    def _take(self) -> list:
        with self.lock:
            while not self.pending and not self.closed:
                self.lock.wait()
            if not self.pending:
                return []
            hids = sorted(self.pending, key=lambda h: self.pending[h].t_arrival)[: self.max_batch]
            # headset order: the same set of headsets gives the same batch
            # shape, so its captured graph is replayed
            return [self.pending.pop(h) for h in sorted(hids)]

    def _render(self, w: int, batch: list, t0: float) -> None:
        handle, stream = self.handles[w], self.streams[w]
        keep: list = []
        scene_ids = sorted({self.admitted[r.headset] for r in batch})
        sidx = {s: k for k, s in enumerate(scene_ids)}
        structs, device = [], None
        for s in scene_ids:
            st, device = _binding.scene_struct(*self.scenes[s], keep=keep)
            structs.append(st)
        views, vscene, vslot, images = [], [], [], []
        with torch.cuda.stream(stream):
            for r in batch:
                pair = []
                for eye, st in enumerate(r.settings):
                    fr, bw, qr = self.foveation
                    v, img = _binding.view_struct(device, *st.as_tuple(), False, False, False, fr, bw,
                                                  False, keep=keep, quarter_radius_px=qr)
                    views.append(v)
                    vscene.append(sidx[self.admitted[r.headset]])
                    vslot.append(2 * r.headset + eye)
                    pair.append(img)
                images.append(tuple(pair))
            start = time.perf_counter() - t0
            handle.render_batch(structs, views, vscene, stream.cuda_stream, vslot)
        end = time.perf_counter() - t0
        for r, im in zip(batch, images):
            r.t_start, r.t_done, r.images = start, end, (im if self.keep_images else None)
        with self.lock:
            self.done.extend(batch)

    def _worker(self, w: int, t0: float) -> None:
        torch.cuda.set_device(self.device)
        while True:
            batch = self._take()
            if not batch:
                return
            self._render(w, batch, t0)

    def warm_up(self, requests: list, frames: int = 2) -> None:
        """Renders each headset's request `frames` times on every worker,
        untimed: the first frame sizes the headset's slots (an exact frame),
        the second captures the worker's graph for it, so that no timed
        request pays for either. With batches of two or more headsets every
        pair of headsets is rendered the same way too (a batch's graph
        depends on which headsets, and so which slots, it holds); larger
        batch shapes are captured when they first occur."""
        first = {r.headset: r for r in reversed(requests)}
        shapes = [(h,) for h in sorted(first)]
        if self.max_batch >= 2:
            hs = sorted(first)
            shapes += [(a, b) for i, a in enumerate(hs) for b in hs[i + 1:]]
        t0 = time.perf_counter()
        for w in range(len(self.handles)):
            for _ in range(frames):
                for shape in shapes:
                    self._render(w, [Request(h, -1, 0.0, first[h].settings) for h in shape], t0)
        torch.cuda.synchronize(self.device)
        self.done = []

    def reset(self, max_batch_headsets: int | None = None) -> None:
        """Makes the server ready for another run (same handles, slots and
        captured graphs), optionally with another batch limit."""
        with self.lock:
            self.pending = {}
            self.done = []
            self.closed = False
            if max_batch_headsets is not None:
                self.max_batch = int(max_batch_headsets)

    def run(self, requests: list) -> dict:
        """Plays the requests at their arrival times and returns statistics."""
        for r in requests:
            if r.headset not in self.admitted:
                raise ValueError(f"headset {r.headset} was not admitted")
        t0 = time.perf_counter()
        threads = [threading.Thread(target=self._worker, args=(w, t0), daemon=True) for w in range(len(self.handles))]
        try:
            for th in threads:
                th.start()
            for r in requests:
                delay = r.t_arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                self.submit(r)
        finally:
            self.close()
            for th in threads:
                th.join()
        torch.cuda.synchronize(self.device)
        return summarize(self.done, time.perf_counter() - t0)


def summarize(done: list, wall_s: float) -> dict:
    served = [r for r in done if not r.dropped]
    lat = np.array([(r.t_done - r.t_arrival) * 1e3 for r in served]) if served else np.zeros(1)
    return {
        "requests": len(done), "served": len(served), "dropped": len(done) - len(served),
        "frames_per_s": len(served) / wall_s if wall_s > 0 else 0.0,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()), "wall_s": wall_s,
    }


def run_plain(scenes: list, admitted: dict, requests: list, device: int = 0,
              foveation: tuple = REFERENCE_FOVEATION, keep_images: bool = True) -> dict:
    """The plain per-request path on the same arrivals: one worker, one
    stereo call per request (exact frames), requests served in arrival order
    with the same latest-frame-wins rule. One untimed frame per headset
    first sizes the buffers, as the server's warm-up does."""
    from .torch_bindings.rasterization import FoveatedHTGSRasterizer
    handle = _lib.handle_for(device)
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_EXACT)
    rast = FoveatedHTGSRasterizer(*foveation)
    for r in {r.headset: r for r in reversed(requests)}.values():
        rast.render_stereo(*scenes[admitted[r.headset]], *r.settings, False, False, False)
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    pending: dict = {}
    done: list = []
    i = 0
    while i < len(requests) or pending:
        now = time.perf_counter() - t0
        while i < len(requests) and requests[i].t_arrival <= now:
            r = requests[i]
            if r.headset in pending:
                pending[r.headset].dropped = True
                done.append(pending[r.headset])
            pending[r.headset] = r
            i += 1
        if not pending:
            time.sleep(max(0.0, requests[i].t_arrival - now))
            continue
        hid = min(pending, key=lambda h: pending[h].t_arrival)
        r = pending.pop(hid)
        r.t_start = time.perf_counter() - t0
        left, right = r.settings
        images = rast.render_stereo(*scenes[admitted[hid]], left, right, False, False, False)
        torch.cuda.synchronize(device)
        r.images = images if keep_images else None
        r.t_done = time.perf_counter() - t0
        done.append(r)
    handle.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
    return summarize(done, time.perf_counter() - t0)
