"""Host logic of the serving loop (no GPU): the headset simulation's arrival
process, gaze bounds and saccades, and the per-slot memory estimate used for
admission."""
import math

import numpy as np

from foveatedhtgscudabackend_amd.serving import HeadsetSim, slot_bytes


def test_arrivals_follow_the_frame_rate_with_small_jitter():
    hs = HeadsetSim(0, 0, seed=5)
    t = np.array(hs.arrivals(2.0))
    assert abs(len(t) - 180) <= 1
    assert np.all(np.diff(t) >= 0) and t.min() >= 0.0
    period = np.diff(t)
    assert abs(period.mean() - 1.0 / 90.0) < 2e-4
    assert 0.5e-3 < period.std() < 2.5e-3      # two jittered ends, sigma 1 ms each


def test_arrivals_are_seeded():
    assert HeadsetSim(1, 0, seed=9).arrivals(0.5) == HeadsetSim(1, 0, seed=9).arrivals(0.5)
    assert HeadsetSim(1, 0, seed=9).arrivals(0.5) != HeadsetSim(1, 0, seed=10).arrivals(0.5)


def test_gaze_stays_central_and_saccades_happen():
    hs = HeadsetSim(2, 0, seed=3)
    t, last, jumps, prev = 0.0, 0.0, 0, np.zeros(2)
    for k in range(900):                       # 10 s at 90 Hz
        t = k / 90.0
        eye, g = hs.pose_at(t, t - last)
        last = t
        assert np.linalg.norm(g) <= 30.0 + 1e-9
        assert abs(np.linalg.norm(eye) - 5.0) < 1e-9
        if np.linalg.norm(g - prev) > 5.0:
            jumps += 1
        prev = g
    # Exp(300 ms) intervals: about 33 saccades in 10 s
    assert 15 <= jumps <= 60


def test_head_motion_is_smooth():
    hs = HeadsetSim(3, 0, seed=4)
    pts, last = [], 0.0
    for k in range(180):
        t = k / 90.0
        eye, _ = hs.pose_at(t, t - last)
        last = t
        pts.append(eye)
    step = np.linalg.norm(np.diff(np.array(pts), axis=0), axis=1)
    assert step.max() < 0.05                   # under 5 cm per frame


def test_slot_estimate_grows_with_scene_and_image():
    a = slot_bytes(1_000_000, 1024, 1024)
    assert slot_bytes(2_000_000, 1024, 1024) > a and slot_bytes(1_000_000, 2048, 2048) > a
    assert a > 80 * 1_000_000
