"""Host-side pieces of bench.py that run without a GPU: the CPU baseline leg
(the oracle rendering whole stereo frames) and the foveation radii."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def test_foveation_radii():
    import bench
    assert bench.foveation_radii(2, 2064, 100.0) == (375, 125, 0)
    fr, bw, qr = bench.foveation_radii(3, 2064, 100.0)
    assert (fr, bw, qr) == (232, 77, 606)


def test_cpu_baseline_renders_whole_stereo_frames(oracle_lib):
    import bench
    from foveatedhtgscudabackend_amd import scene as S
    sc = S.make_scene(5000, seed=3, cube=10.0)
    left, right = S.stereo_cameras(256, 272, 100.0, seed=7)
    gaze = S.seeded_gaze(left, 11)
    radii = bench.foveation_radii(3, 256, 100.0)
    cb = bench.cpu_baseline(sc, (left, right), gaze, radii, 8, 2)
    assert cb["frame_ms"] > 0.0 and cb["wall_s"] >= 2 * cb["frame_ms"] / 1e3 * 0.999
    for k in ("preprocess", "sort", "blend", "upsample"):
        assert k in cb["stage_ms"] and cb["stage_ms"][k] >= 0.0
    # per-frame stage times of both eyes cannot exceed the frame's wall time
    assert sum(cb["stage_ms"].values()) <= cb["frame_ms"] * 1.001
