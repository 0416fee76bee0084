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


def test_traffic_record_is_read_from_the_profile_and_checked_against_the_source(tmp_path, monkeypatch):
    import bench
    for levels in (3, 2):
        rec = bench.traffic_record(levels)
        assert rec is not None and rec["bytes"] > 1e8 and len(rec["blend_source"]) == 16
        assert rec["profile"].startswith("profiles/")
    # a record measured on another blend source is reported as stale: copy the
    # repository's profile with a different hash into a scratch root
    (tmp_path / "profiles").mkdir()
    src = open(os.path.join(ROOT, "profiles", "stereo_3M_2064x2208_3level_pmc_traffic.md")).read()
    rec = bench.traffic_record(3)
    (tmp_path / "profiles" / "stereo_3M_2064x2208_3level_pmc_traffic.md").write_text(
        src.replace(rec["blend_source"], "0" * 16))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.traffic_record(3)["blend_source"] == "0" * 16
    assert bench.traffic_record(2) is None
    monkeypatch.setattr(bench, "ROOT", ROOT)
    assert bench.blend_source_sha() != "0" * 16


def test_cpu_config_leg_times_the_plain_configuration(oracle_lib):
    import bench
    r = bench.cpu_config_ms("cfg1", 0, 1)
    assert r["ms_per_frame"] > 0 and r["runs"] == 1 and r["cores"] == 1
    assert "800x800" in r["workload"] and "K=4" in r["workload"]
    assert r["stage_ms"]["blend"] <= r["ms_per_frame"]
