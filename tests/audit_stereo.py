"""Audits the blend's culls on the stereo configuration (3M Gaussians, 20 m
cube, 2 x 2064x2208 at 100 deg, K=8, the bench's foveation at 2 or 3 levels):
renders each eye's frame with the oracle, dumps it (tests/test_cull.py
format) and runs tests/cull_audit.cpp on it, then with --waves for the
blend's wave-level loop counts. Not collected by pytest (minutes of CPU).

    python tests/audit_stereo.py [levels=3] [eyes=1]
"""
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from foveatedhtgscudabackend_amd import scene as S  # noqa: E402
from foveatedhtgscudabackend_amd import tables as T  # noqa: E402
from tests.test_cull import dump_frame  # noqa: E402


def main(levels: int = 3, eyes: int = 1) -> int:
    import bench
    w, h, fov = 2064, 2208, 100.0
    fr, bw, qr = bench.foveation_radii(levels, w, fov)
    sc = S.make_scene(3_000_000, seed=0, cube=20.0)
    left, right = S.stereo_cameras(w, h, fov, seed=1000)
    gaze = S.seeded_gaze(left, 2000)
    tmp = tempfile.mkdtemp()
    exe = os.path.join(tmp, "cull_audit")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "cull_audit.cpp"), "-o", exe], check=True)
    rc = 0
    for eye, cam in (("left", left), ("right", right))[:eyes]:
        words = T.visibility_mask_np(w, h)
        tabs = (words, T.visibility_area_table_np(words, w, h),
                T.fovea_area_table_np(T.fovea_radius_tiles_for(w, h, fr, bw)))
        v = oracle.view_from_camera(cam, gaze, tabs, K=8, foveation_radius_px=fr, blend_width_px=bw,
                                    quarter_radius_px=qr)
        v["blend_tile_stride"] = 1 << 30  # the audit needs the per-Gaussian stages only
        t = time.time()
        o = oracle.run(sc, v)
        path = os.path.join(tmp, f"frame_{eye}.bin")
        dump_frame(path, o, w, h)
        print(f"{eye}: oracle frame in {time.time() - t:.1f} s", flush=True)
        for extra in ([], ["--waves"]):
            r = subprocess.run([exe, path] + extra, capture_output=True, text=True)
            print(r.stdout, end="", flush=True)
            rc |= r.returncode
        os.remove(path)
    return rc


if __name__ == "__main__":
    args = [int(a) for a in sys.argv[1:]]
    sys.exit(main(*args))
