"""The blend's sample masks (fh_mask_prep / fh_mask_row_bits in
csrc/fhtgs_cull.h, row intervals of the ellipse record fh_ellipse_record that
the preprocess stores per Gaussian) are its tile cull -- a listed Gaussian
whose mask holds no sample of the tile is dropped -- and its first pass -- a
pixel evaluates only the staged Gaussians whose masks hold its bit. They must
not change the image: every sample a Gaussian reaches must be in its mask.
tests/cull_audit.cpp checks that over every (tile, instance) pair and sample
of oracle frames, with the blend's own float32 operations, and reports how
tight the masks are."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.cases import Case

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    Case("cull_fov", 40000, 640, 480, seed=51),
    Case("cull_wide_fov", 30000, 640, 480, seed=52, fov_deg=110.0, cube=6.0),
    Case("cull_close", 20000, 480, 352, seed=53, cube=3.0),
    Case("cull_scaled", 20000, 320, 240, seed=54, scale_modifier=3.0),
    Case("cull_aaa", 20000, 320, 240, seed=55, aaa=True),
    Case("cull_nofov", 20000, 320, 240, seed=56, foveation_radius_px=-1),
    Case("cull_quarter", 20000, 480, 352, seed=57, foveation_radius_px=100, blend_width_px=40,
         quarter_radius_px=140),
]


@pytest.fixture(scope="module")
def audit_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("cull") / "cull_audit"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(HERE, "cull_audit.cpp"), "-o", str(exe)], check=True)
    return exe


def dump_frame(path, o, width, height):
    c = o["count"]
    hdr = np.array([c["N"], c["instances"], c["n_small"], c["n_active"], c["gw"], width, height], dtype=np.int64)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.asarray(o["partition"], dtype=np.uint32).tobytes())
        f.write(np.asarray(o["tile_map_part"], dtype=np.uint32).tobytes())
        f.write(np.asarray(o["ranges"], dtype=np.uint32).tobytes())
        f.write(np.asarray(o["vals_sorted"], dtype=np.uint32).tobytes())
        for k in ("vpmt1", "vpmt2", "vpmt4", "rgba"):
            f.write(np.asarray(o[k], dtype=np.float32).tobytes())


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_cull_drops_only_unreachable(case, oracle_lib, audit_exe, tmp_path):
    sc, cam, gaze, visible, tables, bg = case.build()
    view = case.oracle_view(cam, gaze, tables, bg)
    view["blend_tile_stride"] = 1 << 30  # the audit needs the per-Gaussian stages only
    o = oracle_lib.run(sc, view)
    dump = tmp_path / "frame.bin"
    dump_frame(dump, o, case.width, case.height)
    r = subprocess.run([str(audit_exe), str(dump)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    m = re.search(r"pairs (\d+) kept (\d+) reaching (\d+) violations (\d+)", r.stdout)
    assert m, r.stdout + r.stderr
    pairs, kept, reach, bad = map(int, m.groups())
    assert bad == 0 and r.returncode == 0, r.stdout
    assert pairs == o["count"]["instances"] or pairs <= o["count"]["instances"]
    assert reach <= kept <= pairs
    assert pairs > 0
    # the blend's per-pixel pre-test keeps every sample a kept Gaussian reaches
    m = re.search(r"pixels (\d+) reached (\d+) pretest (\d+) pixel_violations (\d+) no_conic (\d+) min_slack (\S+)",
                  r.stdout)
    assert m, r.stdout
    evals, reached, pre, pbad = map(int, m.groups()[:4])
    assert pbad == 0, r.stdout
    assert reached <= pre <= evals and reached > 0
    assert float(m.group(6)) > 0.0
    # the 16x8 and 16x16 work tiles' masks (one 16-column interval per row for
    # two sub-tiles) keep every FOVEA / BLENDED sample a Gaussian reaches
    m = re.search(r"work_tiles reached (\d+) violations (\d+)", r.stdout)
    assert m, r.stdout
    assert int(m.group(2)) == 0, r.stdout
    if case.foveation_radius_px != -1:
        assert int(m.group(1)) > 0, "the case has FOVEA or BLENDED tiles"
