"""The blend's work tiles (FHTGS_OPT_BLEND_WORK_TILE: 8x8, 16x8 or 16x16
pixels rendered by one workgroup from one staged list) must not change a bit
of any image: every pixel sees the Gaussians of its tile's list that can
reach it, in list order, whichever workgroup renders it. Every case of the
parity suite (bit-identical to the oracle with 8x8 work tiles) is rendered
with the wider work tiles in every frame mode, and so is the stereo
configuration the frame rate is measured on; the wider tiles' work order must
list every unit once, longest first."""
import numpy as np
import pytest
import torch

from tests.test_gpu_parity import CASES, _gpu_render

pytestmark = pytest.mark.gpu

NAMES = {1: "8x8", 2: "16x8", 4: "16x16"}


@pytest.fixture
def handle():
    from foveatedhtgscudabackend_amd import _lib
    h = _lib.handle_for(0)
    yield _lib, h
    h.set_option(_lib.OPT_BLEND_WORK_TILE, 1)
    h.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)


def _same(a, b, what):
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=what)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_work_tiles_leave_images_unchanged(case, handle):
    lib, h = handle
    sc, cam, gaze, visible, tables, bg = case.build()
    h.set_option(lib.OPT_BLEND_WORK_TILE, 1)
    h.set_option(lib.OPT_FRAME_MODE, lib.FRAME_EXACT)
    ref = _gpu_render(case, sc, cam, gaze, visible, bg)
    for wt in (2, 4):
        h.set_option(lib.OPT_BLEND_WORK_TILE, wt)
        assert h.get_option(lib.OPT_BLEND_WORK_TILE) == wt
        assert h.tuned_work_tile(case.width, case.height) == wt
        for mode in (lib.FRAME_EXACT, lib.FRAME_DEVICE, lib.FRAME_GRAPH):
            h.set_option(lib.OPT_FRAME_MODE, mode)
            # graph mode: captured on its first device-count frame, replayed after
            for rep in range(3 if mode == lib.FRAME_GRAPH else 1):
                got = _gpu_render(case, sc, cam, gaze, visible, bg)
                _same(got, ref, f"{case.name}: work tile {NAMES[wt]}, frame mode {mode}, rep {rep}")


def test_work_tile_option_is_validated(handle):
    lib, h = handle
    for bad in (0, 3, 8, -1):
        with pytest.raises(ValueError):
            h.set_option(lib.OPT_BLEND_WORK_TILE, bad)
    assert h.get_option(lib.OPT_BLEND_WORK_TILE) == 1


def _wide_units(h, n_small, W):
    """(order, cost per unit) of the last frame's view 0 with W waves per
    blend workgroup: FOVEA work tiles (W sub-tiles, one list), PERIPHERY tiles
    W at a time, BLENDED work tiles, QUARTER tiles 4 W at a time, in partition
    order; a unit's cost is its longest instance list."""
    n_active = h.counts(0)[3]
    part = np.frombuffer(h.debug_read(0, "partition", 24), dtype=np.uint32).astype(np.int64)
    quart = np.frombuffer(h.debug_read(0, "partition_quarter", 8), dtype=np.uint32).astype(np.int64)
    tmp = np.frombuffer(h.debug_read(0, "tile_map_part", 4 * max(n_active, 1)), dtype=np.uint32)
    rng = np.frombuffer(h.debug_read(0, "ranges", 8 * n_small), dtype=np.uint32).reshape(-1, 2).astype(np.int64)
    length = (rng[:, 1] - rng[:, 0])[tmp[:n_active]] if n_active else np.zeros(0, np.int64)
    cost = []
    for (a, b), step in (((part[0], part[1]), W), ((part[2], part[3]), W), ((part[4], part[5]), W),
                         ((quart[0], quart[1]), 4 * W)):
        shared = step == W and (a, b) != (part[2], part[3])
        for u0 in range(int(a), int(b), step):
            lens = length[u0:min(int(b), u0 + step)]
            if shared:
                assert np.all(lens == lens[0]), "a work tile's sub-tiles share one list length"
            cost.append(int(lens.max()))
    order = np.frombuffer(h.debug_read(0, "blend_order", 4 * len(cost)), dtype=np.uint32)
    return order, np.array(cost, np.int64)


@pytest.mark.parametrize("w,hh,fr,bw,qr,n", [(800, 600, 120, 40, 250, 60000),
                                             (4096, 2304, -1, 125, 0, 20000)])
def test_wide_work_order_is_a_longest_first_permutation(handle, w, hh, fr, bw, qr, n):
    """With 16x8 and 16x16 work tiles every unit appears once in the work
    order and none comes after one with less than half its cost, in every
    frame mode (the 4096x2304 frame without foveation has 36864 16x16 work
    tiles)."""
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import scene as S
    lib, h = handle
    sc = S.make_scene(n, seed=66, cube=10.0)
    args = S.scene_to_torch(sc, "cuda")
    cam = S.orbit_camera(w, hh, 90.0, seed=1066)
    st = S.settings_for(cam, np.array((0.4 * w, 0.4 * hh), np.float32), foveation_radius_px=fr, blend_width_px=bw)
    gw, gh = -(-w // 16), -(-hh // 16)
    r = FoveatedHTGSRasterizer(fr, bw, qr)
    for W in (2, 4):
        h.set_option(lib.OPT_BLEND_WORK_TILE, W)
        for mode in (lib.FRAME_EXACT, lib.FRAME_DEVICE, lib.FRAME_GRAPH):
            h.set_option(lib.OPT_FRAME_MODE, mode)
            for _ in range(2):
                r.render(*args, st, False, False, False)
            torch.cuda.synchronize()
            order, cost = _wide_units(h, 4 * gw * gh, W)
            if fr < 0:
                assert len(cost) == gw * gh * 4 // W
            np.testing.assert_array_equal(np.sort(order), np.arange(len(cost)), err_msg=f"W {W} mode {mode}")
            v = cost[order] + 1
            assert np.all(v[1:] < 2 * np.minimum.accumulate(v)[:-1]), f"W {W} mode {mode}: not longest first"


def test_work_tiles_on_the_stereo_configuration(handle):
    """3M Gaussians, two 2064x2208 eyes at 100 deg, K=8, three levels,
    replayed from its captured graph: both eyes bit-identical with each work
    tile."""
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import scene as S
    from bench import foveation_radii
    lib, h = handle
    w, h_px, fov = 2064, 2208, 100.0
    fr, bw, qr = foveation_radii(3, w, fov)
    sc = S.make_scene(3_000_000, seed=0, cube=20.0)
    left, right = S.stereo_cameras(w, h_px, fov, seed=1000)
    gaze = S.seeded_gaze(left, 2000)
    sl = S.settings_for(left, gaze, K=8, foveation_radius_px=fr, blend_width_px=bw)
    sr = S.settings_for(right, gaze, K=8, foveation_radius_px=fr, blend_width_px=bw)
    args = S.scene_to_torch(sc, "cuda")
    r = FoveatedHTGSRasterizer(fr, bw, qr)
    h.set_option(lib.OPT_FRAME_MODE, lib.FRAME_GRAPH)
    ref = None
    for W in (1, 2, 4):
        h.set_option(lib.OPT_BLEND_WORK_TILE, W)
        for rep in range(3):  # exact, then captured, then replayed
            images = r.render_stereo(*args, sl, sr, False, False, False)
            torch.cuda.synchronize()
            got = [im.cpu().numpy() for im in images]
            if ref is None:
                ref = got
                continue
            for eye in range(2):
                _same(got[eye], ref[eye], f"work tile {NAMES[W]}, rep {rep}, eye {eye}")
