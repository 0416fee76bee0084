"""Pins the oracle's debug overlays (visualization.cu / visualization.cuh):
the gaze marker against its geometry (square, lattice disk, 3-pixel cross)
and the tile borders against a brute-force tile classification."""
import numpy as np
import pytest

import oracle
from foveatedhtgscudabackend_amd import tables as T

COL = np.array([0.25, 0.5, 0.75], np.float32)
FOVEA, BLENDED, PERIPH = (0.902, 0.624, 0.0), (0.337, 0.706, 0.914), (0.0, 0.620, 0.451)


def _marked(img, base):
    return np.any(img != base, axis=-1)


@pytest.mark.parametrize("size", [1, 2, 3, 5, 6, 7, 9, 12, 13, 21])
@pytest.mark.parametrize("vtype", [0, 1, 2])
def test_gaze_marker_geometry(oracle_lib, size, vtype):
    W, H = 64, 48
    base = np.zeros((H, W, 3), np.float32)
    got = oracle.visualize_gaze(base, (30.7, 20.2), COL, vtype, size, False)
    m = _marked(got, base)
    k = size // 2
    dy, dx = np.mgrid[-k:k + 1, -k:k + 1]
    if vtype == 1:
        want = np.ones_like(dx, bool)
    elif vtype == 0:
        want = dx * dx + dy * dy <= (size / 2.0) ** 2
    else:
        want = (np.abs(dx) <= 1) | (np.abs(dy) <= 1)
    exp = np.zeros((H, W), bool)
    exp[20 - k:20 + k + 1, 30 - k:30 + k + 1] = want
    np.testing.assert_array_equal(m, exp)
    assert np.all(got[m] == COL)


def test_gaze_marker_multiple_of_eight_drops_last_row_and_column(oracle_lib):
    # the reference covers offsets i - size/2 for i < 8 ceil(size/8): with
    # size a multiple of 8 the offset +size/2 is never reached
    base = np.zeros((64, 64, 3), np.float32)
    m = _marked(oracle.visualize_gaze(base, (32.0, 32.0), COL, 1, 8, False), base)
    ys, xs = np.nonzero(m)
    assert (xs.min(), xs.max(), ys.min(), ys.max()) == (28, 35, 28, 35)


def test_gaze_marker_clamps_gaze_and_clips_at_the_border(oracle_lib):
    base = np.zeros((3, 40, 50), np.float32)
    got = oracle.visualize_gaze(base, (-12.0, 99.0), COL, 1, 7, True)
    m = np.any(got != base, axis=0)
    ys, xs = np.nonzero(m)
    # gaze clamps to (0, 39); the square of half-width 3 is cut by the border
    assert (xs.min(), xs.max(), ys.min(), ys.max()) == (0, 3, 36, 39)


def _brute_classes(W, H, gaze, fr, bw, vis):
    gw, gh = T.grid_size(W, H)
    R = T.fovea_radius_tiles_for(W, H, fr, bw)
    B = R if fr < 0 else max(0, fr - bw) // 16
    gx = max(0, min(2 * gw - 1, (int(min(max(gaze[0], 0.0), W - 1)) + 8) // 16))
    gy = max(0, min(2 * gh - 1, (int(min(max(gaze[1], 0.0), H - 1)) + 8) // 16))
    ty, tx = np.mgrid[0:gh, 0:gw]
    d2 = (tx + 0.5 - gx) ** 2 + (ty + 0.5 - gy) ** 2
    split = vis & (d2 < R * R)
    return split & (d2 < B * B), split & ~(d2 < B * B), vis & ~split


@pytest.mark.parametrize("W,H,gaze,fr,bw,frac", [
    (320, 240, (150.0, 110.0), 120, 50, 0.0),
    (333, 257, (10.0, 250.0), 90, 40, 0.3),
    (256, 192, (128.0, 96.0), -1, 125, 0.2),
])
def test_tile_borders_follow_the_classification(oracle_lib, W, H, gaze, fr, bw, frac):
    gw, gh = T.grid_size(W, H)
    vis = np.random.default_rng(W).random((gh, gw)) >= frac if frac else np.ones((gh, gw), bool)
    base = np.full((H, W, 3), 0.5, np.float32)
    got = oracle.visualize_tile_boundaries(base, T.visibility_mask_np(W, H, vis), gaze, fr, bw, False)
    fovea, blended, periph = _brute_classes(W, H, gaze, fr, bw, vis)
    want = base.copy()
    yy, xx = np.mgrid[0:H, 0:W]
    lx, ly = xx // 16, yy // 16
    sub_edge = (xx % 8 == 0) | (xx % 8 == 7) | (yy % 8 == 0) | (yy % 8 == 7)
    big_edge = (xx % 16 == 0) | (xx % 16 == 15) | (yy % 16 == 0) | (yy % 16 == 15)
    want[fovea[ly, lx] & sub_edge] = FOVEA
    want[blended[ly, lx] & sub_edge] = BLENDED
    want[periph[ly, lx] & big_edge] = PERIPH
    np.testing.assert_array_equal(got, want.astype(np.float32))
