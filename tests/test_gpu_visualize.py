"""GPU overlays against the oracle, bit for bit, through the reference's
operator surface (FoveatedHTGSVisualizer)."""
import numpy as np
import pytest
import torch

import oracle
from foveatedhtgscudabackend_amd import FoveatedHTGSVisualizer, GazeVisualizationType
from foveatedhtgscudabackend_amd import tables as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chw", [False, True])
@pytest.mark.parametrize("vtype", list(GazeVisualizationType))
@pytest.mark.parametrize("size,gaze", [(1, (5.0, 5.0)), (8, (100.3, 50.9)), (13, (0.0, 119.0)), (40, (250.0, -3.0))])
def test_gaze_marker_matches_oracle(chw, vtype, size, gaze):
    W, H = 160, 120
    rng = np.random.default_rng(size)
    shape = (3, H, W) if chw else (H, W, 3)
    base = rng.random(shape).astype(np.float32)
    col = np.array([0.1, 0.9, 0.4], np.float32)
    img = torch.from_numpy(base.copy()).cuda()
    out = FoveatedHTGSVisualizer().visualize_gaze(img, torch.tensor(gaze), torch.from_numpy(col), size, vtype, chw)
    torch.cuda.synchronize()
    want = oracle.visualize_gaze(base, gaze, col, vtype.value, size, chw)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("chw", [False, True])
@pytest.mark.parametrize("W,H,gaze,fr,bw,frac", [
    (640, 480, (300.0, 200.0), 375, 125, 0.0),
    (333, 257, (332.9, 0.0), 120, 50, 0.4),
    (2064, 2208, (1000.0, 1200.0), 375, 125, 0.1),
    (256, 192, (128.0, 96.0), -1, 125, 0.0),
])
def test_tile_borders_match_oracle(chw, W, H, gaze, fr, bw, frac):
    gw, gh = T.grid_size(W, H)
    vis = np.random.default_rng(W).random((gh, gw)) >= frac if frac else None
    words = T.visibility_mask_np(W, H, vis)
    shape = (3, H, W) if chw else (H, W, 3)
    base = np.random.default_rng(H).random(shape).astype(np.float32)
    img = torch.from_numpy(base.copy()).cuda()
    mask = T.build_tables(W, H, "cuda", vis, fr, bw)[0]
    out = FoveatedHTGSVisualizer(fr, bw).visualize_tile_boundaries(img, mask, torch.tensor(gaze), chw)
    torch.cuda.synchronize()
    want = oracle.visualize_tile_boundaries(base, words, gaze, fr, bw, chw)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
