"""Culling tables built on the GPU equal the host builder's, and a frame
whose view brings no tables (built inside the frame, also under graph replay)
renders the same image as one with host-built tables."""
import numpy as np
import pytest
import torch

from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer, _lib
from foveatedhtgscudabackend_amd import scene as S
from foveatedhtgscudabackend_amd import tables as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,fr,bw,frac", [(640, 480, 375, 125, 0.0), (333, 257, 120, 50, 0.3),
                                             (2064, 2208, 375, 125, 0.1), (256, 192, -1, 125, 0.0)])
def test_device_tables_equal_host_tables(w, h, fr, bw, frac):
    gw, gh = T.grid_size(w, h)
    vis = np.random.default_rng(w).random((gh, gw)) >= frac if frac else None
    want = [t.cpu().numpy() for t in T.build_tables(w, h, "cuda", vis, fr, bw)]
    vt = torch.from_numpy(vis).cuda() if vis is not None else None
    got = [t.cpu().numpy() for t in T.build_tables_device(w, h, vt, fr, bw)]
    torch.cuda.synchronize()
    for g, e in zip(got, want):
        np.testing.assert_array_equal(g.view(np.uint32), e.view(np.uint32))


def test_frames_without_tables_render_the_same():
    h = _lib.handle_for(0)
    sc = S.scene_to_torch(S.make_scene(30000, seed=95, cube=10.0), "cuda")
    r = FoveatedHTGSRasterizer()
    cams = [S.orbit_camera(480, 352, 80.0, seed=1095 + k) for k in range(2)]
    gw, gh = T.grid_size(480, 352)
    vis = np.random.default_rng(3).random((gh, gw)) >= 0.2
    for mode in (_lib.FRAME_EXACT, _lib.FRAME_GRAPH):
        h.set_option(_lib.OPT_FRAME_MODE, mode)
        for rep in range(2):
            for k, cam in enumerate(cams):
                gaze = S.seeded_gaze(cam, 2095 + k + rep)
                st = S.settings_for(cam, gaze, visible=vis if k else None)
                ref = r.render(*sc, st, False, False, False).cpu().numpy()
                mask = st.visibility_mask if k else None
                auto = st._replace(visibility_mask=mask, visibility_mask_area_table=None, fovea_mask_area_table=None)
                got = r.render(*sc, auto, False, False, False).cpu().numpy()
                np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                              err_msg=f"mode {mode} rep {rep} camera {k}")
    h.set_option(_lib.OPT_FRAME_MODE, _lib.FRAME_GRAPH)
