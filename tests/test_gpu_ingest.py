"""A scene and camera read from 3DGS files render on the GPU bit-identically
to the oracle on the same arrays."""
import json

import numpy as np
import pytest
import torch

from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
from foveatedhtgscudabackend_amd import ingest as I
from foveatedhtgscudabackend_amd import scene as S
from foveatedhtgscudabackend_amd import tables as T

pytestmark = pytest.mark.gpu


def test_ply_scene_renders_like_the_oracle(tmp_path, oracle_lib):
    ply, cj = tmp_path / "point_cloud.ply", tmp_path / "cameras.json"
    I.write_ply(str(ply), S.make_scene(30000, seed=91, cube=8.0))
    cj.write_text(json.dumps([I.camera_to_json(S.orbit_camera(512, 384, 75.0, seed=1091))]))
    sc = I.read_ply(str(ply))
    cam = I.read_cameras_json(str(cj))[0]
    gaze = S.seeded_gaze(cam, 2091)
    img = FoveatedHTGSRasterizer().render(*S.scene_to_torch(sc, "cuda"), S.settings_for(cam, gaze), False, False, False)
    torch.cuda.synchronize()
    words = T.visibility_mask_np(cam.width, cam.height)
    tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height), T.fovea_area_table_np(24))
    want = oracle_lib.run(sc, oracle_lib.view_from_camera(cam, gaze, tabs))["image"].reshape(cam.height, cam.width, 3)
    np.testing.assert_array_equal(img.cpu().numpy().view(np.uint32), want.view(np.uint32))
