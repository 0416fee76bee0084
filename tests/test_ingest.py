"""3DGS PLY and cameras.json ingestion (CPU): conversions to the rasterizer's
activated layout, format variants, round trips and camera conventions."""
import json
import math

import numpy as np
import pytest

from foveatedhtgscudabackend_amd import ingest as I
from foveatedhtgscudabackend_amd import scene as S


def test_binary_round_trip(tmp_path):
    sc = S.make_scene(2000, seed=5)
    p = tmp_path / "s.ply"
    I.write_ply(str(p), sc)
    got = I.read_ply(str(p))
    np.testing.assert_array_equal(got["positions"], sc["positions"])
    np.testing.assert_array_equal(got["sh_0"], sc["sh_0"])
    np.testing.assert_array_equal(got["sh_rest"], sc["sh_rest"])
    np.testing.assert_allclose(got["scales"], sc["scales"], rtol=2e-6)
    np.testing.assert_allclose(got["opacities"], sc["opacities"], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(got["rotations"], sc["rotations"], atol=2e-7)


def test_ascii_degree1_with_extra_properties(tmp_path):
    # one Gaussian: logit opacity 0 -> 0.5, log scales -> linear, an unnormalised
    # quaternion, SH degree 1 stored channel-major (R coefficients, then G, then B)
    props = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{j}" for j in range(9)] + \
            ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "red"]
    vals = [1, 2, 3, 0, 0, 1, 0.1, 0.2, 0.3] + [10, 11, 12, 20, 21, 22, 30, 31, 32] + \
           [0.0, math.log(0.5), math.log(0.25), 0.0, 2, 0, 0, 0, 255]
    lines = ["ply", "format ascii 1.0", "comment made by a test", "element vertex 1"]
    lines += [f"property {'uchar' if p == 'red' else 'float'} {p}" for p in props]
    lines += ["element face 0", "property list uchar int vertex_indices", "end_header",
              " ".join(str(v) for v in vals)]
    p = tmp_path / "a.ply"
    p.write_text("\n".join(lines) + "\n")
    g = I.read_ply(str(p))
    np.testing.assert_array_equal(g["positions"], [[1, 2, 3]])
    np.testing.assert_allclose(g["scales"], [[0.5, 0.25, 1.0]], rtol=1e-7)
    np.testing.assert_allclose(g["opacities"], [0.5])
    np.testing.assert_array_equal(g["rotations"], [[1, 0, 0, 0]])
    assert g["sh_rest"].shape == (1, 3, 3)
    np.testing.assert_array_equal(g["sh_rest"][0], [[10, 20, 30], [11, 21, 31], [12, 22, 32]])


def test_big_endian_and_degree0(tmp_path):
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
             "rot_0", "rot_1", "rot_2", "rot_3"]
    data = np.zeros(3, dtype=[(n, ">f4") for n in names])
    data["x"] = [1, 2, 3]
    data["rot_1"] = [1, 1, 1]
    data["opacity"] = [-100, 0, 100]
    p = tmp_path / "b.ply"
    with open(p, "wb") as f:
        f.write(("ply\nformat binary_big_endian 1.0\nelement vertex 3\n" +
                 "".join(f"property float {n}\n" for n in names) + "end_header\n").encode())
        data.tofile(f)
    g = I.read_ply(str(p))
    assert g["sh_rest"].shape == (3, 0, 3)
    np.testing.assert_array_equal(g["positions"][:, 0], [1, 2, 3])
    np.testing.assert_array_equal(g["rotations"], np.tile([0, 1, 0, 0], (3, 1)))
    np.testing.assert_allclose(g["opacities"], [0.0, 0.5, 1.0], atol=1e-30)


def test_rejects_incomplete_files(tmp_path):
    p = tmp_path / "c.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n0\n")
    with pytest.raises(ValueError, match="lacks"):
        I.read_ply(str(p))
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
             "rot_0", "rot_1", "rot_2", "rot_3", "f_rest_0", "f_rest_1", "f_rest_2", "f_rest_3", "f_rest_4", "f_rest_5"]
    p.write_text("ply\nformat ascii 1.0\nelement vertex 1\n" + "".join(f"property float {n}\n" for n in names)
                 + "end_header\n" + " ".join(["0"] * len(names)) + "\n")
    with pytest.raises(ValueError, match="SH bases"):
        I.read_ply(str(p))


def test_cameras_json_round_trip_and_convention(tmp_path):
    cams = [S.orbit_camera(640, 480, 70.0, seed=k) for k in range(3)]
    p = tmp_path / "cameras.json"
    p.write_text(json.dumps([I.camera_to_json(c, k) for k, c in enumerate(cams)]))
    back = I.read_cameras_json(str(p))
    for a, b in zip(cams, back):
        np.testing.assert_allclose(b.VPM, a.VPM, rtol=1e-6, atol=1e-3)
        np.testing.assert_allclose(b.M, a.M, atol=1e-6)
        assert abs(I.fov_x_deg(b) - 70.0) < 1e-9
    # the rotation's columns are the camera axes in world space: a point one
    # metre along the third column projects to the principal point
    rot = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=np.float64)  # columns: camera x, y, z; z = world +x
    cam = I.camera_from_json({"width": 200, "height": 100, "fx": 150, "fy": 150,
                              "position": [1.0, 2.0, 3.0], "rotation": rot.tolist()})
    pt = np.array([1.0, 2.0, 3.0]) + rot[:, 2] * 4.0
    h = cam.VPM.astype(np.float64) @ np.append(pt, 1.0)
    assert abs(h[0] / h[3] - 100.0) < 1e-4 and abs(h[1] / h[3] - 50.0) < 1e-4 and abs(h[3] - 4.0) < 1e-5
