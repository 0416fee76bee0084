"""Seeded synthetic scenes and cameras for tests and benchmarks.

Distributions: means uniform in a cube of side L, log-scales N(-4.5, 0.6) per
axis (stored as linear sigma, as the kernels read them), uniform unit
quaternions (Shoemake) in (w,x,y,z) order, opacity sigmoid(N(0, 1.5)),
sh_0 ~ N(0.5, 0.2) and 15 higher SH coefficients per channel ~ N(0, 0.05).
Cameras: OpenCV axes (x right, y down, z forward) in a y-down world, placed
on a sphere and looking at the origin; VPM maps view depth [near, far] to
[-1, 1] as the preprocess' z-extent cull expects (kernel_utils.cuh:143-149).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def make_scene(n: int, seed: int = 0, cube: float = 10.0, sh_rest_bases: int = 15) -> dict:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-cube / 2, cube / 2, size=(n, 3)).astype(np.float32)
    scales = np.exp(rng.normal(-4.5, 0.6, size=(n, 3))).astype(np.float32)
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    rot = q.astype(np.float32)
    opa = (1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.5, size=n)))).astype(np.float32)
    sh0 = rng.normal(0.5, 0.2, size=(n, 1, 3)).astype(np.float32)
    shr = rng.normal(0.0, 0.05, size=(n, sh_rest_bases, 3)).astype(np.float32)
    return {"positions": pos, "scales": scales, "rotations": rot, "opacities": opa,
            "sh_0": sh0, "sh_rest": shr}


@dataclass
class Camera:
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    near: float
    far: float
    M: np.ndarray        # 4x4 world -> view
    VPM: np.ndarray      # 4x4 world -> screen
    VPR_inv: np.ndarray  # 4x4 screen -> world direction
    position: np.ndarray  # 3

    def gaze_center(self) -> np.ndarray:
        return np.array([self.cx, self.cy], dtype=np.float32)


def look_at_rotation(eye: np.ndarray, target: np.ndarray) -> np.ndarray:
    """Rows: camera x (right), y (down), z (forward) in a y-down world."""
    f = target - eye
    f = f / np.linalg.norm(f)
    down = np.array([0.0, 1.0, 0.0])
    x = np.cross(down, f)
    if np.linalg.norm(x) < 1e-6:
        x = np.cross(np.array([0.0, 0.0, 1.0]), f)
    x = x / np.linalg.norm(x)
    y = np.cross(f, x)
    return np.stack([x, y, f])


def make_camera(width: int, height: int, fov_x_deg: float, eye: np.ndarray, R: np.ndarray,
                near: float = 0.2, far: float = 1000.0) -> Camera:
    """Pinhole camera with square pixels and a centred principal point from
    the horizontal field of view; R: world -> camera rotation (rows)."""
    fx = 0.5 * width / math.tan(math.radians(fov_x_deg) / 2)
    return camera_from_intrinsics(width, height, fx, fx, width / 2.0, height / 2.0, eye, R, near, far)


def camera_from_intrinsics(width: int, height: int, fx: float, fy: float, cx: float, cy: float,
                           eye: np.ndarray, R: np.ndarray, near: float = 0.2, far: float = 1000.0) -> Camera:
    """Camera matrices from pinhole intrinsics, the camera centre and the
    world -> camera rotation R (rows: camera x right, y down, z forward)."""
    eye = np.asarray(eye, dtype=np.float64)
    R = np.asarray(R, dtype=np.float64)
    t = -R @ eye
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = t
    K = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1.0]])
    a = (far + near) / (far - near)
    b = -2.0 * far * near / (far - near)
    VPM = np.zeros((4, 4))
    VPM[0, :3] = fx * R[0] + cx * R[2]
    VPM[0, 3] = fx * t[0] + cx * t[2]
    VPM[1, :3] = fy * R[1] + cy * R[2]
    VPM[1, 3] = fy * t[1] + cy * t[2]
    VPM[2, :3] = a * R[2]
    VPM[2, 3] = a * t[2] + b
    VPM[3, :3] = R[2]
    VPM[3, 3] = t[2]
    VPR_inv = np.zeros((4, 4))
    VPR_inv[:3, :3] = R.T @ np.linalg.inv(K)
    VPR_inv[3, 3] = 1.0
    return Camera(width, height, fx, fy, cx, cy, near, far, M.astype(np.float32), VPM.astype(np.float32),
                  VPR_inv.astype(np.float32), eye.astype(np.float32))


def random_eye(rng: np.random.Generator, radius: float = 5.0) -> np.ndarray:
    v = rng.normal(size=3)
    return radius * v / np.linalg.norm(v)


def orbit_camera(width: int, height: int, fov_x_deg: float, seed: int, radius: float = 5.0,
                 near: float = 0.2, far: float = 1000.0) -> Camera:
    """A seeded camera on a sphere of ``radius`` looking at the origin."""
    rng = np.random.default_rng(seed)
    eye = random_eye(rng, radius)
    return make_camera(width, height, fov_x_deg, eye, look_at_rotation(eye, np.zeros(3)), near, far)


def stereo_cameras(width: int, height: int, fov_x_deg: float, seed: int, ipd: float = 0.063,
                   radius: float = 5.0) -> tuple[Camera, Camera]:
    """Two eyes +-ipd/2 along the head x axis, parallel optical axes."""
    rng = np.random.default_rng(seed)
    head = random_eye(rng, radius)
    R = look_at_rotation(head, np.zeros(3))
    left = make_camera(width, height, fov_x_deg, head - 0.5 * ipd * R[0], R)
    right = make_camera(width, height, fov_x_deg, head + 0.5 * ipd * R[0], R)
    return left, right


def seeded_gaze(cam: Camera, seed: int, max_deg: float = 15.0) -> np.ndarray:
    """A gaze pixel whose direction lies within ``max_deg`` of the optical axis."""
    rng = np.random.default_rng(seed)
    theta = math.radians(max_deg) * math.sqrt(rng.random())
    phi = 2 * math.pi * rng.random()
    r = math.tan(theta)
    return np.array([cam.cx + cam.fx * r * math.cos(phi), cam.cy + cam.fy * r * math.sin(phi)],
                    dtype=np.float32)


def settings_for(cam: Camera, gaze: np.ndarray, device="cuda", K: int = 8, active_sh_bases: int = 16,
                 foveation_radius_px: int = 375, blend_width_px: int = 125, visible=None,
                 background=None, scale_modifier: float = 1.0):
    """RasterizerSettings for a camera, with culling tables built for it."""
    import torch

    from .tables import build_tables
    from .torch_bindings.rasterization import BackgroundModel, RasterizerSettings

    vis, vsat, fsat = build_tables(cam.width, cam.height, device, visible, foveation_radius_px, blend_width_px)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)
    if background is None:
        bg_model, bg_data = BackgroundModel.NONE, torch.zeros(1, dtype=torch.float32, device=device)
    else:
        bg_model, bg_data = background[0], dev(background[1])
    return RasterizerSettings(
        M=dev(cam.M), VPM=dev(cam.VPM), VPR_inv=dev(cam.VPR_inv), cam_position=dev(cam.position),
        gaze_position=torch.from_numpy(np.asarray(gaze, dtype=np.float32).copy()),
        visibility_mask=vis, visibility_mask_area_table=vsat, fovea_mask_area_table=fsat,
        background_model_data=bg_data, background_model=bg_model, K=K, active_sh_bases=active_sh_bases,
        width=cam.width, height=cam.height, focal_x=float(cam.fx), focal_y=float(cam.fy),
        center_x=float(cam.cx), center_y=float(cam.cy), near_plane=float(cam.near), far_plane=float(cam.far),
        scale_modifier=float(scale_modifier))


def scene_to_torch(scene: dict, device="cuda") -> tuple:
    import torch
    return tuple(torch.from_numpy(scene[k]).to(device) for k in
                 ("positions", "scales", "rotations", "opacities", "sh_0", "sh_rest"))
