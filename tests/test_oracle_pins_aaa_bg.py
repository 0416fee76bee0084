"""Pins the oracle's anti-aliased preprocess (kernel_utils_aaa.cuh) and its
two background models (background_model.cuh) against closed forms, computed
here in float64 from the mathematics, not from the oracle's expressions."""
import math

import numpy as np
import pytest

from tests.test_oracle_math import _real_sh_textbook
from tests.test_oracle_pins import _axis_camera, _scene, _tables


def _run(oracle_lib, sc, cam, **kw):
    kw.setdefault("foveation_radius_px", -1)
    tabs = _tables(cam.width, cam.height, kw["foveation_radius_px"])
    return oracle_lib.run(sc, oracle_lib.view_from_camera(cam, (cam.cx, cam.cy), tabs, **kw))


# ---------------------------------------------------------------- AAA preprocess

@pytest.mark.parametrize("sigma,z,opac", [(0.01, 3.0, 0.9), (0.002, 6.0, 0.8), (0.05, 2.0, 0.6), (0.001, 4.0, 0.95)])
def test_aaa_opacity_dilation_closed_form(oracle_lib, sigma, z, opac):
    """An isotropic Gaussian of variance s^2 convolved with the pixel filter
    (variance 0.3 (z/f)^2 per axis) keeps its integral: the peak opacity
    scales by sqrt(det S / det S') = s^2 / (s^2 + 0.3 (z/f)^2)."""
    cam = _axis_camera(800, 800, fov=60.0)
    sc = _scene([[0.0, 0.0, z]], [sigma], [opac])
    o = _run(oracle_lib, sc, cam, anti_aliasing=True)
    f = max(cam.fx, cam.fy)
    var = sigma * sigma + 0.3 * (z / f) ** 2
    want = opac * sigma * sigma / var
    if want < 1.0 / 255.0:
        assert o["n_touched"][0] == 0
        return
    assert o["n_touched"][0] > 0
    got = float(o["rgba"].reshape(-1, 4)[0, 3])
    assert abs(got - want) <= 1e-5 * want


@pytest.mark.parametrize("sigma,z,opac", [(0.05, 3.0, 0.9), (0.2, 5.0, 0.5), (0.02, 1.5, 0.7)])
def test_aaa_extent_is_the_tangent_cone(oracle_lib, sigma, z, opac):
    """On the optical axis the screen extent of the dilated sphere of radius
    r = s' sqrt(rho') at distance z is f r / sqrt(z^2 - r^2) (the tangent
    cone); the anti-aliased bounds have no half-pixel shift."""
    w = h = 1024
    cam = _axis_camera(w, h, fov=70.0)
    sc = _scene([[0.0, 0.0, z]], [sigma], [opac])
    o = _run(oracle_lib, sc, cam, anti_aliasing=True)
    f = max(cam.fx, cam.fy)
    var = sigma * sigma + 0.3 * (z / f) ** 2
    op = opac * sigma * sigma / var
    rho = 2.0 * math.log(op * 255.0)
    r = math.sqrt(var) * math.sqrt(rho)
    e = cam.fx * r / math.sqrt(z * z - r * r)
    lo, hi = (cam.cx - e) / 16.0, (cam.cx + e) / 16.0
    assert min(abs(lo - round(lo)), abs(hi - round(hi))) > 1e-3, "ambiguous case"
    gw = w // 16
    want = [max(0, min(gw, math.floor(lo))), max(0, min(gw, math.ceil(hi)))]
    b = o["bounds"].reshape(-1, 4)[0]
    assert list(b[:2]) == want and list(b[2:]) == want


def test_aaa_culls_camera_inside_and_outside_frustum(oracle_lib):
    cam = _axis_camera(512, 512, fov=60.0)
    # 0: the camera lies inside the Gaussian's cutoff ellipsoid; 1: far to the
    # side of the frustum; 2: in view; 3: just outside the left edge but
    # reaching into the image (kept)
    z = 4.0
    half = z * math.tan(math.radians(30.0))
    sc = _scene([[0.0, 0.0, 0.3], [40.0, 0.0, z], [0.0, 0.0, z], [-half - 0.05, 0.0, z]],
                [0.5, 0.05, 0.05, 0.05], [0.9, 0.9, 0.9, 0.9])
    o = _run(oracle_lib, sc, cam, anti_aliasing=True)
    assert list(o["n_touched"] > 0) == [False, False, True, True]


# ---------------------------------------------------------------- backgrounds

def _pixel_dirs(cam):
    ys, xs = np.mgrid[0:cam.height, 0:cam.width].astype(np.float64)
    pc = np.stack([xs, ys, np.ones_like(xs), np.ones_like(xs)], axis=-1)
    t = pc @ cam.VPR_inv.astype(np.float64)[:3].T
    return t / np.linalg.norm(t, axis=-1, keepdims=True)


def _fade(t, col):
    ty = t[..., 1:2]
    col = np.where(ty > 0.0, col * (1.0 - ty / 0.1), col)
    return np.where(ty > 0.1, 0.0, col)


def test_sh_background_closed_form(oracle_lib):
    """Empty scene, no foveation: every pixel is the SH background
    0.5 tanh(0.5 + sum_lm Y_lm(d) k_lm) + 0.5 along its ray d, faded to black
    over 0 < d.y <= 0.1 (y down)."""
    cam = _axis_camera(96, 80, fov=90.0, forward=(0.3, 0.25, 1.0))
    k = np.random.default_rng(11).normal(0.0, 0.4, size=(16, 3))
    sc = _scene(np.zeros((0, 3)), [], [])
    o = _run(oracle_lib, sc, cam, background_model_type=1, background_model_data=k.astype(np.float32))
    t = _pixel_dirs(cam)
    Y = _real_sh_textbook(t.reshape(-1, 3)).reshape(t.shape[0], t.shape[1], 16)
    want = 0.5 * np.tanh(0.5 + Y @ k.astype(np.float32).astype(np.float64)) + 0.5
    want = _fade(t, np.clip(want, 0.0, 1.0))
    got = o["image"].reshape(cam.height, cam.width, 3)
    assert np.any(t[..., 1] > 0.1) and np.any(t[..., 1] < 0.0)  # the horizon is in view
    assert np.abs(got - want).max() <= 2e-5


@pytest.mark.parametrize("raw", [False, True])
def test_texture_background_closed_form(oracle_lib, raw):
    """Empty scene: equirectangular lookup u = atan2(d.x, d.z)/(2 pi) + 1/2,
    v = asin(d.y)/pi + 1/2, bilinear with horizontal wrap and vertical clamp,
    a sigmoid for the raw texture, then the horizon fade."""
    cam = _axis_camera(96, 80, fov=100.0, forward=(-0.6, 0.2, -1.0))
    rng = np.random.default_rng(12 + raw)
    H, W = 16, 32
    tex = rng.random((3, H, W)).astype(np.float32)
    if raw:
        tex = ((tex - 0.5) * 6.0).astype(np.float32)
    sc = _scene(np.zeros((0, 3)), [], [])
    o = _run(oracle_lib, sc, cam, background_model_type=3 if raw else 2, background_model_data=tex)
    t = _pixel_dirs(cam)
    u = np.arctan2(t[..., 0], t[..., 2]) / math.pi * 0.5 + 0.5
    v = np.arcsin(t[..., 1]) / (0.5 * math.pi) * 0.5 + 0.5
    tx, ty = u * W - 0.5, v * H - 0.5
    x0, y0 = np.floor(tx).astype(int), np.floor(ty).astype(int)
    fx, fy = tx - x0, ty - y0
    xa, xb = x0 % W, (x0 + 1) % W
    ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
    d = tex.astype(np.float64)
    c0 = d[:, ya, xa] * (1 - fx) + d[:, ya, xb] * fx
    c1 = d[:, yb, xa] * (1 - fx) + d[:, yb, xb] * fx
    val = c0 * (1 - fy) + c1 * fy
    if raw:
        val = 1.0 / (1.0 + np.exp(-val))
    want = np.clip(_fade(t, np.moveaxis(val, 0, -1)), 0.0, 1.0)
    got = o["image"].reshape(cam.height, cam.width, 3)
    assert np.abs(got - want).max() <= 5e-5
