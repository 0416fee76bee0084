"""Pins the oracle's quarter-resolution level (an extension with no reference
counterpart) against brute force and the mathematics of the layout:

* classification: a visible tile is QUARTER iff its centre lies at least the
  quarter radius (in tiles) from the gaze tile; its partition follows the
  reference's three, stable in raster order;
* blend: QUARTER samples sit at stride 4 and, since a large tile's instance
  list is the same Gaussians in the same order as each of its sub-tiles'
  lists in a foveation-off render, equal that render bit for bit;
* fill: every other pixel of a QUARTER tile is the bilinear interpolation of
  the four surrounding stride-4 samples (fractions k/4), with samples in
  invisible tiles clamped into the tile and samples past the image edge
  clamped to the last one inside; PERIPHERY and BLENDED stride-2 fills treat
  QUARTER neighbours like invisible tiles;
* with no tile beyond the quarter radius the image is the two-level one.
"""
import numpy as np
import pytest

from foveatedhtgscudabackend_amd import scene as S
from foveatedhtgscudabackend_amd import tables as T
from tests.test_oracle_pins import _axis_camera, _scene, _tables


def _render(oracle_lib, sc, cam, gaze, fr, bw, qr, visible=None, **kw):
    tabs = _tables(cam.width, cam.height, fr, bw, visible)
    return oracle_lib.run(sc, oracle_lib.view_from_camera(
        cam, gaze, tabs, foveation_radius_px=fr, blend_width_px=bw, quarter_radius_px=qr, **kw))


def _tile_types(o, gw, gh):
    """Per large tile: -1 invisible, else the type of its first partitioned entry."""
    t = np.full(gw * gh, -1, np.int64)
    lt = o["tile_map_part"] // 4
    t[lt] = o["tile_type_part"]
    return t.reshape(gh, gw)


@pytest.mark.parametrize("w,h,gaze,fr,bw,qr,frac", [
    (640, 480, (300.0, 200.0), 150, 60, 250, 0.0),
    (1000, 700, (40.0, 650.0), 200, 80, 200, 0.2),
    (333, 257, (332.9, 0.0), 64, 32, 97, 0.4),
])
def test_quarter_classification_brute_force(oracle_lib, w, h, gaze, fr, bw, qr, frac):
    gw, gh = T.grid_size(w, h)
    vis = np.random.default_rng(w).random((gh, gw)) >= frac if frac else None
    o = _render(oracle_lib, _scene(np.zeros((0, 3)), [], []), _axis_camera(w, h), gaze, fr, bw, qr, vis)
    c = o["count"]
    R, B, Q = c["fovea_tiles"], c["blend_tiles"], c["quarter_tiles"]
    assert Q == -(-qr // 16)
    gx, gy = c["gaze_tx"], c["gaze_ty"]
    ty, tx = np.mgrid[0:gh, 0:gw]
    d2 = (tx + 0.5 - gx) ** 2 + (ty + 0.5 - gy) ** 2
    visible = np.ones((gh, gw), bool) if vis is None else vis
    split = visible & (d2 < R * R)
    fovea = split & (d2 < B * B)
    quarter = visible & (d2 >= Q * Q)
    periph = visible & ~split & ~quarter
    assert not (quarter & split).any()
    nF, nB, nP, nQ = 4 * int(fovea.sum()), 4 * int((split & ~fovea).sum()), int(periph.sum()), int(quarter.sum())
    assert nQ > 0 and (nP > 0) == (qr > fr)
    assert c["n_active"] == nF + nP + nB + nQ
    # tiles-touched counts are the two-level ones: a QUARTER tile has one list
    np.testing.assert_array_equal(o["tile_num"], np.where(split, 4, np.where(visible, 1, 0)).reshape(-1))
    types = o["tile_type_part"]
    want = np.concatenate([np.zeros(nF), np.ones(nP), np.full(nB, 2), np.full(nQ, 3)]).astype(np.uint8)
    np.testing.assert_array_equal(types, want)
    ends = np.cumsum([0, nF, nP, nB, nQ])
    part = o["partition"].reshape(4, 2)
    for t in range(4):
        a, b = ends[t], ends[t + 1]
        assert tuple(part[t]) == ((a, b) if b > a else (0, 0))
    lt = o["tile_map_part"] // 4
    np.testing.assert_array_equal(np.sort(lt[types == 3]), np.flatnonzero(quarter.reshape(-1)))
    for t in range(4):
        assert np.all(np.diff(lt[types == t].astype(np.int64)) >= 0)
    # the compacted raster-order map carries the same types
    tt = _tile_types(o, gw, gh)
    assert np.array_equal(tt == 3, quarter)


def _seeded(n, w, h, seed):
    sc = S.make_scene(n, seed=seed, cube=10.0)
    cam = S.orbit_camera(w, h, 60.0, seed=seed + 1000)
    gaze = S.seeded_gaze(cam, seed + 2000)
    return sc, cam, gaze


@pytest.mark.parametrize("w,h,seed,frac,bg", [(320, 240, 70, 0.0, 0), (333, 257, 71, 0.3, 2)])
def test_quarter_samples_equal_full_resolution(oracle_lib, w, h, seed, frac, bg):
    sc, cam, gaze = _seeded(6000, w, h, seed)
    gw, gh = T.grid_size(w, h)
    vis = np.random.default_rng(seed).random((gh, gw)) >= frac if frac else None
    kw = {}
    if bg:
        kw = dict(background_model_type=2,
                  background_model_data=np.random.default_rng(seed).random((3, 16, 32)).astype(np.float32))
    full = _render(oracle_lib, sc, cam, gaze, -1, 125, 0, vis, **kw)["image"].reshape(h, w, 3)
    o = _render(oracle_lib, sc, cam, gaze, 64, 32, 80, vis, **kw)
    img = o["image"].reshape(h, w, 3)
    tt = _tile_types(o, gw, gh)
    assert (tt == 3).sum() > 10
    ys, xs = np.mgrid[0:h, 0:w]
    tq = tt[ys // 16, xs // 16] == 3
    smp = tq & (xs % 4 == 0) & (ys % 4 == 0)
    assert smp.sum() > 100
    np.testing.assert_array_equal(img[smp].view(np.uint32), full[smp].view(np.uint32))
    # the same holds for the stride-2 samples of PERIPHERY tiles
    tp = tt[ys // 16, xs // 16] == 1
    smp2 = tp & (xs % 2 == 0) & (ys % 2 == 0)
    assert smp2.sum() > 100
    np.testing.assert_array_equal(img[smp2].view(np.uint32), full[smp2].view(np.uint32))


def _expected_fill(img, tt, lx, ly, S_, src_ok):
    """float64 bilinear fill of large tile (lx, ly) from its stride-S_ grid."""
    h, w = img.shape[:2]
    bx, by = 16 * lx, 16 * ly
    mx, my = (w - bx - 1) // S_, (h - by - 1) // S_

    def sample(sx, sy):
        px, py = bx + min(sx, mx) * S_, by + min(sy, my) * S_
        if not src_ok(tt[py // 16, px // 16]):
            px = max(bx, min(bx + 16 - S_, px))
            py = max(by, min(by + 16 - S_, py))
        return img[py, px].astype(np.float64)

    out = {}
    for ty in range(16):
        for tx in range(16):
            x, y = bx + tx, by + ty
            if x >= w or y >= h:
                continue
            xm, ym = tx % S_, ty % S_
            if xm == 0 and ym == 0:
                continue
            fx, fy = xm / S_, ym / S_
            i, j = tx // S_, ty // S_
            a, b, c, d = sample(i, j), sample(i + 1, j), sample(i, j + 1), sample(i + 1, j + 1)
            v = (1 - fx) * (1 - fy) * a + fx * (1 - fy) * b + (1 - fx) * fy * c + fx * fy * d
            out[(y, x)] = np.clip(v, 0.0, 1.0)
    return out


@pytest.mark.parametrize("w,h,seed,frac,chw", [(320, 240, 72, 0.0, False), (333, 257, 73, 0.3, True)])
def test_quarter_and_periphery_fill_is_bilinear(oracle_lib, w, h, seed, frac, chw):
    sc, cam, gaze = _seeded(5000, w, h, seed)
    gw, gh = T.grid_size(w, h)
    vis = np.random.default_rng(seed).random((gh, gw)) >= frac if frac else None
    o = _render(oracle_lib, sc, cam, gaze, 64, 32, 96, vis, to_chw=chw)
    img = o["image"].reshape((3, h, w) if chw else (h, w, 3))
    if chw:
        img = img.transpose(1, 2, 0)
    tt = _tile_types(o, gw, gh)
    checked = {1: 0, 3: 0}
    worst = 0.0
    for ly in range(gh):
        for lx in range(gw):
            t = tt[ly, lx]
            if t == 3:
                exp = _expected_fill(img, tt, lx, ly, 4, lambda k: k >= 0)
            elif t == 1:
                exp = _expected_fill(img, tt, lx, ly, 2, lambda k: k >= 0 and k != 3)
            else:
                continue
            for (y, x), v in exp.items():
                worst = max(worst, float(np.abs(img[y, x] - v).max()))
            checked[t] += 1
    assert checked[1] > 5 and checked[3] > 20
    assert worst <= 1e-6, worst


def test_quarter_level_with_no_quarter_tile_is_two_level(oracle_lib):
    sc, cam, gaze = _seeded(5000, 320, 240, 74)
    a = _render(oracle_lib, sc, cam, gaze, 100, 40, 0)
    b = _render(oracle_lib, sc, cam, gaze, 100, 40, 2000)  # beyond every tile centre
    assert b["partition"][6:].tolist() == [0, 0]
    np.testing.assert_array_equal(a["image"].view(np.uint32), b["image"].view(np.uint32))
    np.testing.assert_array_equal(a["tile_type"], b["tile_type"])


def test_constant_background_stays_constant_with_quarter_level(oracle_lib):
    w, h = 640, 480
    cam = _axis_camera(w, h, fov=90.0, forward=(0.0, -1.0, 0.0))
    gw, gh = T.grid_size(w, h)
    vis = np.random.default_rng(5).random((gh, gw)) >= 0.2
    tex = np.full((3, 8, 16), 0.5, np.float32)
    o = _render(oracle_lib, _scene(np.zeros((0, 3)), [], []), cam, (300.0, 260.0), 120, 50, 180, vis,
                background_model_type=2, background_model_data=tex)
    p = o["partition"].reshape(4, 2)
    assert (p[:, 1] > p[:, 0]).all(), "all four tile types present"
    img = o["image"].reshape(h, w, 3)
    px_vis = np.repeat(np.repeat(vis, 16, 0), 16, 1)[:h, :w]
    assert np.all(img[~px_vis] == 0.0)
    assert np.all(img[px_vis] == np.float32(0.5))


@pytest.mark.parametrize("fr,qr,blur", [(150, 100, False), (-1, 300, False), (150, 300, True)])
def test_quarter_radius_rejected(oracle_lib, fr, qr, blur):
    with pytest.raises(ValueError):
        _render(oracle_lib, _scene(np.zeros((0, 3)), [], []), _axis_camera(128, 96), (60.0, 40.0), fr, 30, qr,
                blur_periphery=blur)
