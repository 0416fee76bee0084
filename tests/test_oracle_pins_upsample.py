"""Pins the oracle's blur weights and blend-ring cross-fade against what the
mathematics fixes, on caller images through the oracle's upsampling stage
alone (oracle.upsample):

* blur (interpolation.cuh:270-413): the fused interpolate_and_blur equals a 2x
  bilinear upsample of the stride-2 samples followed by a separable
  [1 2 1]/4 binomial blur -- evaluated here in float64 in that order, on
  random images with invisible tiles (their samples clamped into the tile,
  interpolation.cuh:301-332) and the image edges;
* its impulse response is outer(r, r) with r = [1/8, 1/2, 3/4, 1/2, 1/8];
* blend ring (interpolation.cuh:244-265, 540-583): a BLENDED pixel keeps its
  full-resolution value up to blend_radius + 16 px from the clamped gaze,
  takes the stride-2 fill (bilinear, or with blur the blurred fill) from
  blend_radius + blend_width on, and in between is a convex mix of the two
  whose weight grows with the distance, linearly from 0 to 1 across the
  ring (the reference's cross-fade is a clamped linear ramp in the distance).

A swapped pair of parity branches, a wrong ring offset or width, or swapped
bf / 1 - bf each fail one of these.
"""
import numpy as np
import pytest

from foveatedhtgscudabackend_amd import tables as T
from tests.test_oracle_pins import _axis_camera, _scene, _tables

TL = 16  # large tile


def _view(w, h, gaze, fr, bw, vis=None, blur=False, chw=False):
    cam = _axis_camera(w, h)
    tabs = _tables(w, h, fr, bw, vis)
    import oracle
    return oracle.view_from_camera(cam, gaze, tabs, foveation_radius_px=fr, blend_width_px=bw,
                                   blur_periphery=blur, to_chw=chw)


def _types(oracle_lib, view):
    """Per large tile: -1 invisible, else the type of its first partitioned entry
    (the classification is pinned by brute force in test_oracle_pins.py)."""
    o = oracle_lib.run(_scene(np.zeros((0, 3)), [], []), view, stop_after=1)
    gw, gh = T.grid_size(view["width"], view["height"])
    t = np.full(gw * gh, -1, np.int64)
    t[o["tile_map_part"] // 4] = o["tile_type_part"]
    return t.reshape(gh, gw), o["count"]


def _to_hwc(img, chw, h, w):
    return img.reshape(3, h, w).transpose(1, 2, 0) if chw else img.reshape(h, w, 3)


def _samples(img, tt, lx, ly, ox, oy, n):
    """float64 stride-2 sample grid of n x n starting at pixel (ox, oy), as the
    reference loads it (interpolation.cuh:301-332): the last sample inside the
    image repeats past its edge, a sample in an invisible tile is clamped into
    large tile (lx, ly), coordinates are clamped to the image when read."""
    h, w = img.shape[:2]
    mx, my = (w - ox - 1) // 2, (h - oy - 1) // 2
    s = np.empty((n, n, 3))
    for j in range(n):
        for i in range(n):
            px, py = ox + min(i, mx) * 2, oy + min(j, my) * 2
            tx, ty = int(px / TL), int(py / TL)  # C division: toward zero
            if tt[ty, tx] < 0:
                px = min(max(px, TL * lx), TL * lx + TL - 2)
                py = min(max(py, TL * ly), TL * ly + TL - 2)
            s[j, i] = img[min(max(py, 0), h - 1), min(max(px, 0), w - 1)]
    return s


def _upsample_then_blur(s, size):
    """2x bilinear upsample of samples s (sample k at pixel offset 2(k-1)),
    then the separable [1 2 1]/4 blur; returns pixel offsets 0..size-1."""
    n = s.shape[0]
    up = np.empty((2 * n - 1, 2 * n - 1, 3))  # pixel offsets -2 .. 2n-4
    up[0::2, 0::2] = s
    up[1::2, 0::2] = 0.5 * (s[:-1] + s[1:])
    up[:, 1::2] = 0.5 * (up[:, 0:-1:2] + up[:, 2::2])
    # pixel offset x sits at index x + 2; the blur at x reads x-1, x, x+1
    k = np.array([0.25, 0.5, 0.25])
    rows = sum(k[a] * up[1 + a:1 + a + size + 0, :] for a in range(3))
    return sum(k[b] * rows[:, 1 + b:1 + b + size] for b in range(3))


def _bilinear(s, size):
    """2x bilinear fill from stride-2 samples (sample k at pixel offset 2k)."""
    n = s.shape[0]
    up = np.empty((2 * n - 1, 2 * n - 1, 3))
    up[0::2, 0::2] = s
    up[1::2, 0::2] = 0.5 * (s[:-1] + s[1:])
    up[:, 1::2] = 0.5 * (up[:, 0:-1:2] + up[:, 2::2])
    return up[:size, :size]


@pytest.mark.parametrize("w,h,gaze,fr,bw,frac,chw", [
    (320, 240, (40.0, 30.0), 64, 32, 0.0, False),
    (333, 257, (300.0, 200.0), 96, 48, 0.25, True),
    (200, 150, (-30.0, 500.0), 48, 20, 0.4, False),
])
def test_blur_is_upsample_then_binomial(oracle_lib, w, h, gaze, fr, bw, frac, chw):
    gw, gh = T.grid_size(w, h)
    rng = np.random.default_rng(w + h)
    vis = rng.random((gh, gw)) >= frac if frac else None
    view = _view(w, h, gaze, fr, bw, vis, blur=True, chw=chw)
    tt, _ = _types(oracle_lib, view)
    shape = (3, h, w) if chw else (h, w, 3)
    temp = rng.random(shape, dtype=np.float32)
    final = rng.random(shape, dtype=np.float32)
    out = _to_hwc(oracle_lib.upsample(view, final, temp), chw, h, w)
    t = _to_hwc(temp, chw, h, w)
    worst, tiles = 0.0, 0
    for ly in range(gh):
        for lx in range(gw):
            if tt[ly, lx] != 1:  # PERIPHERY
                continue
            s = _samples(t, tt, lx, ly, TL * lx - 2, TL * ly - 2, 10)
            exp = np.clip(_upsample_then_blur(s, TL), 0.0, 1.0)
            ys, xs = slice(TL * ly, min(h, TL * ly + TL)), slice(TL * lx, min(w, TL * lx + TL))
            got = out[ys, xs]
            worst = max(worst, float(np.abs(got - exp[:got.shape[0], :got.shape[1]]).max()))
            tiles += 1
    assert tiles >= 20
    assert worst <= 1e-6, worst


@pytest.mark.parametrize("blur", [False, True])
def test_impulse_response(oracle_lib, blur):
    w = h = 160
    view = _view(w, h, (0.0, 0.0), 32, 20, blur=blur)
    tt, _ = _types(oracle_lib, view)
    assert (tt[3:7, 3:7] == 1).all(), "the impulse's neighbourhood is PERIPHERY"
    src = np.zeros((h, w, 3), np.float32)
    src[80, 80] = 1.0
    if blur:
        out = oracle_lib.upsample(view, np.zeros_like(src), src)
        r = np.array([1 / 8, 1 / 2, 3 / 4, 1 / 2, 1 / 8])
    else:
        out = oracle_lib.upsample(view, src, None)
        r = np.array([0.0, 1 / 2, 1.0, 1 / 2, 0.0])  # the sample itself keeps its value
    want = np.zeros((h, w))
    want[78:83, 78:83] = np.outer(r, r)
    # dyadic weights: exact in float32
    region = tt.repeat(TL, 0).repeat(TL, 1) == 1
    for c in range(3):
        np.testing.assert_array_equal(out[..., c][region], want[region].astype(np.float32))


@pytest.mark.parametrize("w,h,gaze,fr,bw,frac,chw,blur", [
    (640, 480, (320.0, 240.0), 200, 100, 0.0, False, False),
    (640, 480, (301.7, 255.2), 200, 100, 0.0, False, True),
    (1280, 960, (700.3, 450.6), 375, 125, 0.15, True, False),
    (1280, 960, (611.9, 470.1), 375, 125, 0.15, False, True),
])
def test_blend_ring_cross_fade(oracle_lib, w, h, gaze, fr, bw, frac, chw, blur):
    gw, gh = T.grid_size(w, h)
    rng = np.random.default_rng(int(gaze[0]))
    vis = rng.random((gh, gw)) >= frac if frac else None
    view = _view(w, h, gaze, fr, bw, vis, blur=blur, chw=chw)
    tt, c = _types(oracle_lib, view)
    r_blend = TL * c["blend_tiles"]
    shape = (3, h, w) if chw else (h, w, 3)
    img = rng.random(shape, dtype=np.float32)
    temp = rng.random(shape, dtype=np.float32) if blur else None
    out = _to_hwc(oracle_lib.upsample(view, img, temp), chw, h, w)
    full = _to_hwc(temp if blur else img, chw, h, w)  # the full-resolution render of the ring
    gx, gy = np.clip(gaze[0], 0, w - 1), np.clip(gaze[1], 0, h - 1)
    near, far = r_blend + TL, r_blend + bw
    n_near = n_far = 0
    mix_d, mix_w = [], []
    for ly in range(gh):
        for lx in range(gw):
            if tt[ly, lx] != 2:  # BLENDED
                continue
            for st in range(4):
                ox, oy = TL * lx + 8 * (st % 2), TL * ly + 8 * (st // 2)
                if blur:
                    fill = _upsample_then_blur(_samples(full, tt, lx, ly, ox - 2, oy - 2, 6), 8)
                else:
                    fill = _bilinear(_samples(full, tt, lx, ly, ox, oy, 5), 8)
                for ty in range(8):
                    for tx in range(8):
                        px, py = ox + tx, oy + ty
                        if px >= w or py >= h:
                            continue
                        d = float(np.hypot(gx - px, gy - py))
                        got, ful = out[py, px].astype(np.float64), full[py, px].astype(np.float64)
                        if not blur and tx % 2 == 0 and ty % 2 == 0:
                            assert np.array_equal(out[py, px], full[py, px]), "samples keep their value"
                            continue
                        if d <= near - 1e-3:
                            assert np.array_equal(out[py, px], full[py, px]), (px, py, d)
                            n_near += 1
                        elif d >= far + 1e-3:
                            assert np.abs(got - np.clip(fill[ty, tx], 0, 1)).max() <= 1e-6, (px, py, d)
                            n_far += 1
                        else:
                            diff = fill[ty, tx] - ful
                            ok = np.abs(diff) > 0.05
                            if ok.sum() < 2:
                                continue
                            wts = (got - ful)[ok] / diff[ok]
                            assert np.ptp(wts) <= 1e-4, (px, py, wts)
                            assert -1e-5 <= wts.mean() <= 1 + 1e-5
                            mix_d.append(d)
                            mix_w.append(float(wts.mean()))
    assert n_near > 50 and n_far > 50 and len(mix_d) > 200
    d, wt = np.array(mix_d), np.array(mix_w)
    order = np.argsort(d)
    assert np.all(np.diff(wt[order]) >= -1e-4), "the weight does not decrease with the distance"
    # linear in the distance, 0 at the inner and 1 at the outer edge of the ring
    slope, icpt = np.polyfit(d, wt, 1)
    assert np.abs(wt - (slope * d + icpt)).max() <= 1e-4
    assert abs(-icpt / slope - near) <= 0.05 and abs((1 - icpt) / slope - far) <= 0.05
