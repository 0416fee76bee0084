"""GPU parity: the gfx950 path through the reference-shaped operator surface
against the CPU oracle on the same seeded inputs.

Bar: bit-exact on every integer stage (tile classification and partition,
tiles-touched, bounds, sort keys, sorted order, tile ranges) and on the
float per-Gaussian outputs (VPMT rows, colour); the image within
max|dRGB| <= 2e-3 and PSNR >= 60 dB (the contract), and we also report how
many pixels are not bit-identical.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from tests.cases import Case, psnr

pytestmark = pytest.mark.gpu

TOL_MAX_ABS = 2e-3
TOL_PSNR = 60.0

CASES = [
    Case("nofov_256x192", 20000, 256, 192, seed=1, foveation_radius_px=-1),
    Case("fov_640x480", 40000, 640, 480, seed=3),
    Case("fov_gaze_corner", 40000, 640, 480, seed=4, gaze=(630.0, 10.0)),
    Case("fov_small_radius", 30000, 512, 384, seed=5, foveation_radius_px=150, blend_width_px=60),
    Case("fov_blur", 40000, 640, 480, seed=6, blur=True),
    Case("fov_chw_blur", 30000, 480, 352, seed=7, to_chw=True, blur=True),
    Case("odd_size", 25000, 333, 257, seed=9, foveation_radius_px=100, blend_width_px=40),
    Case("invisible_tiles", 30000, 640, 480, seed=10, invisible_frac=0.3),
    Case("invisible_blur", 30000, 640, 480, seed=11, invisible_frac=0.3, blur=True),
    Case("k1", 20000, 320, 240, seed=12, K=1),
    Case("k4", 20000, 320, 240, seed=13, K=4),
    Case("k16", 20000, 320, 240, seed=14, K=16),
    Case("k32_chw", 20000, 320, 240, seed=15, K=32, to_chw=True),
    Case("sh_deg1", 20000, 320, 240, seed=16, active_sh=4),
    Case("sh_deg0", 20000, 320, 240, seed=17, active_sh=1),
    Case("bg_sh", 15000, 320, 240, seed=18, background=1),
    Case("bg_tex", 15000, 320, 240, seed=19, background=2),
    Case("bg_tex_raw", 15000, 320, 240, seed=20, background=3, foveation_radius_px=120, blend_width_px=50),
    Case("scale_mod", 20000, 320, 240, seed=21, scale_modifier=1.7),
    Case("empty_scene", 0, 128, 96, seed=22),
    # the plain GPU drop-in configuration: 1M Gaussians in a 20 m cube, one 1920x1080 view, K=8,
    # the reference's foveation (375/125 px)
    Case("plain_1M_1920x1080", 1_000_000, 1920, 1080, seed=2, cube=20.0),
    # the plain CPU configuration: 100k Gaussians in a 10 m cube, 800x800 at 60 deg, K=4, no foveation
    Case("cfg1_100k_800x800_k4_nofov", 100_000, 800, 800, seed=0, cube=10.0, K=4, foveation_radius_px=-1),
    # the plain GPU configuration without foveation: 1M Gaussians, 20 m cube, 1920x1080 at 60 deg, K=8
    Case("cfg2_1M_1920x1080_nofov", 1_000_000, 1920, 1080, seed=0, cube=20.0, K=8, foveation_radius_px=-1),
    # scenes whose instance kernel takes 512 and 1024 Gaussians a workgroup (fewer than one
    # 2048-Gaussian workgroup per CU; 0.1M above takes 256, the 1M scenes 2048)
    Case("inst_512_per_wg_150k", 150_000, 640, 480, seed=24, cube=12.0),
    Case("inst_1024_per_wg_300k", 300_000, 800, 600, seed=25, cube=14.0, foveation_radius_px=200,
         blend_width_px=60),
    # more than 65535 small tiles: 32-bit sort keys (end_bit 17) through the whole pipeline
    Case("u32_keys_2048x2080", 20000, 2048, 2080, seed=23, foveation_radius_px=300, blend_width_px=100),
    Case("aaa_fov", 30000, 640, 480, seed=41, aaa=True),
    Case("aaa_blur_chw", 30000, 480, 352, seed=42, aaa=True, blur=True, to_chw=True),
    Case("aaa_nofov_k16", 20000, 320, 240, seed=43, aaa=True, K=16, foveation_radius_px=-1),
    # quarter-resolution level beyond quarter_radius_px
    Case("quarter_640x480", 40000, 640, 480, seed=60, foveation_radius_px=150, blend_width_px=60,
         quarter_radius_px=250),
    Case("quarter_invisible_chw", 30000, 640, 480, seed=61, foveation_radius_px=120, blend_width_px=50,
         quarter_radius_px=160, invisible_frac=0.3, to_chw=True),
    Case("quarter_odd_size", 25000, 333, 257, seed=62, foveation_radius_px=64, blend_width_px=32,
         quarter_radius_px=96),
    Case("quarter_at_fovea_k1", 20000, 480, 352, seed=63, foveation_radius_px=100, blend_width_px=40,
         quarter_radius_px=100, K=1),
    Case("quarter_bg_tex_aaa", 20000, 480, 352, seed=64, foveation_radius_px=120, blend_width_px=50,
         quarter_radius_px=200, background=2, aaa=True),
]


def _gpu_render(case: Case, sc, cam, gaze, visible, bg, benchmark=False):
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import scene as S
    r = FoveatedHTGSRasterizer(case.foveation_radius_px, case.blend_width_px, case.quarter_radius_px)
    st = case.settings(cam, gaze, visible, bg)
    args = S.scene_to_torch(sc, "cuda")
    fn = r.benchmark if benchmark else r.render
    img = fn(*args, st, case.to_chw, case.blur, case.aaa)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def _stage(h, name, dtype, count, view=0):
    raw = h.debug_read(view, name, int(count) * np.dtype(dtype).itemsize)
    return np.frombuffer(raw, dtype=dtype).copy()


def _compare_stages(case, o, h, view=0):
    c = o["count"]
    g = h.counts(view)
    N, n_large, n_small, n_active, I = c["N"], c["n_large"], c["n_small"], c["n_active"], c["instances"]
    assert g[0] == N and g[1] == n_large and g[2] == n_small
    assert g[3] == n_active, f"{case.name}: n_active {g[3]} != {n_active}"
    assert g[4] == I, f"{case.name}: instances {g[4]} != {I}"
    assert g[5] == c["valid_instances"], f"{case.name}: valid instances {g[5]} != {c['valid_instances']}"
    assert g[8] == c["n_visible"], f"{case.name}: visible {g[8]} != {c['n_visible']}"
    np.testing.assert_array_equal(_stage(h, "tile_num", np.uint32, n_large, view), o["tile_num"])
    np.testing.assert_array_equal(_stage(h, "tile_map", np.uint32, n_active, view), o["tile_map"])
    np.testing.assert_array_equal(_stage(h, "tile_type", np.uint8, n_active, view), o["tile_type"])
    np.testing.assert_array_equal(_stage(h, "tile_map_part", np.uint32, n_active, view), o["tile_map_part"])
    np.testing.assert_array_equal(_stage(h, "partition", np.uint32, 6, view), o["partition"][:6])
    np.testing.assert_array_equal(_stage(h, "partition_quarter", np.uint32, 2, view), o["partition"][6:8])
    nt = _stage(h, "n_touched", np.uint32, N, view)
    np.testing.assert_array_equal(nt, o["n_touched"])
    vis = nt > 0
    b = _stage(h, "bounds", np.uint16, 4 * N, view).reshape(-1, 4)
    np.testing.assert_array_equal(b[vis], o["bounds"].reshape(-1, 4)[vis])
    for name in ("vpmt1", "vpmt2", "vpmt4", "rgba"):
        a = _stage(h, name, np.float32, 4 * N, view).reshape(-1, 4)[vis]
        e = o[name].reshape(-1, 4)[vis]
        bad = np.count_nonzero(a.view(np.uint32) != e.view(np.uint32))
        assert bad == 0, f"{case.name}: {name} differs in {bad} words (max abs {np.abs(a - e).max()})"
    keys = _stage(h, "keys_sorted", np.uint32, I, view)
    vals = _stage(h, "vals_sorted", np.uint32, I, view)
    np.testing.assert_array_equal(keys, o["keys_sorted"])
    sent = 0xFFFF if c["key_bytes"] == 2 else 0xFFFFFFFF
    valid = keys != sent
    np.testing.assert_array_equal(vals[valid], o["vals_sorted"][valid])
    np.testing.assert_array_equal(_stage(h, "ranges", np.uint32, 2 * n_small, view), o["ranges"])


def _compare_image(name, got, want):
    assert got.shape == want.shape
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    n_diff = int(np.count_nonzero(got.view(np.uint32) != want.view(np.uint32)))
    p = psnr(got, want)
    print(f"{name}: max|d|={d.max():.3g} psnr={p:.1f} non-identical={n_diff}/{got.size}")
    assert d.max() <= TOL_MAX_ABS, f"{name}: max |dRGB| {d.max()} > {TOL_MAX_ABS}"
    assert p >= TOL_PSNR, f"{name}: PSNR {p} < {TOL_PSNR}"
    return n_diff


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_render_matches_oracle(case, oracle_lib):
    from foveatedhtgscudabackend_amd import _lib
    sc, cam, gaze, visible, tables, bg = case.build()
    o = oracle_lib.run(sc, case.oracle_view(cam, gaze, tables, bg))
    print(f"{case.name}: {o['count']}")
    got = _gpu_render(case, sc, cam, gaze, visible, bg)
    h = _lib.handle_for(0)
    _compare_stages(case, o, h)
    shape = (3, case.height, case.width) if case.to_chw else (case.height, case.width, 3)
    n_diff = _compare_image(case.name, got, o["image"].reshape(shape))
    assert n_diff == 0, f"{case.name}: {n_diff} values not bit-identical to the oracle"


def test_benchmark_matches_render_on_visible_tiles(oracle_lib):
    case = Case("bench_vs_render", 30000, 640, 480, seed=31, invisible_frac=0.25)
    sc, cam, gaze, visible, tables, bg = case.build()
    ref = _gpu_render(case, sc, cam, gaze, visible, bg)
    got = _gpu_render(case, sc, cam, gaze, visible, bg, benchmark=True)
    px_vis = np.repeat(np.repeat(visible, 16, 0), 16, 1)[: case.height, : case.width]
    np.testing.assert_array_equal(got[px_vis], ref[px_vis])
    assert np.all(ref[~px_vis] == 0.0)


def test_render_is_deterministic():
    case = Case("determinism", 40000, 640, 480, seed=32, blur=True)
    sc, cam, gaze, visible, tables, bg = case.build()
    a = _gpu_render(case, sc, cam, gaze, visible, bg)
    b = _gpu_render(case, sc, cam, gaze, visible, bg)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_blend_xcd_run_does_not_change_image():
    """The tile-to-workgroup map is a bijection: every run length (including
    ones longer than the tile count) gives the same image."""
    from foveatedhtgscudabackend_amd import _lib
    case = Case("xcd_run", 30000, 640, 480, seed=34, blur=True)
    sc, cam, gaze, visible, tables, bg = case.build()
    h = _lib.handle_for(0)
    try:
        h.set_option(_lib.OPT_BLEND_XCD_RUN, 0)
        ref = _gpu_render(case, sc, cam, gaze, visible, bg)
        for run in (1, 3, 16, 64, 4096):
            h.set_option(_lib.OPT_BLEND_XCD_RUN, run)
            got = _gpu_render(case, sc, cam, gaze, visible, bg)
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=f"run={run}")
    finally:
        h.set_option(_lib.OPT_BLEND_XCD_RUN, 16)


def test_stereo_shared_preprocess_equals_two_mono(oracle_lib):
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import _lib
    from foveatedhtgscudabackend_amd import scene as S
    case = Case("stereo", 40000, 640, 480, seed=33, fov_deg=100.0)
    sc = S.make_scene(case.n, seed=case.seed, cube=case.cube)
    left, right = S.stereo_cameras(case.width, case.height, case.fov_deg, seed=case.seed + 1000)
    gaze = S.seeded_gaze(left, case.seed + 2000)
    r = FoveatedHTGSRasterizer()
    sl = S.settings_for(left, gaze, K=case.K)
    sr = S.settings_for(right, gaze, K=case.K)
    args = S.scene_to_torch(sc, "cuda")
    li, ri = r.render_stereo(*args, sl, sr, False, False, False)
    torch.cuda.synchronize()
    li, ri = li.cpu().numpy(), ri.cpu().numpy()
    h = _lib.handle_for(0)
    counts_stereo = [h.counts(0), h.counts(1)]
    lm = r.render(*args, sl, False, False, False).cpu().numpy()
    rm = r.render(*args, sr, False, False, False).cpu().numpy()
    np.testing.assert_array_equal(li.view(np.uint32), lm.view(np.uint32))
    np.testing.assert_array_equal(ri.view(np.uint32), rm.view(np.uint32))
    import oracle
    from foveatedhtgscudabackend_amd import tables as T
    for cam, img, cnt in ((left, li, counts_stereo[0]), (right, ri, counts_stereo[1])):
        words = T.visibility_mask_np(cam.width, cam.height)
        tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height), T.fovea_area_table_np(24))
        o = oracle.run(sc, oracle.view_from_camera(cam, gaze, tabs, K=case.K))
        assert cnt[4] == o["count"]["instances"]
        _compare_image("stereo eye", img, o["image"].reshape(case.height, case.width, 3))


@pytest.mark.parametrize("levels", [3, 2])
def test_stereo_configuration_matches_oracle(levels, oracle_lib):
    """The stereo configuration the frame rate is measured on, replayed from
    its captured graph: 3M Gaussians, two 2064x2208 eyes at 100 deg, K=8,
    with three levels (full resolution within 15 deg of the gaze, half to
    35 deg, quarter beyond) or the reference's two (375/125 px). Every stage
    and the whole image of both eyes against the oracle's mono render of each
    eye, bit for bit."""
    from foveatedhtgscudabackend_amd import FoveatedHTGSRasterizer
    from foveatedhtgscudabackend_amd import _lib
    from foveatedhtgscudabackend_amd import scene as S
    from foveatedhtgscudabackend_amd import tables as T
    from bench import foveation_radii
    w, h_px, fov = 2064, 2208, 100.0
    fr, bw, qr = foveation_radii(levels, w, fov)
    sc = S.make_scene(3_000_000, seed=0, cube=20.0)
    left, right = S.stereo_cameras(w, h_px, fov, seed=1000)
    gaze = S.seeded_gaze(left, 2000)
    sl = S.settings_for(left, gaze, K=8, foveation_radius_px=fr, blend_width_px=bw)
    sr = S.settings_for(right, gaze, K=8, foveation_radius_px=fr, blend_width_px=bw)
    args = S.scene_to_torch(sc, "cuda")
    r = FoveatedHTGSRasterizer(fr, bw, qr)
    h = _lib.handle_for(0)
    for _ in range(3):  # exact, then captured, then replayed
        images = r.render_stereo(*args, sl, sr, False, False, False)
    torch.cuda.synchronize()
    views = []
    for cam in (left, right):
        words = T.visibility_mask_np(cam.width, cam.height)
        tabs = (words, T.visibility_area_table_np(words, cam.width, cam.height),
                T.fovea_area_table_np(T.fovea_radius_tiles_for(cam.width, cam.height, fr, bw)))
        views.append(oracle_lib.view_from_camera(cam, gaze, tabs, K=8, foveation_radius_px=fr,
                                                 blend_width_px=bw, quarter_radius_px=qr))
    # the oracle is a pure C function called without the GIL: one thread per eye
    oracle_lib.load()
    with ThreadPoolExecutor(max_workers=2) as ex:
        outs = list(ex.map(lambda ov: oracle_lib.run(sc, ov), views))
    for view, (o, img) in enumerate(zip(outs, images)):
        case = Case(f"stereo_3M_{levels}level_eye{view}", 3_000_000, w, h_px, seed=0)
        _compare_stages(case, o, h, view)
        n_diff = _compare_image(case.name, img.cpu().numpy(), o["image"].reshape(h_px, w, 3))
        assert n_diff == 0, f"{case.name}: {n_diff} values not bit-identical to the oracle"


@pytest.mark.parametrize("n,key_bytes,end_bit", [
    (0, 2, 12), (1, 2, 12), (4095, 2, 15), (4096, 4, 17), (4097, 4, 17), (6143, 4, 17), (6144, 2, 16),
    (6145, 4, 17), (12_289, 4, 22), (100_000, 2, 16),
    (1_000_003, 4, 17), (2_000_000, 4, 22), (300_000, 4, 31)])
def test_sort_pairs_is_stable_lsd(n, key_bytes, end_bit):
    from foveatedhtgscudabackend_amd import _lib
    rng = np.random.default_rng(n + end_bit)
    hi = 1 << min(end_bit, 8 * key_bytes)
    kd = np.uint16 if key_bytes == 2 else np.uint32
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(kd)
    if n > 10:
        keys[rng.random(n) < 0.1] = np.iinfo(kd).max  # sentinels
    vals = np.arange(n, dtype=np.uint32)
    tk = torch.from_numpy(keys.view(np.int16 if key_bytes == 2 else np.int32).copy()).cuda()
    tv = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    h = _lib.handle_for(0)
    h.sort_pairs(tk.data_ptr(), tv.data_ptr(), n, key_bytes, end_bit, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    gk = tk.cpu().numpy().view(kd)
    gv = tv.cpu().numpy().view(np.uint32)
    mask = (1 << end_bit) - 1
    order = np.argsort(keys.astype(np.uint64) & mask, kind="stable")
    np.testing.assert_array_equal(gk, keys[order])
    np.testing.assert_array_equal(gv, vals[order])
