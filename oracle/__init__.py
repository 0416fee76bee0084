"""CPU oracle of the foveated HTGS rasterizer: TEST INFRASTRUCTURE ONLY.

A single-threaded scalar C++ restatement of the reference inference path
(oracle/fhtgs_oracle.cpp, built by oracle/Makefile into oracle/_build/). Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use it,
and only as the checker / reported baseline -- never as the product path.

Parity pins (tests/test_oracle_*.py): transcendentals against libm, the SH
basis against the textbook real spherical harmonics, tile classification and
culling tables by brute force, projected extents against the closed form for
an isotropic Gaussian (and, anti-aliased, the dilated opacity and the tangent
cone), the anti-aliased culls, instance enumeration by brute force, the sort
against a stable sort, the blend against its closed forms (single Gaussian,
pure OIT tail, over-compositing when K covers all fragments, order
invariance), both background models against their closed forms, upsampling
and blur on constant images. Not pinned by any reference output: the
reference ships no tests, fixtures or golden images and cannot be built here
(CUDA only).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libfhtgs_oracle.so")

IDS = {
    "counts": 0, "tile_num": 1, "tile_map": 2, "tile_type": 3, "tile_map_part": 4,
    "tile_type_part": 5, "partition": 6, "n_touched": 7, "bounds": 8, "vpmt1": 9, "vpmt2": 10,
    "vpmt4": 11, "rgba": 12, "offsets": 13, "keys_unsorted": 14, "vals_unsorted": 15,
    "keys_sorted": 16, "vals_sorted": 17, "ranges": 18, "image": 19, "image_temp": 20, "timings": 21,
}
DTYPES = {
    "counts": np.int64, "tile_num": np.uint32, "tile_map": np.uint32, "tile_type": np.uint8,
    "tile_map_part": np.uint32, "tile_type_part": np.uint8, "partition": np.uint32,
    "n_touched": np.uint32, "bounds": np.uint16, "vpmt1": np.float32, "vpmt2": np.float32,
    "vpmt4": np.float32, "rgba": np.float32, "offsets": np.uint32, "keys_unsorted": np.uint32,
    "vals_unsorted": np.uint32, "keys_sorted": np.uint32, "vals_sorted": np.uint32,
    "ranges": np.uint32, "image": np.float32, "image_temp": np.float32, "timings": np.float64,
}
COUNT_NAMES = ("N", "n_large", "n_small", "n_active", "instances", "valid_instances", "end_bit",
               "key_bytes", "n_visible", "gw", "gh", "fovea_tiles", "blend_tiles", "gaze_tx",
               "gaze_ty", "evaluations", "pass_cutoff", "pass_alpha", "core_inserts", "quarter_tiles")
STAGES = ("tiles", "preprocess", "scan", "instances", "sort", "ranges", "blend", "upsample")


class OrcScene(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("scales", ctypes.c_void_p),
                ("rotations", ctypes.c_void_p), ("opacities", ctypes.c_void_p),
                ("sh_0", ctypes.c_void_p), ("sh_rest", ctypes.c_void_p),
                ("n_primitives", ctypes.c_int32), ("total_sh_bases", ctypes.c_int32)]


class OrcView(ctypes.Structure):
    _fields_ = [("M", ctypes.c_float * 12), ("VPM", ctypes.c_float * 16), ("VPR_inv", ctypes.c_float * 16),
                ("cam_position", ctypes.c_float * 3), ("gaze_position", ctypes.c_float * 2),
                ("visibility_mask", ctypes.c_void_p), ("visibility_mask_area_table", ctypes.c_void_p),
                ("fovea_mask_area_table", ctypes.c_void_p), ("background_model_data", ctypes.c_void_p),
                ("background_model_type", ctypes.c_int32), ("background_width", ctypes.c_int32),
                ("background_height", ctypes.c_int32), ("K", ctypes.c_int32),
                ("active_sh_bases", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("focal_x", ctypes.c_float), ("focal_y", ctypes.c_float), ("center_x", ctypes.c_float),
                ("center_y", ctypes.c_float), ("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
                ("scale_modifier", ctypes.c_float), ("to_chw", ctypes.c_int32),
                ("blur_periphery", ctypes.c_int32), ("anti_aliasing", ctypes.c_int32),
                ("foveation_radius_px", ctypes.c_int32), ("blend_width_px", ctypes.c_int32),
                ("blend_tile_stride", ctypes.c_int32), ("quarter_radius_px", ctypes.c_int32)]


_lib = None


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (g++, no GPU needed)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_run.argtypes = [ctypes.POINTER(OrcScene), ctypes.POINTER(OrcView), ctypes.c_int32,
                            ctypes.c_char_p, ctypes.c_int32]
    lib.orc_run.restype = ctypes.c_void_p
    lib.orc_result_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    lib.orc_result_bytes.restype = ctypes.c_int64
    lib.orc_result_copy.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64]
    lib.orc_result_copy.restype = ctypes.c_int32
    lib.orc_result_free.argtypes = [ctypes.c_void_p]
    lib.orc_result_free.restype = None
    lib.orc_upsample.argtypes = [ctypes.POINTER(OrcView), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                 ctypes.c_int32]
    lib.orc_upsample.restype = ctypes.c_int32
    lib.orc_build_visibility_sat.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.orc_build_fovea_sat.argtypes = [ctypes.c_int32, ctypes.c_void_p]
    lib.orc_stable_sort_pairs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32]
    for f in ("orc_expf", "orc_logf", "orc_asinf", "orc_tanf", "orc_tanhf"):
        getattr(lib, f).argtypes = [ctypes.c_float]
        getattr(lib, f).restype = ctypes.c_float
    lib.orc_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.orc_atan2f.restype = ctypes.c_float
    lib.orc_sh_eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    lib.orc_sh_eval.restype = None
    lib.orc_visualize_gaze.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32]
    lib.orc_visualize_gaze.restype = None
    lib.orc_visualize_tile_boundaries.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                                  ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_int32]
    lib.orc_visualize_tile_boundaries.restype = None
    _lib = lib
    return lib


def visualize_gaze(image: np.ndarray, gaze, rgb, vis_type: int, size: int, chw: bool) -> np.ndarray:
    """Draws the gaze marker into a copy of a float32 image (H,W,3) or (3,H,W)."""
    lib = load()
    img = np.ascontiguousarray(image, dtype=np.float32).copy()
    H, W = (img.shape[1], img.shape[2]) if chw else (img.shape[0], img.shape[1])
    col = np.ascontiguousarray(rgb, dtype=np.float32)
    lib.orc_visualize_gaze(_ptr(img), W, H, float(gaze[0]), float(gaze[1]), _ptr(col), int(vis_type), int(size),
                           int(bool(chw)))
    return img


def visualize_tile_boundaries(image: np.ndarray, mask_words: np.ndarray, gaze, foveation_radius_px: int = 375,
                              blend_width_px: int = 125, chw: bool = False) -> np.ndarray:
    """Draws the tile-type borders into a copy of a float32 image."""
    lib = load()
    img = np.ascontiguousarray(image, dtype=np.float32).copy()
    H, W = (img.shape[1], img.shape[2]) if chw else (img.shape[0], img.shape[1])
    m = np.ascontiguousarray(mask_words, dtype=np.uint32)
    lib.orc_visualize_tile_boundaries(_ptr(img), W, H, _ptr(m), float(gaze[0]), float(gaze[1]),
                                      int(foveation_radius_px), int(blend_width_px), int(bool(chw)))
    return img


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _make_view(view: dict, arr) -> OrcView:
    """OrcView from a view dict (see run); ``arr`` keeps converted arrays alive."""
    v = OrcView()
    M = np.asarray(view["M"], dtype=np.float32).reshape(-1)[:12]
    v.M[:] = list(M)
    v.VPM[:] = list(np.asarray(view["VPM"], dtype=np.float32).reshape(-1))
    v.VPR_inv[:] = list(np.asarray(view.get("VPR_inv", np.eye(4)), dtype=np.float32).reshape(-1))
    v.cam_position[:] = list(np.asarray(view["cam_position"], dtype=np.float32).reshape(-1))
    v.gaze_position[:] = list(np.asarray(view["gaze_position"], dtype=np.float32).reshape(-1))
    for k in ("visibility_mask", "visibility_mask_area_table", "fovea_mask_area_table"):
        setattr(v, k, _ptr(arr(view[k], np.uint32)))
    bg = view.get("background_model_data")
    v.background_model_type = int(view.get("background_model_type", 0))
    if bg is not None:
        bga = arr(bg, np.float32)
        v.background_model_data = _ptr(bga)
        if bga.ndim == 3:
            v.background_height, v.background_width = bga.shape[1], bga.shape[2]
    v.K = int(view.get("K", 8))
    v.active_sh_bases = int(view.get("active_sh_bases", 16))
    v.width, v.height = int(view["width"]), int(view["height"])
    for k in ("focal_x", "focal_y", "center_x", "center_y", "near_plane", "far_plane"):
        setattr(v, k, float(view[k]))
    v.scale_modifier = float(view.get("scale_modifier", 1.0))
    v.to_chw = int(bool(view.get("to_chw", False)))
    v.blur_periphery = int(bool(view.get("blur_periphery", False)))
    v.anti_aliasing = int(bool(view.get("anti_aliasing", False)))
    v.foveation_radius_px = int(view.get("foveation_radius_px", 375))
    v.blend_width_px = int(view.get("blend_width_px", 125))
    v.blend_tile_stride = int(view.get("blend_tile_stride", 1))
    v.quarter_radius_px = int(view.get("quarter_radius_px", 0))
    return v


def run(scene: dict, view: dict, stop_after: int = 0, want=None) -> dict:
    """Runs the oracle. ``scene``: numpy arrays as from scene.make_scene;
    ``view``: M (>=12 floats), VPM, VPR_inv, cam_position, gaze_position,
    visibility_mask, visibility_mask_area_table, fovea_mask_area_table
    (uint32 arrays), width, height, focal_x, focal_y, center_x, center_y,
    near_plane, far_plane and optional K, active_sh_bases, scale_modifier,
    to_chw, blur_periphery, foveation_radius_px, blend_width_px,
    quarter_radius_px (extension, 0 = off),
    background_model_type / background_model_data."""
    lib = load()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    sc = OrcScene()
    for k in ("positions", "scales", "rotations", "opacities", "sh_0", "sh_rest"):
        setattr(sc, k, _ptr(arr(scene[k], np.float32)))
    sc.n_primitives = int(scene["positions"].shape[0])
    sc.total_sh_bases = int(np.asarray(scene["sh_rest"]).shape[1])
    v = _make_view(view, arr)
    err = ctypes.create_string_buffer(512)
    r = lib.orc_run(ctypes.byref(sc), ctypes.byref(v), int(stop_after), err, 512)
    if not r:
        raise ValueError(err.value.decode())
    out = {}
    try:
        for name, idx in IDS.items():
            if want is not None and name not in want and name not in ("counts", "timings"):
                continue
            nb = lib.orc_result_bytes(r, idx)
            a = np.empty(nb // np.dtype(DTYPES[name]).itemsize, dtype=DTYPES[name])
            if nb:
                lib.orc_result_copy(r, idx, _ptr(a), nb)
            out[name] = a
    finally:
        lib.orc_result_free(r)
    out["count"] = dict(zip(COUNT_NAMES, (int(x) for x in out["counts"])))
    out["stage_ms"] = dict(zip(STAGES, (float(x) for x in out["timings"])))
    return out


def upsample(view: dict, image: np.ndarray, image_temp: np.ndarray = None) -> np.ndarray:
    """Runs the oracle's upsampling stage alone on caller images (the view
    dict as for run; only the mask, gaze, radii, size and flags are read).
    Returns the filled copy of ``image``; with blur_periphery the blur reads
    ``image_temp``. Images are float32 (H,W,3), or (3,H,W) with to_chw."""
    lib = load()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    v = _make_view(view, arr)
    out = np.ascontiguousarray(image, dtype=np.float32).copy()
    tmp = None if image_temp is None else np.ascontiguousarray(image_temp, dtype=np.float32)
    err = ctypes.create_string_buffer(512)
    rc = lib.orc_upsample(ctypes.byref(v), _ptr(out), None if tmp is None else _ptr(tmp), err, 512)
    if rc != 0:
        raise ValueError(err.value.decode())
    return out


def view_from_camera(cam, gaze, tables: tuple, **kw) -> dict:
    """Oracle view dict from a scene.Camera and numpy (mask, sat, fovea_sat)."""
    v = {"M": cam.M, "VPM": cam.VPM, "VPR_inv": cam.VPR_inv, "cam_position": cam.position,
         "gaze_position": np.asarray(gaze, dtype=np.float32), "visibility_mask": tables[0],
         "visibility_mask_area_table": tables[1], "fovea_mask_area_table": tables[2],
         "width": cam.width, "height": cam.height, "focal_x": cam.fx, "focal_y": cam.fy,
         "center_x": cam.cx, "center_y": cam.cy, "near_plane": cam.near, "far_plane": cam.far}
    v.update(kw)
    return v
