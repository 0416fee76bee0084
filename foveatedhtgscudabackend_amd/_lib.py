"""ctypes binding of libfhtgs.so, the C ABI declared in include/fhtgs.h.

The library is built in-tree (``foveatedhtgscudabackend_amd/_build/libfhtgs.so``)
by ``__graft_entry__.build()`` / ``make -C foveatedhtgscudabackend_amd/csrc``.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import functools
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libfhtgs.so")

FHTGS_OK = 0
FHTGS_E_INVALID = -1
FHTGS_E_HIP = -2
FHTGS_E_OUT_OF_MEMORY = -3
FHTGS_E_TABLES = -4
FHTGS_E_UNSUPPORTED = -5

DEFAULT_FOVEATION_RADIUS_PX = 375
DEFAULT_BLEND_WIDTH_PX = 125
FOVEATION_OFF = -1

DBG = {
    "tile_num": 1, "tile_map": 2, "tile_type": 3, "tile_map_part": 4, "partition": 6,
    "n_touched": 7, "bounds": 8, "vpmt1": 9, "vpmt2": 10, "vpmt4": 11, "rgba": 12,
    "keys_unsorted": 14, "vals_unsorted": 15, "keys_sorted": 16, "vals_sorted": 17,
    "ranges": 18, "partition_quarter": 19, "blend_order": 20,
}
OPT_KEEP_UNSORTED = 1
OPT_SORT_DIGIT_BITS = 2
OPT_STAGE_TIMING = 3
OPT_BLEND_XCD_RUN = 4
OPT_FRAME_MODE = 5
OPT_VIEW_BRANCHES = 6
OPT_BLEND_WORK_TILE = 7  # waves per blend workgroup: 1 (8x8), 2 (16x8), 4 (16x16)
FRAME_EXACT, FRAME_DEVICE, FRAME_GRAPH = 0, 1, 2
NUM_STAGES = 11
STAGE_NAMES = ("clear", "tiles", "preprocess", "scan", "sync", "instances", "sort", "ranges",
               "blend", "upsample", "total")

EXPORTED_SYMBOLS = (
    "fhtgs_create", "fhtgs_destroy", "fhtgs_last_error", "fhtgs_version", "fhtgs_render",
    "fhtgs_workspace_bytes", "fhtgs_workspace_query", "fhtgs_sort_pairs", "fhtgs_fovea_radius_tiles",
    "fhtgs_debug_read", "fhtgs_debug_counts", "fhtgs_set_option", "fhtgs_stage_times",
    "fhtgs_render_batch", "fhtgs_autotune", "fhtgs_tuned_digit_bits", "fhtgs_tuned_work_tile",
    "fhtgs_get_option",
    "fhtgs_visualize_gaze", "fhtgs_visualize_tile_boundaries", "fhtgs_build_tables",
    "fhtgs_validate_view", "fhtgs_sort_pairs_noalloc",
)
SORT_NOALLOC_MAX_PAIRS = 1 << 20  # FHTGS_SORT_NOALLOC_MAX_PAIRS
MAX_LARGE_TILES = 65536  # 16-px tiles per view (4096 x 4096 pixels)


class Scene(ctypes.Structure):
    _fields_ = [
        ("positions", ctypes.c_void_p), ("scales", ctypes.c_void_p),
        ("rotations", ctypes.c_void_p), ("opacities", ctypes.c_void_p),
        ("sh_0", ctypes.c_void_p), ("sh_rest", ctypes.c_void_p),
        ("n_primitives", ctypes.c_int64), ("total_sh_bases", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class View(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_void_p), ("VPM", ctypes.c_void_p), ("VPR_inv", ctypes.c_void_p),
        ("cam_position", ctypes.c_void_p),
        ("gaze_x", ctypes.c_float), ("gaze_y", ctypes.c_float),
        ("visibility_mask", ctypes.c_void_p), ("visibility_mask_area_table", ctypes.c_void_p),
        ("fovea_mask_area_table", ctypes.c_void_p), ("background_model_data", ctypes.c_void_p),
        ("background_model_type", ctypes.c_int32), ("background_width", ctypes.c_int32),
        ("background_height", ctypes.c_int32), ("K", ctypes.c_int32),
        ("active_sh_bases", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("focal_x", ctypes.c_float), ("focal_y", ctypes.c_float),
        ("center_x", ctypes.c_float), ("center_y", ctypes.c_float),
        ("near_plane", ctypes.c_float), ("far_plane", ctypes.c_float),
        ("scale_modifier", ctypes.c_float),
        ("to_chw", ctypes.c_int32), ("blur_periphery", ctypes.c_int32),
        ("anti_aliasing", ctypes.c_int32), ("foveation_radius_px", ctypes.c_int32),
        ("blend_width_px", ctypes.c_int32), ("clear_output", ctypes.c_int32),
        ("image", ctypes.c_void_p), ("image_temp", ctypes.c_void_p),
        ("quarter_radius_px", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


_lib = None
_lib_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Loads libfhtgs.so once; raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libfhtgs.so not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or "
                "`make -C foveatedhtgscudabackend_amd/csrc`")
        lib = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        lib.fhtgs_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        lib.fhtgs_create.restype = ctypes.c_int
        lib.fhtgs_destroy.argtypes = [vp]
        lib.fhtgs_destroy.restype = ctypes.c_int
        lib.fhtgs_last_error.argtypes = [vp]
        lib.fhtgs_last_error.restype = ctypes.c_char_p
        lib.fhtgs_version.argtypes = []
        lib.fhtgs_version.restype = ctypes.c_char_p
        lib.fhtgs_render.argtypes = [vp, ctypes.POINTER(Scene), ctypes.POINTER(View), i32, i32, vp]
        lib.fhtgs_render.restype = ctypes.c_int
        lib.fhtgs_render_batch.argtypes = [vp, ctypes.POINTER(Scene), i32, ctypes.POINTER(View),
                                           ctypes.POINTER(i32), ctypes.POINTER(i32), i32, vp]
        lib.fhtgs_render_batch.restype = ctypes.c_int
        lib.fhtgs_autotune.argtypes = [vp, ctypes.POINTER(Scene), i32, ctypes.POINTER(View),
                                       ctypes.POINTER(i32), ctypes.POINTER(i32), i32, i32, vp]
        lib.fhtgs_autotune.restype = ctypes.c_int
        lib.fhtgs_tuned_digit_bits.argtypes = [vp, i32, i32, ctypes.POINTER(i32)]
        lib.fhtgs_tuned_digit_bits.restype = ctypes.c_int
        lib.fhtgs_tuned_work_tile.argtypes = [vp, i32, i32, ctypes.POINTER(i32)]
        lib.fhtgs_tuned_work_tile.restype = ctypes.c_int
        lib.fhtgs_get_option.argtypes = [vp, i32, ctypes.POINTER(i64)]
        lib.fhtgs_get_option.restype = ctypes.c_int
        f32 = ctypes.c_float
        lib.fhtgs_visualize_gaze.argtypes = [vp, vp, i32, i32, f32, f32, ctypes.POINTER(f32), i32, i32, i32, vp]
        lib.fhtgs_visualize_gaze.restype = ctypes.c_int
        lib.fhtgs_visualize_tile_boundaries.argtypes = [vp, vp, i32, i32, vp, f32, f32, i32, i32, i32, vp]
        lib.fhtgs_visualize_tile_boundaries.restype = ctypes.c_int
        lib.fhtgs_build_tables.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]
        lib.fhtgs_build_tables.restype = ctypes.c_int
        lib.fhtgs_workspace_bytes.argtypes = [vp, ctypes.POINTER(i64)]
        lib.fhtgs_workspace_bytes.restype = ctypes.c_int
        lib.fhtgs_workspace_query.argtypes = [ctypes.POINTER(View), i64, i64, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        lib.fhtgs_workspace_query.restype = ctypes.c_int
        lib.fhtgs_sort_pairs.argtypes = [vp, vp, vp, i64, i32, i32, vp]
        lib.fhtgs_sort_pairs.restype = ctypes.c_int
        lib.fhtgs_fovea_radius_tiles.argtypes = [i32, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fhtgs_fovea_radius_tiles.restype = ctypes.c_int
        lib.fhtgs_debug_read.argtypes = [vp, i32, i32, vp, i64, ctypes.POINTER(i64)]
        lib.fhtgs_debug_read.restype = ctypes.c_int
        lib.fhtgs_debug_counts.argtypes = [vp, i32, ctypes.POINTER(i64)]
        lib.fhtgs_debug_counts.restype = ctypes.c_int
        lib.fhtgs_set_option.argtypes = [vp, i32, i64]
        lib.fhtgs_set_option.restype = ctypes.c_int
        lib.fhtgs_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
        lib.fhtgs_stage_times.restype = ctypes.c_int
        lib.fhtgs_validate_view.argtypes = [ctypes.POINTER(View), ctypes.c_char_p, i32]
        lib.fhtgs_validate_view.restype = ctypes.c_int
        lib.fhtgs_sort_pairs_noalloc.argtypes = [vp, vp, i64, i32, i32, vp, vp]
        lib.fhtgs_sort_pairs_noalloc.restype = ctypes.c_int
        _lib = lib
        return lib


@functools.lru_cache(maxsize=256)
def fovea_radius_tiles(width: int, height: int, foveation_radius_px: int = DEFAULT_FOVEATION_RADIUS_PX,
                       blend_width_px: int = DEFAULT_BLEND_WIDTH_PX) -> tuple[int, int]:
    """(fovea, blend-ring) radii in 16-px tiles; the fovea table is built for the first."""
    lib = load()
    r, b = ctypes.c_int32(), ctypes.c_int32()
    rc = lib.fhtgs_fovea_radius_tiles(width, height, foveation_radius_px, blend_width_px,
                                      ctypes.byref(r), ctypes.byref(b))
    if rc != FHTGS_OK:
        raise ValueError("invalid image size")
    return r.value, b.value


def validate_view(view: View) -> tuple[int, str]:
    """(status, reason) of the library's own checks of a view, on the host
    only (no device call)."""
    lib = load()
    msg = ctypes.create_string_buffer(256)
    rc = lib.fhtgs_validate_view(ctypes.byref(view), msg, 256)
    return rc, msg.value.decode()


def workspace_query(view: View, n_primitives: int, n_instances: int) -> tuple[int, int]:
    """(fixed, instance) device bytes one view slot needs for `view` over
    n_primitives Gaussians with room for n_instances tile instances, before
    any render (host only, no device call); raises ValueError for a view the
    library rejects."""
    lib = load()
    f, i = ctypes.c_int64(0), ctypes.c_int64(0)
    rc = lib.fhtgs_workspace_query(ctypes.byref(view), int(n_primitives), int(n_instances),
                                   ctypes.byref(f), ctypes.byref(i))
    if rc != FHTGS_OK:
        raise ValueError(f"fhtgs_workspace_query: {lib.fhtgs_last_error(None).decode()} (status {rc})")
    return f.value, i.value


def sort_pairs_noalloc(keys_ptr: int, vals_ptr: int, n: int, key_bytes: int, end_bit: int,
                       error_ptr: int | None, stream: int) -> None:
    """fhtgs_sort_pairs_noalloc: the stable radix sort with the library's fixed
    per-device scratch (no handle, no allocation; enqueued on `stream` of the
    current device)."""
    lib = load()
    rc = lib.fhtgs_sort_pairs_noalloc(keys_ptr, vals_ptr, n, key_bytes, end_bit, error_ptr, stream)
    if rc != FHTGS_OK:
        err = RuntimeError if rc == FHTGS_E_HIP else ValueError
        raise err(f"fhtgs_sort_pairs_noalloc: {lib.fhtgs_last_error(None).decode()} (status {rc})")


class Handle:
    """One libfhtgs context (streams' scratch, options) on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.fhtgs_create(int(device), ctypes.byref(h))
        if rc != FHTGS_OK or not h.value:
            raise RuntimeError(f"fhtgs_create(device={device}) failed with status {rc}")
        self.ptr = h
        self.device = int(device)
        self.lock = threading.Lock()

    def check(self, rc: int, what: str) -> None:
        if rc != FHTGS_OK:
            msg = self.lib.fhtgs_last_error(self.ptr).decode(errors="replace")
            if rc == FHTGS_E_INVALID:
                raise ValueError(f"{what}: {msg}")
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    def set_option(self, option: int, value: int) -> None:
        self.check(self.lib.fhtgs_set_option(self.ptr, option, int(value)), "fhtgs_set_option")

    def render(self, scene: Scene, views: list, share_preprocess: bool, stream: int) -> None:
        arr = (View * len(views))(*views)
        with self.lock:
            rc = self.lib.fhtgs_render(self.ptr, ctypes.byref(scene), arr, len(views),
                                       1 if share_preprocess else 0, ctypes.c_void_p(stream))
            self.check(rc, "fhtgs_render")

    @staticmethod
    def _batch_arrays(scenes, views, view_scene, view_slot):
        sarr = (Scene * len(scenes))(*scenes)
        varr = (View * len(views))(*views)
        iarr = (ctypes.c_int32 * len(views))(*[int(k) for k in view_scene])
        slots = None if view_slot is None else (ctypes.c_int32 * len(views))(*[int(k) for k in view_slot])
        return sarr, varr, iarr, slots

    def render_batch(self, scenes: list, views: list, view_scene: list, stream: int,
                     view_slot: list | None = None) -> None:
        sarr, varr, iarr, slots = self._batch_arrays(scenes, views, view_scene, view_slot)
        with self.lock:
            rc = self.lib.fhtgs_render_batch(self.ptr, sarr, len(scenes), varr, iarr, slots, len(views),
                                             ctypes.c_void_p(stream))
            self.check(rc, "fhtgs_render_batch")

    def autotune(self, scenes: list, views: list, view_scene: list, stream: int, frames: int = 3,
                 view_slot: list | None = None) -> None:
        sarr, varr, iarr, slots = self._batch_arrays(scenes, views, view_scene, view_slot)
        with self.lock:
            rc = self.lib.fhtgs_autotune(self.ptr, sarr, len(scenes), varr, iarr, slots, len(views),
                                         int(frames), ctypes.c_void_p(stream))
            self.check(rc, "fhtgs_autotune")

    def visualize_gaze(self, image_ptr: int, width: int, height: int, gaze: tuple, rgb: tuple, vis_type: int,
                       size: int, to_chw: bool, stream: int) -> None:
        col = (ctypes.c_float * 3)(*[float(c) for c in rgb])
        with self.lock:
            self.check(self.lib.fhtgs_visualize_gaze(self.ptr, ctypes.c_void_p(image_ptr), int(width), int(height),
                                                     float(gaze[0]), float(gaze[1]), col, int(vis_type), int(size),
                                                     1 if to_chw else 0, ctypes.c_void_p(stream)),
                       "fhtgs_visualize_gaze")

    def visualize_tile_boundaries(self, image_ptr: int, width: int, height: int, mask_ptr: int, gaze: tuple,
                                  foveation_radius_px: int, blend_width_px: int, to_chw: bool, stream: int) -> None:
        with self.lock:
            self.check(self.lib.fhtgs_visualize_tile_boundaries(
                self.ptr, ctypes.c_void_p(image_ptr), int(width), int(height), ctypes.c_void_p(mask_ptr),
                float(gaze[0]), float(gaze[1]), int(foveation_radius_px), int(blend_width_px),
                1 if to_chw else 0, ctypes.c_void_p(stream)), "fhtgs_visualize_tile_boundaries")

    def build_tables(self, visible_ptr: int, width: int, height: int, foveation_radius_px: int,
                     blend_width_px: int, mask_ptr: int, vis_sat_ptr: int, fovea_sat_ptr: int, stream: int) -> None:
        with self.lock:
            self.check(self.lib.fhtgs_build_tables(
                self.ptr, ctypes.c_void_p(visible_ptr or None), int(width), int(height), int(foveation_radius_px),
                int(blend_width_px), ctypes.c_void_p(mask_ptr or None), ctypes.c_void_p(vis_sat_ptr or None),
                ctypes.c_void_p(fovea_sat_ptr or None), ctypes.c_void_p(stream)), "fhtgs_build_tables")

    def get_option(self, option: int) -> int:
        v = ctypes.c_int64()
        self.check(self.lib.fhtgs_get_option(self.ptr, option, ctypes.byref(v)), "fhtgs_get_option")
        return v.value

    def tuned_digit_bits(self, width: int, height: int) -> int:
        b = ctypes.c_int32()
        self.check(self.lib.fhtgs_tuned_digit_bits(self.ptr, int(width), int(height), ctypes.byref(b)),
                   "fhtgs_tuned_digit_bits")
        return b.value

    def tuned_work_tile(self, width: int, height: int) -> int:
        """Waves per blend workgroup the handle uses for this image size: 1 (8x8), 2 (16x8), 4 (16x16)."""
        b = ctypes.c_int32()
        self.check(self.lib.fhtgs_tuned_work_tile(self.ptr, int(width), int(height), ctypes.byref(b)),
                   "fhtgs_tuned_work_tile")
        return b.value

    def sort_pairs(self, keys_ptr: int, vals_ptr: int, n: int, key_bytes: int, end_bit: int,
                   stream: int) -> None:
        with self.lock:
            rc = self.lib.fhtgs_sort_pairs(self.ptr, ctypes.c_void_p(keys_ptr), ctypes.c_void_p(vals_ptr),
                                           int(n), int(key_bytes), int(end_bit), ctypes.c_void_p(stream))
            self.check(rc, "fhtgs_sort_pairs")

    def counts(self, view: int = 0) -> list:
        c = (ctypes.c_int64 * 16)()
        self.check(self.lib.fhtgs_debug_counts(self.ptr, int(view), c), "fhtgs_debug_counts")
        return list(c)

    def debug_read(self, view: int, name: str, nbytes: int) -> bytes:
        buf = ctypes.create_string_buffer(max(int(nbytes), 1))
        out = ctypes.c_int64()
        self.check(self.lib.fhtgs_debug_read(self.ptr, int(view), DBG[name], buf, int(nbytes),
                                             ctypes.byref(out)), f"fhtgs_debug_read({name})")
        return buf.raw[: out.value]

    def stage_times(self) -> dict:
        ms = (ctypes.c_float * NUM_STAGES)()
        self.check(self.lib.fhtgs_stage_times(self.ptr, ms), "fhtgs_stage_times")
        return dict(zip(STAGE_NAMES, list(ms)))

    def workspace_bytes(self) -> int:
        b = ctypes.c_int64()
        self.check(self.lib.fhtgs_workspace_bytes(self.ptr, ctypes.byref(b)), "fhtgs_workspace_bytes")
        return b.value

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and self.ptr.value:
                self.lib.fhtgs_destroy(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass


_handles: dict = {}
_handles_lock = threading.Lock()


def handle_for(device: int) -> Handle:
    """The process-wide handle of a device (the reference's entry points are
    free functions; this keeps their call shape while scratch is reused)."""
    with _handles_lock:
        h = _handles.get(device)
        if h is None:
            h = Handle(device)
            _handles[device] = h
        return h
