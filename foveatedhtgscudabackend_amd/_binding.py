"""Tensor -> C-ABI marshalling shared by the ``_C``, ``_C_benchmarking`` and
``_C_stereo`` entry points.

Mirrors the reference torch wrappers (rasterization/src/inference_api.cu:13-122,
fast_inference_api.cu:12-121, inference_stereo_api.cu:13-192): the same
positional arguments, the same output allocation (render: pixels no tile
writes read 0; benchmark/stereo: left unwritten), work on the current stream.
"""
from __future__ import annotations

import torch

from . import _lib


def _dev_ptr(t: torch.Tensor, name: str, dtypes, device) -> tuple[int, torch.Tensor]:
    # called for every tensor of every frame: the cheapest checks first
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype not in dtypes:
        raise TypeError(f"{name} must have dtype in {dtypes}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if device is not None and t.get_device() != device.index:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if not t.is_contiguous():
        t = t.contiguous()
    p = t.data_ptr()
    if p % 16:
        t = t.clone()  # vector loads need 16-byte aligned rows
        p = t.data_ptr()
    return p, t


_F32 = (torch.float32,)

# the current stream's handle: torch's raw accessor where it exists (no
# Stream object per call), else the public one
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _current_stream(index: int) -> int:
    if _raw_stream is not None:
        return _raw_stream(index)
    return torch.cuda.current_stream(index).cuda_stream
_U32 = tuple(d for d in (getattr(torch, "uint32", None), torch.int32) if d is not None)


def scene_struct(positions, scales, rotations, opacities, sh_0, sh_rest, keep: list) -> tuple:
    device = positions.device
    n = positions.shape[0]
    if positions.dim() != 2 or positions.shape[1] != 3:
        raise ValueError(f"positions must be (N,3), got {tuple(positions.shape)}")
    if tuple(scales.shape) != (n, 3):
        raise ValueError(f"scales must be (N,3), got {tuple(scales.shape)}")
    if tuple(rotations.shape) != (n, 4):
        raise ValueError(f"rotations must be (N,4), got {tuple(rotations.shape)}")
    if opacities.numel() != n:
        raise ValueError(f"opacities must have N elements, got {tuple(opacities.shape)}")
    if sh_0.numel() != 3 * n:
        raise ValueError(f"sh_0 must be (N,1,3), got {tuple(sh_0.shape)}")
    if sh_rest.dim() != 3 or sh_rest.shape[0] != n or sh_rest.shape[2] != 3:
        raise ValueError(f"sh_rest must be (N,S,3), got {tuple(sh_rest.shape)}")
    kw = {"n_primitives": n, "total_sh_bases": int(sh_rest.shape[1])}
    for field, t in (("positions", positions), ("scales", scales), ("rotations", rotations),
                     ("opacities", opacities), ("sh_0", sh_0), ("sh_rest", sh_rest)):
        p, tc = _dev_ptr(t, field, _F32, device)
        keep.append(tc)
        kw[field] = p
    return _lib.Scene(**kw), device


def view_struct(device, M, VPM, VPR_inv, cam_position, gaze_position, visibility_mask,
                visibility_mask_area_table, fovea_mask_area_table, background_model_data,
                background_model_type, K, active_sh_bases, width, height, focal_x, focal_y,
                center_x, center_y, near_plane, far_plane, scale_modifier, to_chw,
                blur_periphery, anti_aliasing, foveation_radius_px, blend_width_px,
                clear_output, keep: list, quarter_radius_px: int = 0) -> tuple:
    kw: dict = {}  # View fields, set at once at the end (faster than one setattr each)
    width, height = int(width), int(height)
    if M.numel() < 12:
        raise ValueError("M must hold at least 3 rows of 4")
    if VPM.numel() != 16 or VPR_inv.numel() != 16:
        raise ValueError("VPM and VPR_inv must be 4x4")
    if cam_position.numel() != 3:
        raise ValueError("cam_position must have 3 elements")
    for field, t in (("M", M), ("VPM", VPM), ("VPR_inv", VPR_inv), ("cam_position", cam_position)):
        p, tc = _dev_ptr(t, field, _F32, device)
        keep.append(tc)
        kw[field] = p
    # the reference reads the gaze through a host pointer (inference.cu:79-82)
    if gaze_position.numel() != 2:
        raise ValueError("gaze_position must have 2 elements")
    g = gaze_position
    if g.is_cuda or g.dtype is not torch.float32:
        g = g.detach().to("cpu", torch.float32)
    kw["gaze_x"], kw["gaze_y"] = g.tolist() if g.dim() == 1 else g.reshape(-1).tolist()
    gw, gh = (width + 15) // 16, (height + 15) // 16
    # the library's own up-front checks (fhtgs_validate_view), before any allocation
    if gw * gh > _lib.MAX_LARGE_TILES:
        raise NotImplementedError(f"{width}x{height} has {gw * gh} 16-px tiles; at most "
                                  f"{_lib.MAX_LARGE_TILES} (4096x4096) are supported")
    if int(foveation_radius_px) >= 0 and int(blend_width_px) <= 16:
        raise ValueError("blend_width_px must exceed 16 with foveation on (the ring cross-fade "
                         "divides by blend_width_px - 16)")
    # extension: both summed-area tables None -> built on the device inside
    # the frame (and a None mask means every tile is visible)
    auto = visibility_mask_area_table is None and fovea_mask_area_table is None
    if (visibility_mask_area_table is None) != (fovea_mask_area_table is None):
        raise ValueError("give both summed-area tables or neither")
    if visibility_mask is None and not auto:
        raise ValueError("visibility_mask is required with explicit summed-area tables")
    if visibility_mask is not None and visibility_mask.numel() < (gw * gh + 31) // 32:
        raise ValueError("visibility_mask is smaller than ceil(tiles/32) words")
    if not auto:
        # summed-area tables are laid out for one image size and one foveation
        # radius: a table of another size would be read with the wrong stride
        if visibility_mask_area_table.numel() != (gw + 1) * (gh + 1):
            raise ValueError(f"visibility_mask_area_table must hold (gw+1)*(gh+1) = {(gw + 1) * (gh + 1)} "
                             f"entries for {width}x{height}, got {visibility_mask_area_table.numel()}")
        R, _ = _lib.fovea_radius_tiles(width, height, int(foveation_radius_px), int(blend_width_px))
        if fovea_mask_area_table.numel() != (2 * R + 1) ** 2:
            raise ValueError(f"fovea_mask_area_table must hold (2*{R}+1)^2 = {(2 * R + 1) ** 2} entries for "
                             f"this foveation radius, got {fovea_mask_area_table.numel()} (tables built for "
                             "another radius?)")
    for field, t in (("visibility_mask", visibility_mask),
                     ("visibility_mask_area_table", visibility_mask_area_table),
                     ("fovea_mask_area_table", fovea_mask_area_table)):
        if t is None:
            kw[field] = None
            continue
        p, tc = _dev_ptr(t, field, _U32, device)
        keep.append(tc)
        kw[field] = p
    bg_type = int(background_model_type)
    p, tc = _dev_ptr(background_model_data, "background_model_data", _F32, device)
    keep.append(tc)
    kw["background_model_data"] = p
    kw["background_model_type"] = bg_type
    if bg_type in (2, 3):
        if background_model_data.dim() != 3 or background_model_data.shape[0] != 3:
            raise ValueError("texture background must be (3,h,w)")
        kw["background_width"] = int(background_model_data.shape[2])
        kw["background_height"] = int(background_model_data.shape[1])
    elif bg_type == 1 and background_model_data.numel() < 48:
        raise ValueError("SH background needs 16x3 coefficients")
    kw["K"] = int(K)
    kw["active_sh_bases"] = int(active_sh_bases)
    kw["width"], kw["height"] = width, height
    kw["focal_x"], kw["focal_y"] = float(focal_x), float(focal_y)
    kw["center_x"], kw["center_y"] = float(center_x), float(center_y)
    kw["near_plane"], kw["far_plane"] = float(near_plane), float(far_plane)
    kw["scale_modifier"] = float(scale_modifier)
    kw["to_chw"] = 1 if to_chw else 0
    kw["blur_periphery"] = 1 if blur_periphery else 0
    kw["anti_aliasing"] = 1 if anti_aliasing else 0
    kw["foveation_radius_px"] = int(foveation_radius_px)
    kw["blend_width_px"] = int(blend_width_px)
    kw["quarter_radius_px"] = int(quarter_radius_px)
    kw["clear_output"] = 1 if clear_output else 0
    shape = (3, height, width) if to_chw else (height, width, 3)
    image = torch.empty(shape, dtype=torch.float32, device=device)
    kw["image"] = image.data_ptr()
    if blur_periphery:
        temp = torch.empty(shape, dtype=torch.float32, device=device)
        keep.append(temp)
        kw["image_temp"] = temp.data_ptr()
    return _lib.View(**kw), image


def render_views(scene_args: tuple, view_args: list, share_preprocess: bool, clear_output: bool,
                 to_chw, blur_periphery, anti_aliasing, foveation_radius_px, blend_width_px,
                 quarter_radius_px: int = 0):
    keep: list = []
    scene, device = scene_struct(*scene_args, keep=keep)
    views, images = [], []
    for va in view_args:
        v, img = view_struct(device, *va, to_chw, blur_periphery, anti_aliasing, foveation_radius_px,
                             blend_width_px, clear_output, keep=keep, quarter_radius_px=quarter_radius_px)
        views.append(v)
        images.append(img)
    index = device.index if device.index is not None else torch.cuda.current_device()
    handle = _lib.handle_for(index)
    stream = _current_stream(index)
    handle.render(scene, views, share_preprocess, stream)
    del keep
    return images
