"""``_C_visualization``: the debug-overlay entry points
(torch_bindings/bindings_visualization.cpp:8-9,
visualization/include/visualization_api.h:8-20). Both draw into ``image`` in
place on the current stream and return it, as the reference does.

``gaze_position`` and ``gaze_color`` are read on the host, as the reference
dereferences them there (visualization_api.cu:28-29, visualization.cu:36-40).
``visualize_tile_boundaries`` takes the keyword-only foveation radii of the
render entry points (default 375 / 125 px) so the overlay shows the
classification a render with the same radii uses.
"""
import torch

from . import _lib
from ._binding import _dev_ptr, _F32, _U32

__all__ = ["visualize_gaze", "visualize_tile_boundaries"]


def _image(image: torch.Tensor, to_chw: bool):
    if not isinstance(image, torch.Tensor) or image.dtype != torch.float32 or image.device.type != "cuda":
        raise TypeError("image must be a float32 GPU tensor")
    if image.dim() != 3 or (image.shape[0] if to_chw else image.shape[2]) != 3:
        raise ValueError(f"image must be {'(3,H,W)' if to_chw else '(H,W,3)'}, got {tuple(image.shape)}")
    if not image.is_contiguous():
        raise ValueError("image must be contiguous (it is drawn into in place)")
    height, width = (image.shape[1], image.shape[2]) if to_chw else (image.shape[0], image.shape[1])
    return int(width), int(height)


def _host2(t, name: str) -> tuple:
    v = torch.as_tensor(t).detach().to("cpu", torch.float32).reshape(-1)
    if v.numel() < 2:
        raise ValueError(f"{name} must have 2 elements")
    return float(v[0]), float(v[1])


def visualize_gaze(image, gaze_position, gaze_color, visualization_type, visualization_size, to_chw):
    width, height = _image(image, to_chw)
    col = torch.as_tensor(gaze_color).detach().to("cpu", torch.float32).reshape(-1)
    if col.numel() < 3:
        raise ValueError("gaze_color must have 3 elements")
    handle = _lib.handle_for(image.device.index if image.device.index is not None else torch.cuda.current_device())
    handle.visualize_gaze(image.data_ptr(), width, height, _host2(gaze_position, "gaze_position"),
                          tuple(float(c) for c in col[:3]), int(visualization_type), int(visualization_size),
                          bool(to_chw), torch.cuda.current_stream(image.device).cuda_stream)
    return image


def visualize_tile_boundaries(image, visibility_mask, gaze_position, to_chw, *,
                              foveation_radius_px=_lib.DEFAULT_FOVEATION_RADIUS_PX,
                              blend_width_px=_lib.DEFAULT_BLEND_WIDTH_PX):
    width, height = _image(image, to_chw)
    gw, gh = (width + 15) // 16, (height + 15) // 16
    if visibility_mask.numel() < (gw * gh + 31) // 32:
        raise ValueError("visibility_mask is smaller than ceil(tiles/32) words")
    mp, mask = _dev_ptr(visibility_mask, "visibility_mask", _U32, image.device)
    handle = _lib.handle_for(image.device.index if image.device.index is not None else torch.cuda.current_device())
    handle.visualize_tile_boundaries(image.data_ptr(), width, height, mp, _host2(gaze_position, "gaze_position"),
                                     int(foveation_radius_px), int(blend_width_px), bool(to_chw),
                                     torch.cuda.current_stream(image.device).cuda_stream)
    del mask
    return image
