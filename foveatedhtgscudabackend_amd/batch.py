"""Batched rendering: many views of several resident scenes in one call.

The reference renders one view (``_C.render``) or one stereo pair
(``_C_stereo.render``) per call. ``render_batch`` hands a whole batch to
``fhtgs_render_batch``: up to four views of one scene that agree on the
preprocess parameters share one read of the Gaussians (stereo pairs, or any
such views of the batch), the
frame is enqueued without a mid-frame host wait, and repeated batches of the
same shape replay one captured hipGraph. Every image equals what
``FoveatedHTGSRasterizer.render`` returns for that view alone.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _binding, _lib
from .torch_bindings.rasterization import RasterizerSettings


@dataclass
class BatchView:
    """One view of a batch: which scene, its camera settings and render flags."""
    scene: int
    settings: RasterizerSettings
    to_chw: bool = False
    blur_periphery: bool = False
    anti_aliasing: bool = False
    foveation_radius_px: int = _lib.DEFAULT_FOVEATION_RADIUS_PX
    blend_width_px: int = _lib.DEFAULT_BLEND_WIDTH_PX
    clear_output: bool = True
    quarter_radius_px: int = 0  # extension: quarter-resolution level beyond this radius, 0 = off
    slot: int | None = None  # handle slot (persistent buffers), default: position in the batch


def _marshal(scenes: list, views: list):
    if not scenes or not views:
        raise ValueError("a batch needs at least one scene and one view")
    keep: list = []
    structs, device = [], None
    for sc in scenes:
        s, dev = _binding.scene_struct(*sc, keep=keep)
        if device is not None and dev != device:
            raise ValueError("all scenes of a batch must be on one device")
        device = dev
        structs.append(s)
    vstructs, images, index = [], [], []
    for bv in views:
        if not 0 <= bv.scene < len(scenes):
            raise ValueError(f"view refers to scene {bv.scene} of {len(scenes)}")
        v, img = _binding.view_struct(device, *bv.settings.as_tuple(), bv.to_chw, bv.blur_periphery,
                                      bv.anti_aliasing, bv.foveation_radius_px, bv.blend_width_px,
                                      bv.clear_output, keep=keep, quarter_radius_px=bv.quarter_radius_px)
        vstructs.append(v)
        images.append(img)
        index.append(bv.scene)
    given = [bv.slot is not None for bv in views]
    if any(given) and not all(given):
        raise ValueError("give a slot for every view of the batch or for none")
    slots = [bv.slot for bv in views] if all(given) else None
    dev_index = device.index if device.index is not None else torch.cuda.current_device()
    handle = _lib.handle_for(dev_index)
    stream = _binding._current_stream(dev_index)
    return handle, structs, vstructs, index, slots, images, stream, keep


def render_batch(scenes: list, views: list) -> list:
    """scenes: list of (positions, scales, rotations, opacities, sh_0, sh_rest)
    tuples on one GPU; views: list of BatchView. Returns one image per view
    ((H,W,3) or (3,H,W) float32), rendered on the current stream."""
    handle, structs, vstructs, index, slots, images, stream, keep = _marshal(scenes, views)
    handle.render_batch(structs, vstructs, index, stream, slots)
    del keep
    return images


def autotune(scenes: list, views: list, frames: int = 3) -> dict:
    """Tunes the handle for this batch (per view size the sort digit width and
    the blend work tile, 8x8 / 16x8 / 16x16; for the batch the blend XCD run
    and the concurrent view branches; images are unchanged by all of them)
    and returns the choices."""
    handle, structs, vstructs, index, slots, images, stream, keep = _marshal(scenes, views)
    handle.autotune(structs, vstructs, index, stream, frames, slots)
    del keep
    sizes = sorted({(bv.settings.width, bv.settings.height) for bv in views})
    names = {1: "8x8", 2: "16x8", 4: "16x16"}
    return {"digit_bits": {f"{w}x{h}": handle.tuned_digit_bits(w, h) for w, h in sizes},
            "work_tile": {f"{w}x{h}": names[handle.tuned_work_tile(w, h)] for w, h in sizes},
            "blend_xcd_run": handle.get_option(_lib.OPT_BLEND_XCD_RUN),
            "view_branches": handle.get_option(_lib.OPT_VIEW_BRANCHES)}
