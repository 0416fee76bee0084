"""Operator surface of the rasterizer, with the reference's classes, fields
and call shapes (torch_bindings/rasterization.py:9-154 of
FoveatedHTGSCudaBackend). ``render``, ``benchmark`` and ``render_stereo``
dispatch to the gfx950 library through ``_C``, ``_C_benchmarking`` and
``_C_stereo``.
"""
from __future__ import annotations

from enum import Enum
from typing import NamedTuple

import torch

from .. import _C, _C_benchmarking, _C_stereo, _C_visualization
from .._lib import DEFAULT_BLEND_WIDTH_PX, DEFAULT_FOVEATION_RADIUS_PX

__all__ = ["BackgroundModel", "GazeVisualizationType", "RasterizerSettings", "FoveatedHTGSRasterizer",
           "FoveatedHTGSVisualizer"]


class BackgroundModel(Enum):
    NONE = 0
    SH = 1
    TEXTURE = 2
    TEXTURE_RAW = 3  # texture storing pre-activation values, activated after interpolation


class GazeVisualizationType(Enum):
    CIRCLE = 0
    SQUARE = 1
    CROSS = 2


class RasterizerSettings(NamedTuple):
    M: torch.Tensor  # affine transformation from world space to view space (rows 0..2 used)
    VPM: torch.Tensor  # homogeneous transformation from world space to screen space
    VPR_inv: torch.Tensor  # screen space to world directions (no translation)
    cam_position: torch.Tensor  # camera position in world space
    gaze_position: torch.Tensor  # gaze position in screen space (read on the host)
    visibility_mask: torch.Tensor  # bitset over 16x16 tiles, uint32 words
    visibility_mask_area_table: torch.Tensor  # summed-area table of the mask
    fovea_mask_area_table: torch.Tensor  # summed-area table of the fovea disk (x3)
    background_model_data: torch.Tensor  # background model specific data
    background_model: BackgroundModel
    K: int  # size of the core for hybrid transparency
    active_sh_bases: int  # number of spherical harmonics bases used for colour
    width: int
    height: int
    focal_x: float
    focal_y: float
    center_x: float
    center_y: float
    near_plane: float
    far_plane: float
    scale_modifier: float  # scaling factor applied to each Gaussian

    def as_tuple(self) -> tuple:
        return (
            self.M,
            self.VPM,
            self.VPR_inv,
            self.cam_position,
            self.gaze_position,
            self.visibility_mask,
            self.visibility_mask_area_table,
            self.fovea_mask_area_table,
            self.background_model_data,
            self.background_model.value,
            self.K,
            self.active_sh_bases,
            self.width,
            self.height,
            self.focal_x,
            self.focal_y,
            self.center_x,
            self.center_y,
            self.near_plane,
            self.far_plane,
            self.scale_modifier,
        )


class FoveatedHTGSRasterizer(torch.nn.Module):
    """``foveation_radius_px`` / ``blend_width_px`` default to the reference's
    compile-time constants (375 / 125 px); ``foveation_radius_px=-1`` renders
    every visible tile at full resolution. ``quarter_radius_px`` > 0 (an
    extension, default off) renders tiles beyond that radius at quarter
    resolution."""

    def __init__(self, foveation_radius_px: int = DEFAULT_FOVEATION_RADIUS_PX,
                 blend_width_px: int = DEFAULT_BLEND_WIDTH_PX, quarter_radius_px: int = 0):
        super().__init__()
        self.foveation_radius_px = int(foveation_radius_px)
        self.blend_width_px = int(blend_width_px)
        self.quarter_radius_px = int(quarter_radius_px)

    def _fov(self) -> dict:
        return {"foveation_radius_px": self.foveation_radius_px, "blend_width_px": self.blend_width_px,
                "quarter_radius_px": self.quarter_radius_px}

    def render(
            self,
            positions: torch.Tensor,
            scales: torch.Tensor,
            rotations: torch.Tensor,
            opacities: torch.Tensor,
            sh_0: torch.Tensor,
            sh_rest: torch.Tensor,
            settings: RasterizerSettings,
            to_chw: bool,
            blur_periphery: bool,
            aaa_mode: bool,
    ) -> torch.Tensor:
        return _C.render(positions, scales, rotations, opacities, sh_0, sh_rest,
                         *settings.as_tuple(), to_chw, blur_periphery, aaa_mode, **self._fov())

    def benchmark(
            self,
            positions: torch.Tensor,
            scales: torch.Tensor,
            rotations: torch.Tensor,
            opacities: torch.Tensor,
            sh_0: torch.Tensor,
            sh_rest: torch.Tensor,
            settings: RasterizerSettings,
            to_chw: bool,
            blur_periphery: bool,
            aaa_mode: bool,
    ) -> torch.Tensor:
        return _C_benchmarking.benchmark(positions, scales, rotations, opacities, sh_0, sh_rest,
                                         *settings.as_tuple(), to_chw, blur_periphery, aaa_mode,
                                         **self._fov())

    def render_stereo(
            self,
            positions: torch.Tensor,
            scales: torch.Tensor,
            rotations: torch.Tensor,
            opacities: torch.Tensor,
            sh_0: torch.Tensor,
            sh_rest: torch.Tensor,
            settings_left: RasterizerSettings,
            settings_right: RasterizerSettings,
            to_chw: bool,
            blur_periphery: bool,
            aaa_mode: bool,
    ) -> tuple[torch.Tensor, torch.Tensor]:
        return _C_stereo.render(positions, scales, rotations, opacities, sh_0, sh_rest,
                                *settings_left.as_tuple(), *settings_right.as_tuple(),
                                to_chw, blur_periphery, aaa_mode, **self._fov())


class FoveatedHTGSVisualizer(torch.nn.Module):
    """Debug overlays (torch_bindings/rasterization.py:157-192 of the
    reference): the gaze marker and the tile-type borders, drawn in place.
    The foveation radii default to the rasterizer's (375 / 125 px)."""

    def __init__(self, foveation_radius_px: int = DEFAULT_FOVEATION_RADIUS_PX,
                 blend_width_px: int = DEFAULT_BLEND_WIDTH_PX):
        super().__init__()
        self.foveation_radius_px = int(foveation_radius_px)
        self.blend_width_px = int(blend_width_px)

    def visualize_gaze(
            self,
            image: torch.Tensor,
            gaze_position: torch.Tensor,
            gaze_color: torch.Tensor,
            visualization_size: int,
            visualization_type: GazeVisualizationType,
            to_chw: bool,
    ) -> torch.Tensor:
        return _C_visualization.visualize_gaze(image, gaze_position, gaze_color, visualization_type.value,
                                               visualization_size, to_chw)

    def visualize_tile_boundaries(
            self,
            image: torch.Tensor,
            visibility_mask: torch.Tensor,
            gaze_position: torch.Tensor,
            to_chw: bool,
    ) -> torch.Tensor:
        return _C_visualization.visualize_tile_boundaries(image, visibility_mask, gaze_position, to_chw,
                                                          foveation_radius_px=self.foveation_radius_px,
                                                          blend_width_px=self.blend_width_px)
