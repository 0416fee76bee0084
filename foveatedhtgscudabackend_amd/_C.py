"""``_C.render``: the mono render entry point (torch_bindings/bindings.cpp:8,
rasterization/include/inference_api.h:8-38). Same 30 positional arguments;
returns a float32 (3,H,W) or (H,W,3) image in which pixels no tile writes are 0.

Two keyword-only extensions make the reference's compile-time foveation
radii (rasterization/include/config.h:30-32) runtime values:
``foveation_radius_px`` (default 375, ``-1`` = foveation off) and
``blend_width_px`` (default 125). A third, ``quarter_radius_px`` (default 0 =
off), adds a quarter-resolution level beyond that radius (no reference
counterpart).
"""
from ._binding import render_views
from ._lib import DEFAULT_BLEND_WIDTH_PX, DEFAULT_FOVEATION_RADIUS_PX

__all__ = ["render"]


def render(positions, scales, rotations, opacities, sh_0, sh_rest,
           M, VPM, VPR_inv, cam_position, gaze_position, visibility_mask,
           visibility_mask_area_table, fovea_mask_area_table, background_model_data,
           background_model_type, K, active_sh_bases, width, height, focal_x, focal_y,
           center_x, center_y, near_plane, far_plane, scale_modifier,
           to_chw, blur_periphery, anti_aliasing, *,
           foveation_radius_px=DEFAULT_FOVEATION_RADIUS_PX, blend_width_px=DEFAULT_BLEND_WIDTH_PX,
           quarter_radius_px=0):
    view = (M, VPM, VPR_inv, cam_position, gaze_position, visibility_mask,
            visibility_mask_area_table, fovea_mask_area_table, background_model_data,
            background_model_type, K, active_sh_bases, width, height, focal_x, focal_y,
            center_x, center_y, near_plane, far_plane, scale_modifier)
    (image,) = render_views((positions, scales, rotations, opacities, sh_0, sh_rest), [view],
                            False, True, to_chw, blur_periphery, anti_aliasing,
                            foveation_radius_px, blend_width_px, quarter_radius_px)
    return image
