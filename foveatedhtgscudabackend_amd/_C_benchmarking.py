"""``_C_benchmarking.benchmark`` (torch_bindings/bindings_benchmarking.cpp:7,
rasterization/include/fast_inference_api.h:8-38): same arguments and result
as ``_C.render`` except that the output is not cleared first
(fast_inference_api.cu:78-82 allocates it with ``torch::empty``), so pixels of
invisible tiles are undefined.
"""
from ._binding import render_views
from ._lib import DEFAULT_BLEND_WIDTH_PX, DEFAULT_FOVEATION_RADIUS_PX

__all__ = ["benchmark"]


def benchmark(positions, scales, rotations, opacities, sh_0, sh_rest,
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
                            False, False, to_chw, blur_periphery, anti_aliasing,
                            foveation_radius_px, blend_width_px, quarter_radius_px)
    return image
