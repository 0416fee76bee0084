"""``_C_stereo.render`` (torch_bindings/bindings_stereo.cpp:8,
rasterization/include/inference_stereo_api.h:7-61): 6 scene tensors, 21
left-eye settings, 21 right-eye settings, then to_chw, blur_periphery and
anti_aliasing; returns ``(left, right)``.

As in rasterization/src/inference_stereo_api.cu:111-120,180-186 the left
eye's K, active_sh_bases, near/far planes, scale modifier and background
model are used for both eyes. Unlike the reference (two preprocess launches,
inference_stereo.cu:304-373) the scene is read once for both eyes; each
eye's arithmetic is the same as a mono render of that eye. Outputs are
allocated uninitialised like the reference's (``torch::empty``).
"""
from ._binding import render_views
from ._lib import DEFAULT_BLEND_WIDTH_PX, DEFAULT_FOVEATION_RADIUS_PX

__all__ = ["render"]


def render(positions, scales, rotations, opacities, sh_0, sh_rest, *args,
           foveation_radius_px=DEFAULT_FOVEATION_RADIUS_PX, blend_width_px=DEFAULT_BLEND_WIDTH_PX,
           clear_output=False, quarter_radius_px=0):
    if len(args) != 45:
        raise TypeError(f"_C_stereo.render expects 51 positional arguments, got {6 + len(args)}")
    left, right = list(args[0:21]), list(args[21:42])
    to_chw, blur_periphery, anti_aliasing = args[42:45]
    # left-eye values for both eyes: background data/type (8, 9), K (10),
    # active_sh_bases (11), near (18), far (19), scale_modifier (20)
    for k in (8, 9, 10, 11, 18, 19, 20):
        right[k] = left[k]
    images = render_views((positions, scales, rotations, opacities, sh_0, sh_rest),
                          [tuple(left), tuple(right)], True, clear_output, to_chw, blur_periphery,
                          anti_aliasing, foveation_radius_px, blend_width_px, quarter_radius_px)
    return images[0], images[1]
