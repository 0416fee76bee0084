"""C ABI checks that need no GPU: the library loads, exports every entry
point include/fhtgs.h declares, and the ctypes mirrors match the C layout."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fhtgs.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fhtgs_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from foveatedhtgscudabackend_amd import _lib
    lib = _lib.load()
    declared = _declared_functions()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(lib, name), f"libfhtgs.so does not export {name}"
    assert set(declared) == set(_lib.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fhtgs_\w+)", out))
    assert set(declared) <= exported


def test_struct_layout_matches_header(tmp_path):
    import ctypes
    from foveatedhtgscudabackend_amd import _lib
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fhtgs.h"\n'
        "int main(void){\n"
        ' printf("%zu %zu\\n", sizeof(fhtgs_view), sizeof(fhtgs_scene));\n'
        + "".join(f' printf("%zu\\n", offsetof(fhtgs_view, {f}));\n' for f, _ in _lib.View._fields_)
        + "".join(f' printf("%zu\\n", offsetof(fhtgs_scene, {f}));\n' for f, _ in _lib.Scene._fields_)
        + " return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(_lib.View) and vals[1] == ctypes.sizeof(_lib.Scene)
    offs = [getattr(_lib.View, f).offset for f, _ in _lib.View._fields_] + \
           [getattr(_lib.Scene, f).offset for f, _ in _lib.Scene._fields_]
    assert vals[2:] == offs


def test_fovea_radius_rule_matches_table_builder():
    from foveatedhtgscudabackend_amd import _lib, tables
    for w, h, fr, bw in ((1920, 1080, 375, 125), (2064, 2208, 232, 125), (800, 800, -1, 125), (64, 48, 100, 40)):
        R, B = _lib.fovea_radius_tiles(w, h, fr, bw)
        assert R == tables.fovea_radius_tiles_for(w, h, fr, bw)
        if fr >= 0:
            assert R == (fr + 15) // 16 and B == max(0, fr - bw) // 16
        else:
            gw, gh = tables.grid_size(w, h)
            # the gaze tile is in [0, gw] x [0, gh] and tile centres in
            # [0.5, gw - 0.5] x [0.5, gh - 0.5]: every centre is within R
            assert R == B and R * R > gw ** 2 + gh ** 2


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from foveatedhtgscudabackend_amd import _lib
    with pytest.raises(RuntimeError):
        _lib.Handle(0)


def test_binding_rejects_bad_shapes_before_any_gpu_call():
    import torch
    from foveatedhtgscudabackend_amd import _binding
    keep = []
    with pytest.raises(ValueError):
        _binding.scene_struct(torch.zeros(5, 2), torch.zeros(5, 3), torch.zeros(5, 4), torch.zeros(5),
                              torch.zeros(5, 1, 3), torch.zeros(5, 15, 3), keep=keep)
    with pytest.raises(ValueError):
        _binding.scene_struct(torch.zeros(5, 3), torch.zeros(5, 3), torch.zeros(4, 4), torch.zeros(5),
                              torch.zeros(5, 1, 3), torch.zeros(5, 15, 3), keep=keep)
    with pytest.raises(ValueError):  # CPU tensors are rejected: no CPU fallback
        _binding.scene_struct(torch.zeros(5, 3), torch.zeros(5, 3), torch.zeros(5, 4), torch.zeros(5),
                              torch.zeros(5, 1, 3), torch.zeros(5, 15, 3), keep=keep)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def _valid_view(**kw):
    from foveatedhtgscudabackend_amd import _lib
    v = dict(M=8, VPM=8, VPR_inv=8, cam_position=8, image=8, visibility_mask=8,
             visibility_mask_area_table=8, fovea_mask_area_table=8, K=8, active_sh_bases=16,
             width=2064, height=2208, focal_x=866.0, focal_y=866.0, center_x=1032.0, center_y=1104.0,
             near_plane=0.2, far_plane=1000.0, scale_modifier=1.0, foveation_radius_px=232,
             blend_width_px=77, quarter_radius_px=606)
    v.update(kw)
    return _lib.View(**v)


@pytest.mark.parametrize("kw,code,why", [
    ({}, 0, ""),
    ({"foveation_radius_px": -1, "blend_width_px": 0, "quarter_radius_px": 0}, 0, ""),
    ({"blend_width_px": 16}, -1, "blend width"),
    ({"blend_width_px": -5}, -1, "blend width"),
    ({"blend_width_px": 300}, -1, "exceeds the foveation radius"),
    ({"width": 4096, "height": 4096}, 0, ""),
    ({"width": 4097, "height": 4096}, -5, "65536"),
    ({"width": 8192, "height": 300}, 0, ""),
    ({"width": 8193, "height": 8192}, -5, "65536"),
    ({"quarter_radius_px": 100}, -1, "quarter radius"),
    ({"image": None}, -1, "null"),
])
def test_view_checks_run_on_the_host_before_any_enqueue(kw, code, why):
    """fhtgs_validate_view is the check fhtgs_render runs first; it needs no
    device. Images above 65536 16-px tiles fail here with FHTGS_E_UNSUPPORTED,
    not later inside a frame."""
    from foveatedhtgscudabackend_amd import _lib
    rc, msg = _lib.validate_view(_valid_view(**kw))
    assert rc == code, msg
    assert why in msg


def test_workspace_query_sizes_a_slot_from_counts_on_the_host():
    """fhtgs_workspace_query gives a slot's bytes from the view's geometry and
    the element counts, before anything exists on the device (pointers may be
    NULL): the per-Gaussian part grows by the buffers the preprocess writes
    (tiles touched 4, bounds 8, record row 64, ellipse record as one 32-B
    entry, block sums 4/256 B), the instance part by double-buffered 32-bit
    keys and values."""
    from foveatedhtgscudabackend_amd import _lib
    geo = dict(K=8, width=2064, height=2208, foveation_radius_px=232, blend_width_px=77, quarter_radius_px=606)
    v = _lib.View(**geo)
    f0, i0 = _lib.workspace_query(v, 0, 0)
    f1, i1 = _lib.workspace_query(v, 3_000_000, 4_800_000)
    assert 0 < f0 < f1 and 0 < i0 < i1
    per_gauss = (f1 - f0) / 3_000_000
    assert 4 + 8 + 64 + 32 <= per_gauss <= 4 + 8 + 64 + 32 + 1, per_gauss
    per_inst = (i1 - i0) / 4_800_000
    assert 16 <= per_inst <= 17, per_inst  # 2 x (4-B key + 4-B value), plus look-back status
    # the tile part follows the image size; tables brought by the caller
    # need no device-built copies
    f_small, _ = _lib.workspace_query(_lib.View(**dict(geo, width=640, height=480, quarter_radius_px=0)), 0, 0)
    assert f_small < f0
    f_tabs, _ = _lib.workspace_query(_lib.View(**geo, visibility_mask=8, visibility_mask_area_table=8,
                                               fovea_mask_area_table=8), 0, 0)
    assert f_tabs < f0
    # geometry the library rejects is rejected here too
    with pytest.raises(ValueError, match="65536"):
        _lib.workspace_query(_lib.View(**dict(geo, width=8193, height=8192)), 10, 10)
    with pytest.raises(ValueError, match="blend width"):
        _lib.workspace_query(_lib.View(**dict(geo, blend_width_px=16)), 10, 10)
    with pytest.raises(ValueError):
        _lib.workspace_query(v, -1, 0)
