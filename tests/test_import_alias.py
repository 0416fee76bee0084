"""The reference's import lines resolve unchanged to this build (no GPU needed):
``from FoveatedHTGSCudaBackend import _C, _C_benchmarking, _C_stereo,
_C_visualization`` (torch_bindings/rasterization.py:6) and the plugin shim's
``.FoveatedHTGSCudaBackend.torch_bindings.rasterization`` relative to the
directory holding the package (__init__.py:13-18)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_import_lines_give_this_builds_objects():
    from FoveatedHTGSCudaBackend import _C, _C_benchmarking, _C_stereo, _C_visualization
    from FoveatedHTGSCudaBackend.torch_bindings.rasterization import (BackgroundModel, FoveatedHTGSRasterizer,
                                                                      FoveatedHTGSVisualizer, GazeVisualizationType,
                                                                      RasterizerSettings)
    import foveatedhtgscudabackend_amd as impl
    from foveatedhtgscudabackend_amd import _C as iC, _C_benchmarking as iB, _C_stereo as iS, _C_visualization as iV
    from foveatedhtgscudabackend_amd.torch_bindings import rasterization as iR
    assert _C.render is iC.render and _C_stereo.render is iS.render
    assert _C_benchmarking.benchmark is iB.benchmark
    assert _C_visualization.visualize_gaze is iV.visualize_gaze
    assert _C_visualization.visualize_tile_boundaries is iV.visualize_tile_boundaries
    assert FoveatedHTGSRasterizer is iR.FoveatedHTGSRasterizer is impl.FoveatedHTGSRasterizer
    assert FoveatedHTGSVisualizer is iR.FoveatedHTGSVisualizer
    assert RasterizerSettings is iR.RasterizerSettings and BackgroundModel is iR.BackgroundModel
    assert GazeVisualizationType is iR.GazeVisualizationType
    # the 21 settings fields in the reference's order (torch_bindings/rasterization.py:22-43)
    assert len(RasterizerSettings._fields) == 21 and RasterizerSettings._fields[0] == "M"


def test_plugin_shim_relative_import_resolves(tmp_path):
    """A plugin directory that holds the package (as the reference's root
    __init__.py does) imports it relatively; only the repository root's parent
    is on sys.path."""
    plugin = tmp_path / "htgs_plugin"
    plugin.mkdir()
    os.symlink(os.path.join(ROOT, "FoveatedHTGSCudaBackend"), plugin / "FoveatedHTGSCudaBackend")
    (plugin / "__init__.py").write_text(
        "from .FoveatedHTGSCudaBackend.torch_bindings.rasterization import (\n"
        "    FoveatedHTGSRasterizer, FoveatedHTGSVisualizer, RasterizerSettings)\n")
    code = ("import htgs_plugin, foveatedhtgscudabackend_amd as impl; "
            "assert htgs_plugin.FoveatedHTGSRasterizer is impl.FoveatedHTGSRasterizer; print('ok')")
    env = dict(os.environ, PYTHONPATH=str(tmp_path))
    r = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), env=env, capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
