"""MI355X (gfx950) foveated hybrid-transparency Gaussian-splatting rasterizer.

Drop-in for the inference path of FoveatedHTGSCudaBackend: the same
``FoveatedHTGSRasterizer.render / benchmark / render_stereo`` calls and
``RasterizerSettings`` fields, backed by hand-written HIP kernels in
``libfhtgs.so`` (C ABI: include/fhtgs.h).
"""
from .torch_bindings.rasterization import (BackgroundModel, FoveatedHTGSRasterizer, FoveatedHTGSVisualizer,
                                           GazeVisualizationType, RasterizerSettings)

__all__ = ["FoveatedHTGSRasterizer", "FoveatedHTGSVisualizer", "RasterizerSettings", "BackgroundModel",
           "GazeVisualizationType"]
