"""The reference's import name, ``FoveatedHTGSCudaBackend``, served by this build.

Callers written against the reference import
``from FoveatedHTGSCudaBackend import _C, _C_benchmarking, _C_stereo, _C_visualization``
(torch_bindings/rasterization.py:6) and
``FoveatedHTGSCudaBackend.torch_bindings.rasterization`` (the plugin shim,
__init__.py:13-18). Every name here is the object of
``foveatedhtgscudabackend_amd`` itself (an alias, not a second implementation),
so those lines resolve unchanged to the gfx950 library.
"""
from __future__ import annotations

import os
import sys

try:
    import foveatedhtgscudabackend_amd as _impl
except ImportError:  # used from a checkout whose root is not on sys.path
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))
    import foveatedhtgscudabackend_amd as _impl

from . import _C, _C_benchmarking, _C_stereo, _C_visualization  # noqa: E402
from .torch_bindings.rasterization import (BackgroundModel, FoveatedHTGSRasterizer,  # noqa: E402
                                           FoveatedHTGSVisualizer, GazeVisualizationType,
                                           RasterizerSettings)

__all__ = ["_C", "_C_benchmarking", "_C_stereo", "_C_visualization", "FoveatedHTGSRasterizer",
           "FoveatedHTGSVisualizer", "RasterizerSettings", "BackgroundModel", "GazeVisualizationType"]
