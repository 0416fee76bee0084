"""``FoveatedHTGSCudaBackend.torch_bindings.rasterization``: this build's operator
surface (``foveatedhtgscudabackend_amd.torch_bindings.rasterization``)."""
from foveatedhtgscudabackend_amd.torch_bindings.rasterization import *  # noqa: F401,F403
from foveatedhtgscudabackend_amd.torch_bindings.rasterization import __all__  # noqa: F401
