"""``FoveatedHTGSCudaBackend._C_visualization``: this build's ``foveatedhtgscudabackend_amd._C_visualization``."""
from foveatedhtgscudabackend_amd._C_visualization import *  # noqa: F401,F403
from foveatedhtgscudabackend_amd._C_visualization import __all__  # noqa: F401
