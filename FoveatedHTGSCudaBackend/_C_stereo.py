"""``FoveatedHTGSCudaBackend._C_stereo``: this build's ``foveatedhtgscudabackend_amd._C_stereo``."""
from foveatedhtgscudabackend_amd._C_stereo import *  # noqa: F401,F403
from foveatedhtgscudabackend_amd._C_stereo import __all__  # noqa: F401
