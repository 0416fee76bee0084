"""``FoveatedHTGSCudaBackend._C_benchmarking``: this build's ``foveatedhtgscudabackend_amd._C_benchmarking``."""
from foveatedhtgscudabackend_amd._C_benchmarking import *  # noqa: F401,F403
from foveatedhtgscudabackend_amd._C_benchmarking import __all__  # noqa: F401
