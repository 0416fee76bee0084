"""``FoveatedHTGSCudaBackend._C``: this build's ``foveatedhtgscudabackend_amd._C``."""
from foveatedhtgscudabackend_amd._C import *  # noqa: F401,F403
from foveatedhtgscudabackend_amd._C import __all__  # noqa: F401
