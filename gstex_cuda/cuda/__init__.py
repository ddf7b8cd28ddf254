"""``gstex_cuda.cuda``: the native entry points (cuda/__init__.py:14-26 of the
reference), served by the gfx950 library through ``gstex_cuda_amd.cuda``."""
from gstex_cuda_amd.cuda import *  # noqa: F401,F403
from gstex_cuda_amd.cuda import __all__  # noqa: F401
