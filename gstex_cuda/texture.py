"""``gstex_cuda.texture``: re-exports ``gstex_cuda_amd.texture`` (the reference's
texture.py interface on the gfx950 library)."""
from gstex_cuda_amd.texture import *  # noqa: F401,F403
