"""``gstex_cuda.texture_sample``: re-exports ``gstex_cuda_amd.texture_sample`` (the reference's
texture_sample.py interface on the gfx950 library)."""
from gstex_cuda_amd.texture_sample import *  # noqa: F401,F403
