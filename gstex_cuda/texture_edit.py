"""``gstex_cuda.texture_edit``: re-exports ``gstex_cuda_amd.texture_edit`` (the reference's
texture_edit.py interface on the gfx950 library)."""
from gstex_cuda_amd.texture_edit import *  # noqa: F401,F403
