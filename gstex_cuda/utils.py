"""``gstex_cuda.utils``: re-exports ``gstex_cuda_amd.utils`` (the reference's
utils.py interface on the gfx950 library)."""
from gstex_cuda_amd.utils import *  # noqa: F401,F403
