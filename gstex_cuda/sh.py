"""``gstex_cuda.sh``: re-exports ``gstex_cuda_amd.sh`` (the reference's
sh.py interface on the gfx950 library)."""
from gstex_cuda_amd.sh import *  # noqa: F401,F403
