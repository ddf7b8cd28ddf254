"""``gstex_cuda.get_aabb_2d``: re-exports ``gstex_cuda_amd.get_aabb_2d`` (the
reference's get_aabb_2d.py interface on the gfx950 library).  The reference's
pure-torch ``get_aabb_2d_torch`` (get_aabb_2d.py:34-54) is not provided."""
from gstex_cuda_amd.get_aabb_2d import *  # noqa: F401,F403

_ABSENT = {"get_aabb_2d_torch"}


def __getattr__(name):
    if name in _ABSENT:
        raise AttributeError(f"gstex_cuda.get_aabb_2d.{name} is the reference's pure-torch path, "
                             "which this package does not provide; use get_aabb_2d")
    raise AttributeError(f"module 'gstex_cuda.get_aabb_2d' has no attribute {name!r}")
