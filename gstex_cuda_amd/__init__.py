"""gstex_cuda_amd — MI355X (gfx950) implementation of the GStex textured
2D-Gaussian rasteriser behind the ``gstex_cuda`` Python API.

Modules mirror the reference package: ``cuda`` (native entry points),
``texture`` (texture_gaussians), ``utils`` (binning), ``get_aabb_2d``,
``sh`` and ``texture_sample``.  Kernels live in ``libgstex_hip.so`` (built from
``csrc/`` by ``gstex_cuda_amd.build.build_library``) behind the C ABI in
``include/gstex_hip.h``.
"""

__version__ = "0.1.0"
