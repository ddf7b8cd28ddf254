"""The reference's import path, ``gstex_cuda``, served by ``gstex_cuda_amd``.

A caller written against victor-rong/GStex_cuda imports the same module and
function names from here unchanged (example.py:9-14: ``gstex_cuda.texture``,
``gstex_cuda.get_aabb_2d``, ``gstex_cuda._torch_impl``, ``gstex_cuda.timer``;
tests: ``gstex_cuda.sh``, ``gstex_cuda.texture_sample``).  Every module here
re-exports the gfx950 implementation in ``gstex_cuda_amd``; none has code of
its own on the render path, and none falls back to a CPU or torch path.

Provided: ``cuda`` (the native entry points), ``texture``, ``get_aabb_2d``,
``utils``, ``sh``, ``texture_sample``, ``texture_edit``, ``timer`` and, from
``_torch_impl``, the rotation helpers ``normalized_quat_to_rotmat`` /
``quat_to_rotmat`` and ``project_pix``.  The reference's pure-torch render
path (``texture_gaussians(..., use_torch_impl=True)``,
``_torch_impl.texture_forward``, ``sample_texture``, ``compute_sh_color``,
``eval_sh_bases``, ``get_aabb_2d.get_aabb_2d_torch``) is not: those names raise
with a message saying so.
"""
from gstex_cuda_amd import __version__  # noqa: F401
