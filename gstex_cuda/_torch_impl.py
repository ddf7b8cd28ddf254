"""``gstex_cuda._torch_impl``: the caller-side torch helpers of the reference's
_torch_impl.py that its trainer imports (example.py:11, _torch_impl.py:116-147).

Provided, written from the formulas:

* ``normalized_quat_to_rotmat(q)`` -- rotation matrix (..., 3, 3) of unit
  quaternions (w, x, y, z); its columns are the splat axes;
* ``quat_to_rotmat(q)`` -- the same after normalising q;
* ``project_pix(fxfy, p_view, center)`` -- pinhole pixel coordinates
  (x / (z + 1e-6)) f + c.

Not provided: the reference's pure-torch render path (``texture_forward``,
``sample_texture``, ``compute_sh_color``, ``eval_sh_bases``).  Rendering,
sampling and SH run only on the gfx950 library (``gstex_cuda.texture``,
``gstex_cuda.texture_sample``, ``gstex_cuda.sh``); those names raise
AttributeError with a message saying so.
"""
from __future__ import annotations

import torch
from torch import Tensor

_ABSENT = {"texture_forward", "sample_texture", "compute_sh_color", "eval_sh_bases"}


def normalized_quat_to_rotmat(quat: Tensor) -> Tensor:
    """R(q) for unit q = (w, x, y, z), shape (..., 3, 3)."""
    if quat.shape[-1] != 4:
        raise AssertionError(quat.shape)
    w, x, y, z = quat.unbind(-1)
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    rows = (
        (1 - 2 * (yy + zz), 2 * (xy - wz), 2 * (xz + wy)),
        (2 * (xy + wz), 1 - 2 * (xx + zz), 2 * (yz - wx)),
        (2 * (xz - wy), 2 * (yz + wx), 1 - 2 * (xx + yy)),
    )
    return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)


def quat_to_rotmat(quat: Tensor) -> Tensor:
    """R(q / |q|), shape (..., 3, 3)."""
    if quat.shape[-1] != 4:
        raise AssertionError(quat.shape)
    return normalized_quat_to_rotmat(torch.nn.functional.normalize(quat, dim=-1))


def project_pix(fxfy, p_view: Tensor, center, eps: float = 1e-6) -> Tensor:
    """Pixel coordinates (..., 2) of view-space points (..., 3)."""
    fx, fy = fxfy
    cx, cy = center
    rw = 1.0 / (p_view[..., 2] + 1e-6)
    return torch.stack([p_view[..., 0] * rw * fx + cx, p_view[..., 1] * rw * fy + cy], dim=-1)


def __getattr__(name):
    if name in _ABSENT:
        raise AttributeError(
            f"gstex_cuda._torch_impl.{name} is the reference's pure-torch render path, which "
            "this package does not provide: rendering, sampling and SH run on the gfx950 "
            "library only")
    raise AttributeError(f"module 'gstex_cuda._torch_impl' has no attribute {name!r}")
