"""Drop-in for ``gstex_cuda.get_aabb_2d`` (get_aabb_2d.py:1-92): screen-space
AABBs, point projection and tile counts, all on the GPU."""
from __future__ import annotations

from torch import Tensor
import torch

from . import cuda as _C


def merge_aabbs(centers1: Tensor, extents1: Tensor, centers2: Tensor, extents2: Tensor):
    """Smallest AABB containing both boxes (get_aabb_2d.py:15-20)."""
    lo = torch.minimum(centers1 - extents1, centers2 - extents2)
    hi = torch.maximum(centers1 + extents1, centers2 + extents2)
    return 0.5 * (lo + hi), 0.5 * (hi - lo)


def project_points(points: Tensor, viewmat: Tensor, intrins, clip: bool = False):
    """Pixel coordinates and view depth of points (get_aabb_2d.py:22-32).
    Not differentiable (the depths only feed the tile sort)."""
    fx, fy, cx, cy = intrins
    return _C.project_points(points.contiguous().float(), viewmat.contiguous().float(), fx, fy,
                             cx, cy, clip=clip)


def get_aabb_2d(means: Tensor, scales: Tensor, glob_scale, quats: Tensor, viewmat: Tensor,
                intrins):
    """AABB of the +-3 sigma splat corners (get_aabb_2d.py:56-68)."""
    fx, fy, cx, cy = intrins
    return _C.get_aabb_2d(means.contiguous(), scales.contiguous(), glob_scale,
                          quats.contiguous(), viewmat.contiguous(), fx, fy, cx, cy)


def get_num_tiles_hit_2d(centers: Tensor, extents: Tensor, img_height: int, img_width: int,
                         block_width: int, wrapped: bool = False) -> Tensor:
    """Tiles touched by each AABB (get_aabb_2d.py:70-92), int32.  ``wrapped``
    (not in the reference) counts the wrapped rectangle that
    ``map_gaussian_to_intersects(..., wrapped=True)`` walks."""
    return _C.num_tiles_hit(centers.contiguous(), extents.contiguous(), img_height, img_width,
                            block_width, wrapped=wrapped)


def project_gaussians(means: Tensor, scales: Tensor, glob_scale, quats: Tensor, viewmat: Tensor,
                      intrins, img_height: int, img_width: int, block_width: int):
    """get_aabb_2d + depths + get_num_tiles_hit_2d in one kernel ->
    (centers, extents, depths, num_tiles_hit)."""
    fx, fy, cx, cy = intrins
    return _C.project_gaussians(means.contiguous(), scales.contiguous(), glob_scale,
                                quats.contiguous(), viewmat.contiguous(), fx, fy, cx, cy,
                                img_height, img_width, block_width)
