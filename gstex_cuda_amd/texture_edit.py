"""Drop-in for ``gstex_cuda.texture_edit`` (texture_edit.py:14-239): splat an
edit canvas onto the Gaussians' texel charts.

``texture_edit`` bins the Gaussians like ``texture_gaussians`` and runs the
gfx950 edit kernel (``gstex_texture_edit`` in the C ABI).  The result has the
texture's row count and ``texture_info[2]`` (>= 5) channels: channels 0..4 hold
(r*a, g*a, b*a, a, weight) summed over every composited pixel/Gaussian pair
whose depth lies in the pixel's [depth_lower, depth_upper], each pair spread
over its bilinear texel quad.  Dividing channels 0..3 by channel 4 gives the
texels that reproduce the edit.  There is no gradient (as in the reference).
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor

from . import cuda as _C
from .utils import compute_cumulative_intersects


def texture_edit(texture_info: Tuple[int, int, int], texture_dims: Tensor, updated_img: Tensor,
                 updated_alpha: Tensor, depth_lower: Tensor, depth_upper: Tensor,
                 centers: Tensor, extents: Tensor, depths: Tensor, num_tiles_hit: Tensor,
                 opacity: Tensor, means: Tensor, scales: Tensor, glob_scale, quats: Tensor,
                 uv0: Tensor, umap: Tensor, vmap: Tensor, viewmat: Tensor, c2w: Tensor,
                 fx: float, fy: float, cx: float, cy: float, img_height: int, img_width: int,
                 block_width: int, settings: int, background: Optional[Tensor] = None,
                 use_torch_impl: bool = False) -> Tensor:
    """texture_edit.py:14-239.  ``settings`` bit 0 is the edit kernel's blur switch."""
    assert block_width > 1 and block_width <= 16, "block_width must be between 2 and 16"
    if use_torch_impl:
        raise NotImplementedError("only the GPU implementation exists (as in the reference)")
    tile_bounds = ((img_width + block_width - 1) // block_width,
                   (img_height + block_width - 1) // block_width, 1)
    num_points = centers.size(0)
    # the reference sizes the output as the sum of the charts' texel counts
    texture_total_size = int(torch.sum(texture_dims[:, 0] * texture_dims[:, 1]).item())
    num_intersects, _ = compute_cumulative_intersects(num_tiles_hit)
    if num_intersects < 1:
        raise RuntimeError("texture_edit: no Gaussian intersects the image")
    gaussian_ids_sorted, tile_bins = _C.bin_gaussians(
        num_points, num_intersects, centers.contiguous(), extents.contiguous(),
        depths.contiguous(), num_tiles_hit.to(torch.int32).contiguous(), tile_bounds, block_width)
    if background is None:
        background = torch.zeros(3, dtype=torch.float32, device=means.device)
    return _C.texture_edit(
        tile_bounds, (block_width, block_width, 1), (img_width, img_height, 1), texture_info,
        texture_total_size, texture_dims.contiguous(), updated_img.contiguous(),
        updated_alpha.contiguous(), depth_lower.contiguous(), depth_upper.contiguous(),
        gaussian_ids_sorted, tile_bins, opacity.contiguous(), means.contiguous(),
        scales.contiguous(), glob_scale, quats.contiguous(), uv0.contiguous(), umap.contiguous(),
        vmap.contiguous(), viewmat.contiguous(), c2w.contiguous(), fx, fy, cx, cy, settings,
        background.contiguous())
