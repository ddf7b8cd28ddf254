"""Drop-in for ``gstex_cuda.texture`` (texture.py:14-408): differentiable
rasterisation of textured 2D Gaussians on gfx950.

``texture_gaussians`` keeps the reference signature and returns
``(out_img, out_depth, out_reg, out_alpha, out_texture, out_normal)``.  The
backward is the hand-written HIP kernel (transmittance recovered back to
front), exposed through ``torch.autograd.Function``.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor
from torch.autograd import Function

from . import cuda as _C
from .utils import compute_cumulative_intersects


def texture_gaussians(texture_info: Tuple[int, int, int], texture_dims: Tensor, centers: Tensor,
                      extents: Tensor, depths: Tensor, num_tiles_hit: Tensor, colors: Tensor,
                      opacity: Tensor, means: Tensor, scales: Tensor, glob_scale,
                      quats: Tensor, uv0: Tensor, umap: Tensor, vmap: Tensor, texture: Tensor,
                      viewmat: Tensor, c2w: Tensor, fx: float, fy: float, cx: float, cy: float,
                      img_height: int, img_width: int, block_width: int, settings: int,
                      background: Optional[Tensor] = None, use_torch_impl: bool = False):
    """Rasterise textured 2D Gaussians (texture.py:14-150).

    Arguments and outputs are those of the reference; see its docstring.  The
    reference's pure-torch path (``use_torch_impl=True``) is not provided.
    """
    assert 1 < block_width <= 16, "block_width must be between 2 and 16"
    if colors.dtype == torch.uint8:
        colors = colors.float() / 255
    if background is not None:
        assert background.shape[0] == colors.shape[-1], (
            f"incorrect shape of background color tensor, expected shape {colors.shape[-1]}")
    else:
        background = torch.ones(colors.shape[-1], dtype=torch.float32, device=colors.device)
    if colors.ndimension() != 2:
        raise ValueError("colors must have dimensions (N, D)")
    if use_torch_impl:
        raise NotImplementedError("gstex_cuda_amd has no torch implementation path")
    return _TextureGaussians.apply(
        texture_info, texture_dims.contiguous(), centers.contiguous(), extents.contiguous(),
        depths.contiguous(), num_tiles_hit.contiguous(), colors.contiguous(),
        opacity.contiguous(), means.contiguous(), scales.contiguous(), glob_scale,
        quats.contiguous(), uv0.contiguous(), umap.contiguous(), vmap.contiguous(),
        texture.contiguous(), viewmat.contiguous(), c2w.contiguous(), fx, fy, cx, cy,
        img_height, img_width, block_width, settings, background.contiguous())


def _tile_bounds(img_width: int, img_height: int, block_width: int):
    return ((img_width + block_width - 1) // block_width,
            (img_height + block_width - 1) // block_width, 1)


class _TextureGaussians(Function):
    """Forward: depth-first binning + HIP compositing; backward: HIP adjoint."""

    @staticmethod
    def forward(ctx, texture_info, texture_dims, centers, extents, depths, num_tiles_hit, colors,
                opacity, means, scales, glob_scale, quats, uv0, umap, vmap, texture, viewmat,
                c2w, fx, fy, cx, cy, img_height, img_width, block_width, settings, background):
        num_points = centers.size(0)
        tile_bounds = _tile_bounds(img_width, img_height, block_width)
        num_intersects, _ = compute_cumulative_intersects(num_tiles_hit)
        dev = means.device
        if num_intersects < 1:
            # Nothing hits the image: background colour, empty everything else.
            # (The reference leaves its outputs unbound here, texture.py:197-205.)
            gaussian_ids_sorted = torch.zeros((0,), dtype=torch.int32, device=dev)
            tile_bins = torch.zeros((tile_bounds[0] * tile_bounds[1], 2), dtype=torch.int32,
                                    device=dev)
        else:
            gaussian_ids_sorted, tile_bins = _C.bin_gaussians(
                num_points, num_intersects, centers, extents, depths,
                num_tiles_hit.to(torch.int32).contiguous(), tile_bounds, block_width)
        # Splat records are built once and kept for the backward; the wave
        # masks let the backward skip entries no pixel of a wave composited.
        if int(texture_info[1]) != 1:
            raise RuntimeError("texture_info[1] (num_probs) must be 1")
        # texels may be stored as bfloat16 (float32 arithmetic and gradients)
        _C._raster_checks(dict(texture_dims=texture_dims, gaussian_ids_sorted=gaussian_ids_sorted,
                               tile_bins=tile_bins, colors=colors, opacities=opacity,
                               means=means, scales=scales, quats=quats, uv0=uv0, umap=umap,
                               vmap=vmap, texture=texture, viewmat=viewmat, c2w=c2w,
                               background=background, texture_info=texture_info),
                          bf16_texture_ok=True)
        records = _C.build_splat_records(means, scales, glob_scale, quats, colors, opacity, uv0,
                                         umap, vmap, texture_dims, viewmat, c2w, fx, fy, cx, cy)
        C = int(texture_info[2])
        ((out_img, out_depth, out_reg, out_texture, out_normal, final_Ts, final_idx, depth_idx,
          out_reg_s), wave_masks) = _C.raster_forward(
            records, num_points, img_width, img_height, block_width, C, settings,
            gaussian_ids_sorted, tile_bins, texture, viewmat, c2w, fx, fy, cx, cy, background)
        ctx.img_width = img_width
        ctx.img_height = img_height
        ctx.texture_info = texture_info
        ctx.fx, ctx.fy, ctx.cx, ctx.cy = fx, fy, cx, cy
        ctx.glob_scale = glob_scale
        ctx.num_intersects = num_intersects
        ctx.block_width = block_width
        ctx.settings = settings
        ctx.save_for_backward(gaussian_ids_sorted, tile_bins, records, wave_masks, scales, quats,
                              uv0, umap, vmap, texture, viewmat, c2w, background, final_Ts,
                              final_idx, depth_idx, out_reg_s)
        out_alpha = 1 - final_Ts
        return out_img, out_depth, out_reg, out_alpha, out_texture, out_normal

    @staticmethod
    def backward(ctx, v_out_img, v_out_depth, v_out_reg, v_out_alpha, v_out_texture,
                 v_out_normal):
        (gaussian_ids_sorted, tile_bins, records, wave_masks, scales, quats, uv0, umap, vmap,
         texture, viewmat, c2w, background, final_Ts, final_idx, depth_idx,
         out_reg_s) = ctx.saved_tensors

        def grad_or_zeros(g, like):
            return torch.zeros_like(like) if g is None else g.contiguous()

        v_out_img = grad_or_zeros(v_out_img, final_Ts.unsqueeze(-1).expand(-1, -1, 3))
        v_out_depth = grad_or_zeros(v_out_depth, final_Ts)
        v_out_reg = grad_or_zeros(v_out_reg, final_Ts)
        v_out_alpha = grad_or_zeros(v_out_alpha, final_Ts)
        v_out_normal = grad_or_zeros(v_out_normal, final_Ts.unsqueeze(-1).expand(-1, -1, 3))
        C = int(ctx.texture_info[2])
        v_out_texture = grad_or_zeros(
            v_out_texture, final_Ts.unsqueeze(-1).expand(-1, -1, C))
        n = scales.size(0)
        grad_records, v_texture = _C.raster_backward(
            records, n, ctx.img_width, ctx.img_height, ctx.block_width, C, ctx.settings,
            gaussian_ids_sorted, tile_bins, texture, viewmat, c2w, ctx.fx, ctx.fy, ctx.cx, ctx.cy,
            background, final_Ts, final_idx, depth_idx, out_reg_s, v_out_img, v_out_depth,
            v_out_reg, v_out_alpha, v_out_texture, v_out_normal, wave_masks)
        (v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap,
         v_vmap) = _C.finalize_gradients(grad_records, records, quats, scales, viewmat, c2w,
                                         ctx.fx, ctx.fy)
        return (None, None, None, None, None, None, v_colors, v_opacity, v_means, v_scales, None,
                v_quats, v_uv0.view_as(uv0), v_umap.view_as(umap), v_vmap.view_as(vmap),
                v_texture, None, None, None, None, None, None, None, None, None, None, None)
