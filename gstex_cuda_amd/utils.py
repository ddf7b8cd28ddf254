"""Drop-in for ``gstex_cuda.utils`` (utils.py:11-162): tile binning on the GPU.

``bin_and_sort_gaussians`` keeps the reference's five outputs; the sort is the
library's stable radix sort instead of torch.sort, so ties (equal depth bits
inside a tile) are ordered by intersect position, i.e. by Gaussian id.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch import Tensor

from . import cuda as _C


def get_tile_bin_edges(num_intersects: int, isect_ids_sorted: Tensor,
                       tile_bounds: Tuple[int, int, int]) -> Tensor:
    """[first, last+1) of every tile in the sorted intersect ids (utils.py:11-37)."""
    return _C.get_tile_bin_edges(num_intersects, isect_ids_sorted.contiguous(), tile_bounds)


def compute_cumulative_intersects(num_tiles_hit: Tensor) -> Tuple[int, Tensor]:
    """Inclusive int32 cumsum and its total (utils.py:40-59).  The total is read
    back to the host, as in the reference."""
    cum_tiles_hit = torch.cumsum(num_tiles_hit, dim=0, dtype=torch.int32)
    num_intersects = int(cum_tiles_hit[-1].item()) if cum_tiles_hit.numel() else 0
    return num_intersects, cum_tiles_hit


def map_gaussian_to_intersects(num_points: int, num_intersects: int, centers: Tensor,
                               extents: Tensor, depths: Tensor, cum_tiles_hit: Tensor,
                               tile_bounds: Tuple[int, int, int], block_size: int,
                               wrapped: bool = False) -> Tuple[Tensor, Tensor]:
    """(tile << 32 | depth bits, Gaussian id) per intersect (utils.py:61-104)."""
    return _C.map_gaussian_to_intersects(
        num_points, num_intersects, centers.contiguous(), extents.contiguous(),
        depths.contiguous(), cum_tiles_hit.contiguous(), tile_bounds, block_size, wrapped)


def _key_bits(tile_bounds) -> int:
    tiles = max(int(tile_bounds[0]) * int(tile_bounds[1]), 1)
    return 32 + max(tiles - 1, 0).bit_length()


def bin_and_sort_gaussians(num_points: int, num_intersects: int, centers: Tensor, extents: Tensor,
                           depths: Tensor, cum_tiles_hit: Tensor,
                           tile_bounds: Tuple[int, int, int], block_size: int,
                           wrapped: bool = False):
    """Keys, sort and bin edges (utils.py:106-162) ->
    (isect_ids, gaussian_ids, isect_ids_sorted, gaussian_ids_sorted, tile_bins)."""
    isect_ids, gaussian_ids = map_gaussian_to_intersects(
        num_points, num_intersects, centers, extents, depths, cum_tiles_hit, tile_bounds,
        block_size, wrapped=wrapped)
    isect_ids_sorted, gaussian_ids_sorted = _C.sort_intersects(
        isect_ids, gaussian_ids, end_bit=_key_bits(tile_bounds))
    tile_bins = get_tile_bin_edges(num_intersects, isect_ids_sorted, tile_bounds)
    return isect_ids, gaussian_ids, isect_ids_sorted, gaussian_ids_sorted, tile_bins


def bin_gaussians(num_points: int, num_intersects: int, centers: Tensor, extents: Tensor,
                  depths: Tensor, num_tiles_hit: Tensor, tile_bounds: Tuple[int, int, int],
                  block_size: int) -> Tuple[Tensor, Tensor]:
    """(gaussian_ids_sorted, tile_bins) identical to bin_and_sort_gaussians', from
    one depth sort of the Gaussians and a stable tile sort of the intersects."""
    return _C.bin_gaussians(num_points, num_intersects, centers.contiguous(),
                            extents.contiguous(), depths.contiguous(),
                            num_tiles_hit.to(torch.int32).contiguous(), tile_bounds, block_size)
