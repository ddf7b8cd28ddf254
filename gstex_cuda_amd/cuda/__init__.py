"""Drop-in for ``gstex_cuda.cuda``: the ten native entry points of the reference
(cuda/__init__.py:14-26, bound in cuda/csrc/ext.cpp:8-25), same names, same
positional arguments, same return tuples.

Each function validates its tensors like the reference's CHECK_INPUT
(``"<name> must be a CUDA tensor"`` / ``"<name> must be contiguous"`` as
RuntimeError), allocates outputs with torch on the input's device, and runs
the gfx950 kernels through the C ABI on torch's current stream.  There is no
CPU path: without the built library every call raises.
"""
from __future__ import annotations

from typing import Tuple

import torch

from .. import _lib
from .._lib import call

__all__ = [
    "compute_sh_forward", "compute_sh_backward", "map_gaussian_to_intersects",
    "get_tile_bin_edges", "get_aabb_2d", "texture_forward", "texture_backward",
    "texture_sample_forward", "texture_sample_backward", "texture_edit",
    # additions beyond the reference's ten
    "project_gaussians", "project_points", "num_tiles_hit", "count_intersects", "sort_intersects",
    "sort_pairs_u32",
    "bin_gaussians",
    "build_splat_records", "raster_forward", "raster_backward", "finalize_gradients",
    "set_raster_culling", "set_precise_edge_on",
]


def _check(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _dtype(t: torch.Tensor, name: str, dtype: torch.dtype) -> None:
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must have dtype {dtype}, got {t.dtype}")


def _ptr(t: torch.Tensor):
    return t.data_ptr() if t.numel() > 0 else None


def _stream(device: torch.device):
    return torch.cuda.current_stream(device).cuda_stream


def _workspace(nbytes: int, device: torch.device) -> Tuple[torch.Tensor, int]:
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    return ws, int(nbytes)


# --------------------------------------------------------------------------- SH
def compute_sh_forward(num_points: int, degree: int, degrees_to_use: int,
                       viewdirs: torch.Tensor, coeffs: torch.Tensor) -> torch.Tensor:
    """bindings.cu:18-43: colours (N, 3) from coeffs (N, D, 3)."""
    _check(viewdirs, "viewdirs")
    _check(coeffs, "coeffs")
    nb = _num_bases(degree)
    if coeffs.dim() != 3 or coeffs.size(0) != num_points or coeffs.size(1) != nb or coeffs.size(2) != 3:
        raise RuntimeError("coeffs must have dimensions (N, D, 3)")
    _dtype(viewdirs, "viewdirs", torch.float32)
    if coeffs.dtype not in (torch.float32, torch.bfloat16):
        raise RuntimeError(f"coeffs must have dtype float32 or bfloat16, got {coeffs.dtype}")
    colors = torch.empty((num_points, 3), dtype=torch.float32, device=coeffs.device)
    fn = ("gstex_compute_sh_forward_bf16" if coeffs.dtype == torch.bfloat16
          else "gstex_compute_sh_forward")
    with torch.cuda.device(coeffs.device):
        call(fn, num_points, degree, degrees_to_use, _ptr(viewdirs), _ptr(coeffs), _ptr(colors),
             _stream(coeffs.device))
    return colors


def compute_sh_backward(num_points: int, degree: int, degrees_to_use: int,
                        viewdirs: torch.Tensor, v_colors: torch.Tensor) -> torch.Tensor:
    """bindings.cu:45-75: v_coeffs (N, D, 3)."""
    _check(viewdirs, "viewdirs")
    _check(v_colors, "v_colors")
    if viewdirs.dim() != 2 or viewdirs.size(0) != num_points or viewdirs.size(1) != 3:
        raise RuntimeError("viewdirs must have dimensions (N, 3)")
    if v_colors.dim() != 2 or v_colors.size(0) != num_points or v_colors.size(1) != 3:
        raise RuntimeError("v_colors must have dimensions (N, 3)")
    nb = _num_bases(degree)
    v_coeffs = torch.empty((num_points, nb, 3), dtype=torch.float32, device=v_colors.device)
    with torch.cuda.device(v_colors.device):
        call("gstex_compute_sh_backward", num_points, degree, degrees_to_use, _ptr(viewdirs),
             _ptr(v_colors), _ptr(v_coeffs), _stream(v_colors.device))
    return v_coeffs


def _num_bases(degree: int) -> int:
    return {0: 1, 1: 4, 2: 9, 3: 16}.get(int(degree), 25)


# --------------------------------------------------------------------------- projection
def get_aabb_2d(means, scales, glob_scale, quats, viewmat, fx, fy, cx, cy):
    """get_aabb_2d.cu:91-126 -> (centers (N,2), extents (N,2))."""
    for t, n in ((means, "means"), (scales, "scales"), (quats, "quats"), (viewmat, "viewmat")):
        _check(t, n)
        _dtype(t, n, torch.float32)
    n = means.size(0)
    centers = torch.empty((n, 2), dtype=torch.float32, device=means.device)
    extents = torch.empty((n, 2), dtype=torch.float32, device=means.device)
    with torch.cuda.device(means.device):
        call("gstex_get_aabb_2d", n, _ptr(means), _ptr(scales), float(glob_scale), _ptr(quats),
             _ptr(viewmat), float(fx), float(fy), float(cx), float(cy), _ptr(centers),
             _ptr(extents), _stream(means.device))
    return centers, extents


def num_tiles_hit(centers, extents, img_height, img_width, block_width, wrapped=False):
    """get_aabb_2d.py:70-92 on the GPU -> (N,) int32.  ``wrapped``: the tile
    count of each Gaussian's wrapped (unclamped) rectangle instead, which
    map_gaussian_to_intersects(..., wrapped=True) then fills completely."""
    _check(centers, "centers")
    _check(extents, "extents")
    n = centers.size(0)
    out = torch.empty((n,), dtype=torch.int32, device=centers.device)
    with torch.cuda.device(centers.device):
        if wrapped:
            call("gstex_num_tiles_hit_wrapped", n, _ptr(centers), _ptr(extents),
                 int(block_width), _ptr(out), _stream(centers.device))
        else:
            call("gstex_num_tiles_hit", n, _ptr(centers), _ptr(extents), int(img_height),
                 int(img_width), int(block_width), _ptr(out), _stream(centers.device))
    return out


def project_points(points, viewmat, fx, fy, cx, cy, clip=False):
    """get_aabb_2d.py:22-32 -> (pix (N,2), depths (N,))."""
    _check(points, "points")
    _check(viewmat, "viewmat")
    if viewmat.numel() < 12:
        raise RuntimeError("viewmat must hold at least three rows of four")
    n = points.size(0)
    pix = torch.empty((n, 2), dtype=torch.float32, device=points.device)
    depths = torch.empty((n,), dtype=torch.float32, device=points.device)
    with torch.cuda.device(points.device):
        call("gstex_project_points", n, _ptr(points), _ptr(viewmat), float(fx), float(fy),
             float(cx), float(cy), int(bool(clip)), _ptr(pix), _ptr(depths),
             _stream(points.device))
    return pix, depths


def project_gaussians(means, scales, glob_scale, quats, viewmat, fx, fy, cx, cy, img_height,
                      img_width, block_width):
    """Fused get_aabb_2d + depth + get_num_tiles_hit_2d ->
    (centers, extents, depths, num_tiles_hit)."""
    for t, nm in ((means, "means"), (scales, "scales"), (quats, "quats"), (viewmat, "viewmat")):
        _check(t, nm)
        _dtype(t, nm, torch.float32)
    n = means.size(0)
    dev = means.device
    centers = torch.empty((n, 2), dtype=torch.float32, device=dev)
    extents = torch.empty((n, 2), dtype=torch.float32, device=dev)
    depths = torch.empty((n,), dtype=torch.float32, device=dev)
    tiles = torch.empty((n,), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_project_gaussians", n, _ptr(means), _ptr(scales), float(glob_scale),
             _ptr(quats), _ptr(viewmat), float(fx), float(fy), float(cx), float(cy),
             int(img_height), int(img_width), int(block_width), _ptr(centers), _ptr(extents),
             _ptr(depths), _ptr(tiles), _stream(dev))
    return centers, extents, depths, tiles


# --------------------------------------------------------------------------- binning
def map_gaussian_to_intersects(num_points, num_intersects, centers, extents, depths,
                               cum_tiles_hit, tile_bounds, block_width, wrapped=False):
    """bindings.cu:77-121 -> (isect_ids int64 (I,), gaussian_ids int32 (I,))."""
    for t, nm in ((centers, "centers"), (extents, "extents"), (depths, "depths"),
                  (cum_tiles_hit, "cum_tiles_hit")):
        _check(t, nm)
    _dtype(cum_tiles_hit, "cum_tiles_hit", torch.int32)
    dev = centers.device
    isect_ids = torch.empty((num_intersects,), dtype=torch.int64, device=dev)
    gaussian_ids = torch.empty((num_intersects,), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_map_gaussian_to_intersects", int(num_points), int(num_intersects),
             _ptr(centers), _ptr(extents), _ptr(depths), _ptr(cum_tiles_hit),
             int(tile_bounds[0]), int(tile_bounds[1]), int(block_width), int(bool(wrapped)),
             _ptr(isect_ids), _ptr(gaussian_ids), _stream(dev))
    return isect_ids, gaussian_ids


def sort_intersects(isect_ids, gaussian_ids, end_bit=64):
    """Stable radix sort replacing torch.sort + gather (utils.py:159-160)."""
    _check(isect_ids, "isect_ids")
    _check(gaussian_ids, "gaussian_ids")
    m = isect_ids.numel()
    dev = isect_ids.device
    keys_out = torch.empty_like(isect_ids)
    ids_out = torch.empty_like(gaussian_ids)
    if m == 0:
        return keys_out, ids_out
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_sort_intersects_workspace_size(m), dev)
        call("gstex_sort_intersects", m, int(end_bit), _ptr(isect_ids), _ptr(gaussian_ids),
             _ptr(keys_out), _ptr(ids_out), _ptr(ws), nb, _stream(dev))
    return keys_out, ids_out


def count_intersects(num_tiles_hit) -> int:
    """Total intersect count sum(max(num_tiles_hit, 0)), read back to the host
    (the total of compute_cumulative_intersects, utils.py:40-59)."""
    _check(num_tiles_hit, "num_tiles_hit")
    dev = num_tiles_hit.device
    out = torch.zeros((1,), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_count_intersects", num_tiles_hit.numel(), _ptr(num_tiles_hit), _ptr(out),
             _stream(dev))
    return int(out.item())


def sort_pairs_u32(keys, values, end_bit=32):
    """Stable radix sort of (uint32 key, uint32 value) pairs over key bits
    [0, end_bit); int32 / uint32 tensors of equal length."""
    _check(keys, "keys")
    _check(values, "values")
    n = keys.numel()
    if values.numel() != n:
        raise ValueError("keys and values must have the same length")
    dev = keys.device
    keys_out = torch.empty_like(keys)
    values_out = torch.empty_like(values)
    if n == 0:
        return keys_out, values_out
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_sort_pairs_u32_workspace_size(n), dev)
        call("gstex_sort_pairs_u32", n, int(end_bit), _ptr(keys), _ptr(values), _ptr(keys_out),
             _ptr(values_out), _ptr(ws), nb, _stream(dev))
    return keys_out, values_out


def get_tile_bin_edges(num_intersects, isect_ids_sorted, tile_bounds):
    """bindings.cu:123-141 -> tile_bins (T, 2) int32."""
    _check(isect_ids_sorted, "isect_ids_sorted")
    dev = isect_ids_sorted.device
    num_tiles = int(tile_bounds[0]) * int(tile_bounds[1])
    tile_bins = torch.empty((num_tiles, 2), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_get_tile_bin_edges", int(num_intersects), _ptr(isect_ids_sorted),
             int(tile_bounds[0]), int(tile_bounds[1]), _ptr(tile_bins), _stream(dev))
    return tile_bins


def bin_gaussians(num_points, num_intersects, centers, extents, depths, num_tiles_hit,
                  tile_bounds, block_width):
    """Depth-first binning -> (gaussian_ids_sorted (I,) int32, tile_bins (T,2) int32)."""
    for t, nm in ((centers, "centers"), (extents, "extents"), (depths, "depths"),
                  (num_tiles_hit, "num_tiles_hit")):
        _check(t, nm)
    _dtype(num_tiles_hit, "num_tiles_hit", torch.int32)
    dev = centers.device
    num_tiles = int(tile_bounds[0]) * int(tile_bounds[1])
    ids = torch.empty((num_intersects,), dtype=torch.int32, device=dev)
    bins = torch.empty((num_tiles, 2), dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        ws, nb = _workspace(
            _lib.lib().gstex_bin_gaussians_workspace_size(int(num_points), int(num_intersects)),
            dev)
        call("gstex_bin_gaussians", int(num_points), int(num_intersects), _ptr(centers),
             _ptr(extents), _ptr(depths), _ptr(num_tiles_hit), int(tile_bounds[0]),
             int(tile_bounds[1]), int(block_width), _ptr(ids), _ptr(bins), _ptr(ws), nb,
             _stream(dev))
    return ids, bins


# --------------------------------------------------------------------------- raster
_RASTER_INPUTS = ("texture_dims", "gaussian_ids_sorted", "tile_bins", "colors", "opacities",
                  "means", "scales", "quats", "uv0", "umap", "vmap", "texture", "viewmat", "c2w",
                  "background")


def _texture_shape(texture: torch.Tensor, channels: int) -> None:
    """The kernels size every chart as h * w * channels values of ``texture``
    rows, so its row width must be the channel count."""
    if texture.dim() != 2 or texture.size(1) != int(channels):
        raise RuntimeError(f"texture must have dimensions (X, {int(channels)}) to match "
                           f"texture_info[2], got {tuple(texture.shape)}")


def _raster_checks(named, bf16_texture_ok=False):
    for name in _RASTER_INPUTS:
        _check(named[name], name)
    for name in ("texture_dims", "gaussian_ids_sorted", "tile_bins"):
        _dtype(named[name], name, torch.int32)
    for name in ("colors", "opacities", "means", "scales", "quats", "uv0", "umap", "vmap",
                 "viewmat", "c2w", "background"):
        _dtype(named[name], name, torch.float32)
    if bf16_texture_ok:
        _texture_format(named["texture"])
    else:
        _dtype(named["texture"], "texture", torch.float32)
    _texture_shape(named["texture"], named["texture_info"][2])


def texture_forward(tile_bounds, block, img_size, texture_info, texture_dims,
                    gaussian_ids_sorted, tile_bins, colors, opacities, means, scales, glob_scale,
                    quats, uv0, umap, vmap, texture, viewmat, c2w, fx, fy, cx, cy, settings,
                    background):
    """texture.cu:762-901 -> (out_img, out_depth, out_reg, out_texture, out_normal,
    final_Ts, final_idx, depth_idx, out_reg_s)."""
    _raster_checks(locals())
    dev = means.device
    W, H = int(img_size[0]), int(img_size[1])
    C = int(texture_info[2])
    n = means.size(0)
    f32 = dict(dtype=torch.float32, device=dev)
    out_img = torch.empty((H, W, colors.size(1)), **f32)
    out_depth = torch.empty((H, W), **f32)
    out_reg = torch.empty((H, W), **f32)
    out_texture = torch.empty((H, W, C), **f32)
    out_normal = torch.empty((H, W, 3), **f32)
    final_Ts = torch.empty((H, W), **f32)
    final_idx = torch.empty((H, W), dtype=torch.int32, device=dev)
    depth_idx = torch.empty((H, W), dtype=torch.int32, device=dev)
    out_reg_s = torch.empty((H, W, 3), **f32)
    if int(block[0]) != int(block[1]):
        raise RuntimeError("block must be square (block_width, block_width, 1)")
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_texture_forward_workspace_size(n), dev)
        call("gstex_texture_forward", int(tile_bounds[0]), int(tile_bounds[1]), int(block[0]),
             W, H, int(texture_info[0]), int(texture_info[1]), C, _ptr(texture_dims),
             _ptr(gaussian_ids_sorted), gaussian_ids_sorted.numel(), _ptr(tile_bins),
             _ptr(colors), _ptr(opacities), _ptr(means), _ptr(scales), n, float(glob_scale),
             _ptr(quats), _ptr(uv0), _ptr(umap), _ptr(vmap), _ptr(texture),
             texture.size(0) if texture.dim() > 0 else 0, _ptr(viewmat), _ptr(c2w), float(fx),
             float(fy), float(cx), float(cy), int(settings), _ptr(background), _ptr(out_img),
             _ptr(out_depth), _ptr(out_reg), _ptr(out_texture), _ptr(out_normal), _ptr(final_Ts),
             _ptr(final_idx), _ptr(depth_idx), _ptr(out_reg_s), _ptr(ws), nb, _stream(dev))
    return (out_img, out_depth, out_reg, out_texture, out_normal, final_Ts, final_idx, depth_idx,
            out_reg_s)


def texture_backward(img_height, img_width, block_width, texture_info, texture_dims,
                     gaussian_ids_sorted, tile_bins, colors, opacities, means, scales, glob_scale,
                     quats, uv0, umap, vmap, texture, viewmat, c2w, fx, fy, cx, cy, settings,
                     background, final_Ts, final_idx, depth_idx, final_s, v_output,
                     v_output_depth, v_output_reg, v_output_alpha, v_output_texture,
                     v_output_normal):
    """texture.cu:903-1053 -> (v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap,
    v_vmap, v_texture)."""
    named = locals()
    _raster_checks(named)
    for name in ("final_Ts", "final_idx", "depth_idx", "final_s", "v_output", "v_output_depth",
                 "v_output_reg", "v_output_alpha", "v_output_texture", "v_output_normal"):
        _check(named[name], name)
    dev = means.device
    n = means.size(0)
    C = int(texture_info[2])
    num_probs = int(texture_info[1])
    f32 = dict(dtype=torch.float32, device=dev)
    v_colors = torch.empty((n, colors.size(1)), **f32)
    v_opacity = torch.empty((n, 1), **f32)
    v_means = torch.empty((n, 3), **f32)
    v_scales = torch.empty((n, 3), **f32)
    v_quats = torch.empty((n, 4), **f32)
    v_uv0 = torch.empty((n, num_probs, 2), **f32)
    v_umap = torch.empty((n, num_probs, 3), **f32)
    v_vmap = torch.empty((n, num_probs, 3), **f32)
    v_texture = torch.empty((texture.size(0), C), **f32)
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_texture_backward_workspace_size(n), dev)
        call("gstex_texture_backward", int(img_height), int(img_width), int(block_width),
             int(texture_info[0]), num_probs, C, _ptr(texture_dims), _ptr(gaussian_ids_sorted),
             gaussian_ids_sorted.numel(), _ptr(tile_bins), _ptr(colors), _ptr(opacities),
             _ptr(means), _ptr(scales), n, float(glob_scale), _ptr(quats), _ptr(uv0), _ptr(umap),
             _ptr(vmap), _ptr(texture), texture.size(0), _ptr(viewmat), _ptr(c2w), float(fx),
             float(fy), float(cx), float(cy), int(settings), _ptr(background), _ptr(final_Ts),
             _ptr(final_idx), _ptr(depth_idx), _ptr(final_s), _ptr(v_output),
             _ptr(v_output_depth), _ptr(v_output_reg), _ptr(v_output_alpha),
             _ptr(v_output_texture), _ptr(v_output_normal), _ptr(v_colors), _ptr(v_opacity),
             _ptr(v_means), _ptr(v_scales), _ptr(v_quats), _ptr(v_uv0), _ptr(v_umap),
             _ptr(v_vmap), _ptr(v_texture), _ptr(ws), nb, _stream(dev))
    return v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap, v_texture


# --------------------------------------------------------------------------- per-stage raster
# The training path splits texture_forward / texture_backward into stages so
# that splat records are built once per view and the backward visits only the
# entries the forward composited (wave masks).  Same semantics as the two
# reference-shaped ops above.
def _texture_format(texture: torch.Tensor) -> int:
    if texture.dtype == torch.bfloat16:
        return 1
    if texture.dtype == torch.float32:
        return 0
    raise RuntimeError(f"texture must have dtype float32 or bfloat16, got {texture.dtype}")


def build_splat_records(means, scales, glob_scale, quats, colors, opacities, uv0, umap, vmap,
                        texture_dims, viewmat, c2w, fx, fy, cx, cy):
    """Per-Gaussian splat records + screen boxes (uint8 tensor) for
    raster_forward/backward."""
    named = locals()
    for name in ("means", "scales", "quats", "colors", "opacities", "uv0", "umap", "vmap",
                 "texture_dims", "viewmat", "c2w"):
        _check(named[name], name)
    dev = means.device
    n = means.size(0)
    records = torch.empty((max(int(_lib.lib().gstex_splat_records_size(n)), 1),),
                          dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        call("gstex_build_splat_records", n, _ptr(means), _ptr(scales), float(glob_scale),
             _ptr(quats), _ptr(colors), _ptr(opacities), _ptr(uv0), _ptr(umap), _ptr(vmap),
             _ptr(texture_dims), _ptr(viewmat), _ptr(c2w), float(fx), float(fy), float(cx),
             float(cy), _ptr(records), _stream(dev))
    return records


def set_precise_edge_on(enabled: bool) -> bool:
    """Solve near-edge-on ray/plane hits (|n . ray| < 0.05) in float64 (default off:
    the reference's float32 arithmetic); returns the previous setting."""
    return bool(_lib.lib().gstex_set_precise_edge_on(int(bool(enabled))))


def set_raster_culling(enabled: bool) -> bool:
    """Turn the raster kernels' per-wave screen-box culling on or off (results are
    identical either way); returns the previous setting."""
    return bool(_lib.lib().gstex_set_raster_culling(int(bool(enabled))))


def raster_forward(records, num_points, img_width, img_height, block_width, channels, settings,
                   gaussian_ids_sorted, tile_bins, texture, viewmat, c2w, fx, fy, cx, cy,
                   background, wave_masks=True):
    """Forward compositing from splat records -> (the nine texture_forward outputs,
    wave_masks or None)."""
    for name, t in (("gaussian_ids_sorted", gaussian_ids_sorted), ("tile_bins", tile_bins),
                    ("texture", texture), ("viewmat", viewmat), ("c2w", c2w),
                    ("background", background)):
        _check(t, name)
    _texture_shape(texture, channels)
    dev = records.device
    W, H, C = int(img_width), int(img_height), int(channels)
    tx, ty = (W + block_width - 1) // block_width, (H + block_width - 1) // block_width
    f32 = dict(dtype=torch.float32, device=dev)
    outs = (torch.empty((H, W, 3), **f32), torch.empty((H, W), **f32),
            torch.empty((H, W), **f32), torch.empty((H, W, C), **f32),
            torch.empty((H, W, 3), **f32), torch.empty((H, W), **f32),
            torch.empty((H, W), dtype=torch.int32, device=dev),
            torch.empty((H, W), dtype=torch.int32, device=dev), torch.empty((H, W, 3), **f32))
    m = gaussian_ids_sorted.numel()
    masks = None
    if wave_masks:
        nb = int(_lib.lib().gstex_raster_wave_masks_size(m, tx, ty))
        masks = torch.empty((max(nb, 8),), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        call("gstex_raster_forward", W, H, int(block_width), C, int(settings), _ptr(records),
             int(num_points), _ptr(gaussian_ids_sorted), m, _ptr(tile_bins), _ptr(texture),
             texture.size(0), _texture_format(texture), _ptr(viewmat), _ptr(c2w), float(fx),
             float(fy), float(cx), float(cy), _ptr(background), *[_ptr(o) for o in outs],
             None if masks is None else _ptr(masks), 0 if masks is None else masks.numel(),
             _stream(dev))
    return outs, masks


def raster_backward(records, num_points, img_width, img_height, block_width, channels, settings,
                    gaussian_ids_sorted, tile_bins, texture, viewmat, c2w, fx, fy, cx, cy,
                    background, final_Ts, final_idx, depth_idx, final_s, v_output,
                    v_output_depth, v_output_reg, v_output_alpha, v_output_texture,
                    v_output_normal, wave_masks=None):
    """Adjoint of raster_forward -> (grad_records (N, 32), v_texture (X, C))."""
    named = locals()
    for name in ("final_Ts", "final_idx", "depth_idx", "final_s", "v_output", "v_output_depth",
                 "v_output_reg", "v_output_alpha", "v_output_texture", "v_output_normal"):
        _check(named[name], name)
    _check(texture, "texture")
    _texture_shape(texture, channels)
    dev = records.device
    n, C = int(num_points), int(channels)
    grad_records = torch.zeros((n, 32), dtype=torch.float32, device=dev)
    v_texture = torch.zeros((texture.size(0), C), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_raster_backward", int(img_width), int(img_height), int(block_width), C,
             int(settings), _ptr(records), n, _ptr(gaussian_ids_sorted),
             gaussian_ids_sorted.numel(), _ptr(tile_bins), _ptr(texture), texture.size(0),
             _texture_format(texture), _ptr(viewmat), _ptr(c2w), float(fx), float(fy),
             float(cx), float(cy), _ptr(background), _ptr(final_Ts), _ptr(final_idx), _ptr(depth_idx), _ptr(final_s),
             _ptr(v_output), _ptr(v_output_depth), _ptr(v_output_reg), _ptr(v_output_alpha),
             _ptr(v_output_texture), _ptr(v_output_normal), _ptr(grad_records),
             _ptr(v_texture), None if wave_masks is None else _ptr(wave_masks),
             0 if wave_masks is None else wave_masks.numel(), _stream(dev))
    return grad_records, v_texture


def finalize_gradients(grad_records, records, quats, scales, viewmat, c2w, fx, fy,
                       num_probs=1):
    """Reference gradient tensors from the backward's per-Gaussian sums ->
    (v_colors, v_opacity, v_means, v_scales, v_quats, v_uv0, v_umap, v_vmap)."""
    dev = grad_records.device
    n = grad_records.size(0)
    f32 = dict(dtype=torch.float32, device=dev)
    outs = (torch.empty((n, 3), **f32), torch.empty((n, 1), **f32), torch.empty((n, 3), **f32),
            torch.empty((n, 3), **f32), torch.empty((n, 4), **f32),
            torch.empty((n, num_probs, 2), **f32), torch.empty((n, num_probs, 3), **f32),
            torch.empty((n, num_probs, 3), **f32))
    with torch.cuda.device(dev):
        call("gstex_finalize_gradients", n, _ptr(grad_records), _ptr(records), _ptr(quats),
             _ptr(scales), _ptr(viewmat), _ptr(c2w), float(fx), float(fy),
             *[_ptr(o) for o in outs], _stream(dev))
    return outs


# --------------------------------------------------------------------------- sampling
def _sample_checks(texture_info, texture_dims, uvs, texture, v_output=None) -> int:
    """Dtype and shape checks of texture_sample_forward/backward (the reference
    raises through data_ptr<int>/data_ptr<float>, texture_sample.cu:98-101,
    133-137); returns the query count.  Each query reads its own texture_dims
    row, and the kernels index rows of C = texture_info[2] float32 values."""
    named = dict(texture_dims=texture_dims, uvs=uvs, texture=texture)
    if v_output is not None:
        named["v_output"] = v_output
    for nm, t in named.items():
        _check(t, nm)
    _dtype(texture_dims, "texture_dims", torch.int32)
    for nm in ("uvs", "texture", "v_output"):
        if nm in named:
            _dtype(named[nm], nm, torch.float32)
    C = int(texture_info[2])
    if uvs.dim() != 2 or uvs.size(1) != 2:
        raise RuntimeError(f"uvs must have dimensions (Q, 2), got {tuple(uvs.shape)}")
    q = uvs.size(0)
    if texture_dims.dim() != 2 or texture_dims.size(1) != 3 or texture_dims.size(0) < q:
        raise RuntimeError(f"texture_dims must have dimensions (Q, 3) with Q >= {q} queries, "
                           f"got {tuple(texture_dims.shape)}")
    _texture_shape(texture, C)
    if v_output is not None and (v_output.dim() != 2 or v_output.size(0) != q
                                 or v_output.size(1) != C):
        raise RuntimeError(f"v_output must have dimensions ({q}, {C}), got "
                           f"{tuple(v_output.shape)}")
    return q


def texture_sample_forward(texture_info, texture_dims, uvs, texture):
    """texture_sample.cu:72-105 -> (Q, C)."""
    q = _sample_checks(texture_info, texture_dims, uvs, texture)
    dev = texture.device
    C = int(texture_info[2])
    out = torch.empty((q, C), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        call("gstex_texture_sample_forward", int(texture_info[0]), int(texture_info[1]), C,
             _ptr(texture_dims), _ptr(uvs), q, _ptr(texture), texture.size(0), _ptr(out),
             _stream(dev))
    return out


def texture_sample_backward(texture_info, texture_dims, uvs, texture, v_output):
    """Adjoint of texture_sample_forward w.r.t. the texels -> (X, C)
    (texture_sample.cu:107-141, reading v_output)."""
    _sample_checks(texture_info, texture_dims, uvs, texture, v_output)
    dev = texture.device
    C = int(texture_info[2])
    v_texture = torch.empty((texture.size(0), C), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_texture_sample_backward_workspace_size(uvs.size(0)),
                            dev)
        call("gstex_texture_sample_backward", int(texture_info[0]), int(texture_info[1]), C,
             _ptr(texture_dims), _ptr(uvs), uvs.size(0), _ptr(texture), texture.size(0),
             _ptr(v_output), _ptr(v_texture), _ptr(ws), nb, _stream(dev))
    return v_texture


def texture_edit(tile_bounds, block, img_size, texture_info, texture_total_size, texture_dims,
                 updated_img, updated_alpha, depth_lower, depth_upper, gaussian_ids_sorted,
                 tile_bins, opacities, means, scales, glob_scale, quats, uv0, umap, vmap, viewmat,
                 c2w, fx, fy, cx, cy, settings, background):
    """texture_edit.cu:229-346 -> updated_texture (texture_total_size, C): channels
    0..4 hold the edit (r*a, g*a, b*a, a, weight) splatted onto the texel charts."""
    named = locals()
    for name in ("texture_dims", "updated_img", "updated_alpha", "depth_lower", "depth_upper",
                 "gaussian_ids_sorted", "tile_bins", "opacities", "means", "scales", "quats",
                 "uv0", "umap", "vmap", "viewmat", "c2w", "background"):
        _check(named[name], name)
    for name in ("texture_dims", "gaussian_ids_sorted", "tile_bins"):
        _dtype(named[name], name, torch.int32)
    for name in ("updated_img", "updated_alpha", "depth_lower", "depth_upper", "opacities",
                 "means", "scales", "quats", "uv0", "umap", "vmap", "viewmat", "c2w"):
        _dtype(named[name], name, torch.float32)
    if int(block[0]) != int(block[1]):
        raise RuntimeError("block must be square (block_width, block_width, 1)")
    dev = means.device
    W, H = int(img_size[0]), int(img_size[1])
    C = int(texture_info[2])
    X = int(texture_total_size)
    n = means.size(0)
    out = torch.empty((X, C), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        ws, nb = _workspace(_lib.lib().gstex_texture_edit_workspace_size(n), dev)
        call("gstex_texture_edit", int(tile_bounds[0]), int(tile_bounds[1]), int(block[0]), W, H,
             int(texture_info[0]), int(texture_info[1]), C, X, _ptr(texture_dims),
             _ptr(updated_img), _ptr(updated_alpha), _ptr(depth_lower), _ptr(depth_upper),
             _ptr(gaussian_ids_sorted), gaussian_ids_sorted.numel(), _ptr(tile_bins),
             _ptr(opacities), _ptr(means), _ptr(scales), n, float(glob_scale), _ptr(quats),
             _ptr(uv0), _ptr(umap), _ptr(vmap), _ptr(viewmat), _ptr(c2w), float(fx), float(fy),
             float(cx), float(cy), int(settings), _ptr(background), _ptr(out), _ptr(ws), nb,
             _stream(dev))
    return out
