"""Shared test helpers: scene construction, oracle binning and comparison metrics."""
from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_kat():
    with open(os.path.join(HERE, "golden", "kat.json")) as fh:
        return json.load(fh)


def view_depths(means: np.ndarray, viewmat: np.ndarray) -> np.ndarray:
    """View-space z of the means in float32 (the binning depth)."""
    m = means.astype(np.float32)
    v = viewmat.astype(np.float32)
    return (v[2, 0] * m[:, 0] + v[2, 1] * m[:, 1] + v[2, 2] * m[:, 2] + v[2, 3]).astype(np.float32)


def tile_bounds(img_w: int, img_h: int, bw: int):
    return ((img_w + bw - 1) // bw, (img_h + bw - 1) // bw, 1)


def oracle_scene(O, s, ids_sorted=None, tile_bins=None, colors: Optional[np.ndarray] = None,
                 settings: Optional[int] = None, block_width: Optional[int] = None):
    """An oracle Scene from a synthetic Scene (binning done by the oracle if absent)."""
    bw = s.block_width if block_width is None else block_width
    sc = O.Scene(means=s.means, scales=s.scales, quats=s.quats,
                 colors=s.colors if colors is None else colors, opacities=s.opacities,
                 uv0=s.uv0, umap=s.umap, vmap=s.vmap, texture=s.texture,
                 texture_dims=s.texture_dims, viewmat=s.viewmat, c2w=s.c2w,
                 background=s.background, fx=s.fx, fy=s.fy, cx=s.cx, cy=s.cy, img_h=s.img_h,
                 img_w=s.img_w, block_width=bw,
                 settings=s.settings if settings is None else settings)
    if ids_sorted is None:
        c, e = O.get_aabb_2d(s.means, s.scales, 1.0, s.quats, s.viewmat, s.fx, s.fy, s.cx, s.cy)
        d = view_depths(s.means, s.viewmat)
        nt = O.num_tiles_hit(c, e, s.img_h, s.img_w, bw)
        _, _, _, ids_sorted, tile_bins = O.bin_and_sort(c, e, d, nt, tile_bounds(s.img_w, s.img_h, bw), bw)
    sc.ids_sorted = ids_sorted
    sc.tile_bins = tile_bins
    return sc


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.linalg.norm(b.ravel()), 1e-30)
    return float(np.linalg.norm((a - b).ravel()) / den)


def frac_close(a, b, atol: float, rtol: float = 0.0) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ok = np.abs(a - b) <= atol + rtol * np.abs(b)
    return float(ok.mean()) if ok.size else 1.0


def camera_facing_scene(means, scales, opacities, colors, W, H, f, cx, cy, texel=None,
                        background=(0.0, 0.0, 0.0), uv_scale=6.0):
    """A scene of identity-rotation splats (normal along the optical axis)."""
    from gstex_cuda_amd.synthetic import Scene

    n = len(means)
    means = np.asarray(means, np.float32).reshape(n, 3)
    scales = np.asarray(scales, np.float32).reshape(n, 3)
    quats = np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1))
    tex = np.asarray(texel if texel is not None else [0.5, 0.5, 0.5], np.float32).reshape(1, 3)
    texture = np.tile(tex, (n, 1)).astype(np.float32)
    dims = np.stack([np.ones(n), np.ones(n), np.arange(n)], 1).astype(np.int32)
    umap = np.zeros((n, 1, 3), np.float32)
    vmap = np.zeros((n, 1, 3), np.float32)
    umap[:, 0, 0] = 1.0 / (uv_scale * scales[:, 0])
    vmap[:, 0, 1] = 1.0 / (uv_scale * scales[:, 1])
    eye = np.eye(4, dtype=np.float32)
    return Scene(means=means, scales=scales, quats=quats,
                 opacities=np.asarray(opacities, np.float32).reshape(n, 1),
                 colors=np.asarray(colors, np.float32).reshape(n, 3),
                 uv0=np.full((n, 1, 2), 0.5, np.float32), umap=umap, vmap=vmap, texture=texture,
                 texture_dims=dims, viewmat=eye.copy(), c2w=eye.copy(),
                 background=np.asarray(background, np.float32), target=np.zeros((H, W, 3), np.float32),
                 fx=f, fy=f, cx=cx, cy=cy, img_h=H, img_w=W, block_width=16, settings=1 << 8)


def textured_splat_scene(k):
    """The one-splat scene of a ``splat_tex`` known-answer case: a 4 x 5 chart
    with explicit uv0 / umap / vmap and quaternion."""
    from gstex_cuda_amd.synthetic import Scene

    f32 = np.float32
    H, W = k["H"], k["W"]
    eye = np.eye(4, dtype=f32)
    return Scene(means=np.array([k["mean"]], f32), scales=np.array([k["scale"]], f32),
                 quats=np.array([k["quat"]], f32), opacities=np.array([[k["opacity"]]], f32),
                 colors=np.array([k["color"]], f32), uv0=np.array([[k["uv0"]]], f32),
                 umap=np.array([[k["umap"]]], f32), vmap=np.array([[k["vmap"]]], f32),
                 texture=np.array(k["texture"], f32), texture_dims=np.array([k["dims"]], np.int32),
                 viewmat=eye.copy(), c2w=eye.copy(), background=np.array(k["background"], f32),
                 target=np.zeros((H, W, 3), f32), fx=k["f"], fy=k["f"], cx=k["cx"], cy=k["cy"],
                 img_h=H, img_w=W, block_width=16, settings=1 << 8)
