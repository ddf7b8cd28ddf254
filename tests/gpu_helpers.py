"""GPU-side helpers for the parity tests: run the HIP pipeline on a synthetic
scene and hand the same inputs (and binning) to the oracle."""
from __future__ import annotations

import sys

import numpy as np
import torch

from helpers import oracle_scene, tile_bounds

DEV = torch.device("cuda:0")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def np_(t):
    return t.detach().cpu().numpy()


def project_and_bin(s, bw=None):
    """GPU projection + depth-first binning; returns host arrays."""
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.utils import compute_cumulative_intersects

    bw = s.block_width if bw is None else bw
    means, scales, quats, vm = (to_dev(s.means), to_dev(s.scales), to_dev(s.quats),
                                to_dev(s.viewmat))
    c, e, d, nt = _C.project_gaussians(means, scales, 1.0, quats, vm, s.fx, s.fy, s.cx, s.cy,
                                       s.img_h, s.img_w, bw)
    m, _ = compute_cumulative_intersects(nt)
    tb = tile_bounds(s.img_w, s.img_h, bw)
    if m > 0:
        ids, bins = _C.bin_gaussians(s.num_points, m, c, e, d, nt, tb, bw)
    else:
        ids = torch.zeros(0, dtype=torch.int32, device=DEV)
        bins = torch.zeros((tb[0] * tb[1], 2), dtype=torch.int32, device=DEV)
    return dict(centers=c, extents=e, depths=d, num_tiles_hit=nt, num_intersects=m, ids=ids,
                bins=bins)


def raster_inputs(s, colors=None):
    col = s.colors if colors is None else colors
    return dict(texture_dims=to_dev(s.texture_dims), colors=to_dev(col),
                opacities=to_dev(s.opacities), means=to_dev(s.means), scales=to_dev(s.scales),
                quats=to_dev(s.quats), uv0=to_dev(s.uv0), umap=to_dev(s.umap),
                vmap=to_dev(s.vmap), texture=to_dev(s.texture), viewmat=to_dev(s.viewmat),
                c2w=to_dev(s.c2w), background=to_dev(s.background))


def gpu_forward(s, ids, bins, settings=None, bw=None, colors=None):
    from gstex_cuda_amd import cuda as _C

    bw = s.block_width if bw is None else bw
    st = s.settings if settings is None else settings
    r = raster_inputs(s, colors)
    tb = tile_bounds(s.img_w, s.img_h, bw)
    out = _C.texture_forward(tb, (bw, bw, 1), (s.img_w, s.img_h, 1),
                             (s.num_points, 1, s.texture.shape[1]), r["texture_dims"], ids, bins,
                             r["colors"], r["opacities"], r["means"], r["scales"], 1.0, r["quats"],
                             r["uv0"], r["umap"], r["vmap"], r["texture"], r["viewmat"], r["c2w"],
                             s.fx, s.fy, s.cx, s.cy, st, r["background"])
    names = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")
    return dict(zip(names, out)), r


def gpu_backward(s, ids, bins, fwd, grads, settings=None, bw=None, colors=None):
    from gstex_cuda_amd import cuda as _C

    bw = s.block_width if bw is None else bw
    st = s.settings if settings is None else settings
    r = raster_inputs(s, colors)
    g = [to_dev(np.asarray(grads[k], np.float32)) for k in
         ("img", "depth", "reg", "alpha", "tex", "normal")]
    out = _C.texture_backward(s.img_h, s.img_w, bw, (s.num_points, 1, s.texture.shape[1]),
                              r["texture_dims"], ids, bins, r["colors"], r["opacities"],
                              r["means"], r["scales"], 1.0, r["quats"], r["uv0"], r["umap"],
                              r["vmap"], r["texture"], r["viewmat"], r["c2w"], s.fx, s.fy, s.cx,
                              s.cy, st, r["background"], fwd["final_Ts"], fwd["final_idx"],
                              fwd["depth_idx"], fwd["out_reg_s"], *g)
    names = ("v_colors", "v_opacity", "v_means", "v_scales", "v_quats", "v_uv0", "v_umap",
             "v_vmap", "v_texture")
    return dict(zip(names, out))


def oracle_with(O, s, ids, bins, settings=None, bw=None, colors=None):
    return oracle_scene(O, s, ids_sorted=np_(ids), tile_bins=np_(bins), colors=colors,
                        settings=settings, block_width=bw)


def random_grads(rng, H, W, C, scale=1.0):
    return dict(img=scale * rng.standard_normal((H, W, 3)).astype(np.float32),
                depth=scale * rng.standard_normal((H, W)).astype(np.float32),
                reg=scale * rng.standard_normal((H, W)).astype(np.float32),
                alpha=scale * rng.standard_normal((H, W)).astype(np.float32),
                tex=scale * rng.standard_normal((H, W, C)).astype(np.float32),
                normal=scale * rng.standard_normal((H, W, 3)).astype(np.float32))


def grad_close(a, b, rel_l2=2e-3, elem_tol=2e-3, frac=0.99):
    """Relative-L2 and per-element closeness scaled by the group's max magnitude."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    scale = max(np.abs(b).max(), 1e-12)
    l2 = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
    ok = np.abs(a - b) <= elem_tol * scale
    return l2 <= rel_l2 and ok.mean() >= frac, float(l2), float(ok.mean())
