"""Per-wave culling by screen boxes and ellipses leaves every raster output unchanged.

The raster kernels skip a list entry for a wave when the Gaussian's screen box,
or (blur off) its screen ellipse, misses all of the wave's pixels.  These tests
render scenes with culling on and off and require identical forward outputs
(bit for bit) and backward gradients (up to the order of float atomics), on
random scenes and on scenes built to stress the bounds: edge-on and grazing
splats, splats reaching behind the camera, thin splats at every in-plane
angle, sub-pixel splats, extreme and sub-threshold opacities, zero scales,
unnormalised quaternions, and saturated pixels that end on a near-invisible
Gaussian.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, grad_close, log, np_, random_grads

pytestmark = pytest.mark.gpu

FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")
GRAD_NAMES = ("v_colors", "v_opacity", "v_means", "v_scales", "v_quats", "v_uv0", "v_umap",
              "v_vmap", "v_texture")


def _rebuild_uv_maps(s, uv_scale=6.0):
    from gstex_cuda_amd.synthetic import quat_axes_np

    a1, a2, _ = quat_axes_np(s.quats)
    s.umap = (a1 / (uv_scale * s.scales[:, :1].astype(np.float64))).astype(np.float32).reshape(
        -1, 1, 3)
    s.vmap = (a2 / (uv_scale * s.scales[:, 1:2].astype(np.float64))).astype(np.float32).reshape(
        -1, 1, 3)


def _scene(kind, settings=1 << 8):
    from gstex_cuda_amd import synthetic

    spec = dataclasses.replace(synthetic.WORKLOADS[1], settings=settings)
    s = synthetic.make_scene(1, seed=100 + sum(map(ord, kind)), num_points=4000, img_w=144,
                             img_h=112, spec=spec)
    rng = np.random.default_rng(5)
    n = s.num_points
    pick = rng.random(n) < 0.25
    if kind == "edge_on":
        # a3 perpendicular to the view direction of the mean (plane contains the ray)
        m = s.means.astype(np.float64)
        d = m / np.linalg.norm(m, axis=1, keepdims=True)
        a3 = np.cross(d, np.array([0.0, 1.0, 0.0]))
        a3 /= np.linalg.norm(a3, axis=1, keepdims=True)
        a3 += 1e-3 * rng.standard_normal(a3.shape)  # nearly edge-on
        a3 /= np.linalg.norm(a3, axis=1, keepdims=True)
        z = np.array([0.0, 0.0, 1.0])
        axis = np.cross(z, a3)
        sn = np.linalg.norm(axis, axis=1, keepdims=True)
        ang = np.arctan2(sn[:, 0], a3 @ z)
        axis = axis / np.maximum(sn, 1e-12)
        q = np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], 1)
        s.quats[pick] = q[pick].astype(np.float32)
        _rebuild_uv_maps(s)
    elif kind == "near_camera":
        s.means[pick, 2] = rng.uniform(0.02, 0.6, pick.sum()).astype(np.float32)
        s.scales[pick, :2] *= 4.0
        _rebuild_uv_maps(s)
    elif kind == "opacity_extremes":
        vals = np.array([0.0, 0.001, 1 / 255, 0.0039, 0.0040, 0.5, 0.99, 1.0, 1.5], np.float32)
        s.opacities[:, 0] = vals[rng.integers(0, len(vals), n)]
    elif kind == "degenerate":
        s.scales[pick & (rng.random(n) < 0.5), 0] = 0.0
        s.quats[pick] *= 1.3
    elif kind == "thin_rotated":
        s.scales[pick, 1] = s.scales[pick, 0] * rng.uniform(0.005, 0.05, pick.sum())
        s.scales[pick, 0] *= 3.0
        _rebuild_uv_maps(s)
    elif kind == "grazing":
        # |a3 . view direction| between 0.02 and 0.3
        m = s.means.astype(np.float64)
        d = m / np.linalg.norm(m, axis=1, keepdims=True)
        side = np.cross(d, rng.standard_normal(d.shape))
        side /= np.linalg.norm(side, axis=1, keepdims=True)
        c = rng.uniform(0.02, 0.3, n)[:, None] * np.sign(rng.standard_normal((n, 1)))
        a3 = c * d + np.sqrt(1 - c * c) * side
        z = np.array([0.0, 0.0, 1.0])
        axis = np.cross(z, a3)
        sn = np.linalg.norm(axis, axis=1, keepdims=True)
        ang = np.arctan2(sn[:, 0], a3 @ z)
        axis = axis / np.maximum(sn, 1e-12)
        q = np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], 1)
        s.quats[pick] = q[pick].astype(np.float32)
        s.scales[pick, :2] *= 2.0
        _rebuild_uv_maps(s)
    elif kind == "tiny":
        s.scales[pick, :2] *= rng.uniform(0.02, 0.2, (pick.sum(), 1)).astype(np.float32)
        s.opacities[pick, 0] = rng.uniform(0.3, 1.0, pick.sum()).astype(np.float32)
        _rebuild_uv_maps(s)
    elif kind == "saturated":
        s.opacities[:, 0] = 0.99
        s.scales[:, :2] *= 3.0
        s.opacities[rng.random(n) < 0.3, 0] = 0.004  # near-invisible Gaussians in between
        _rebuild_uv_maps(s)
    return s


def _run(s, cull, masks, cot):
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.pipeline import scene_to_step

    prev = _C.set_raster_culling(cull)
    try:
        step = scene_to_step(s, DEV, use_wave_masks=masks)
        step.forward()
        step.v_img.copy_(torch.from_numpy(cot["img"]).to(DEV))
        step.v_tex.copy_(torch.from_numpy(cot["tex"]).to(DEV))
        step.backward(v_depth=torch.from_numpy(cot["depth"]).to(DEV),
                      v_reg=torch.from_numpy(cot["reg"]).to(DEV),
                      v_alpha=torch.from_numpy(cot["alpha"]).to(DEV),
                      v_normal=torch.from_numpy(cot["normal"]).to(DEV))
        torch.cuda.synchronize()
    finally:
        _C.set_raster_culling(prev)
    return ({k: np_(getattr(step, k)) for k in FWD_NAMES},
            {k: np_(v) for k, v in step.grads.items()}, step.num_intersects)


def _grads_match(a, b, rel):
    """Same non-finite positions, and the finite values within ``rel``."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    fa, fb = np.isfinite(a), np.isfinite(b)
    if not np.array_equal(fa, fb) or not np.array_equal(np.isnan(a), np.isnan(b)):
        return False, "non-finite pattern differs"
    ok, l2, fr = grad_close(a[fa], b[fa], rel_l2=rel, elem_tol=rel, frac=1.0)
    return ok, (l2, fr)


KINDS = ["random", "edge_on", "near_camera", "opacity_extremes", "degenerate", "saturated",
         "thin_rotated", "grazing", "tiny"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("settings", [1 << 8, (1 << 8) | (1 << 9) | (1 << 10)])
def test_culling_changes_nothing(kind, settings):
    s = _scene(kind, settings)
    log(f"case {kind} settings={settings}")
    cot = random_grads(np.random.default_rng(3), s.img_h, s.img_w, s.texture.shape[1])
    f_on, g_on, m = _run(s, True, True, cot)
    f_off, g_off, _ = _run(s, False, True, cot)
    _, g_on_nm, _ = _run(s, True, False, cot)
    log(f"  intersects {m}")
    for k in FWD_NAMES:
        assert np.array_equal(f_on[k], f_off[k], equal_nan=True), k
    for k in GRAD_NAMES:
        for other in (g_off, g_on_nm):
            ok, info = _grads_match(g_on[k], other[k], 1e-5)
            assert ok, (k, info)


def test_culling_full_size_1080p_500k():
    """1920x1080 with 500k Gaussians: identical forward with culling on and off."""
    from gstex_cuda_amd import synthetic

    s = synthetic.make_scene(2)
    cot = random_grads(np.random.default_rng(4), s.img_h, s.img_w, s.texture.shape[1], 1e-3)
    f_on, g_on, _ = _run(s, True, True, cot)
    f_off, g_off, _ = _run(s, False, True, cot)
    for k in FWD_NAMES:
        assert np.array_equal(f_on[k], f_off[k], equal_nan=True), k
    for k in GRAD_NAMES + ("v_sh_coeffs",):
        ok, l2, fr = grad_close(g_on[k], g_off[k], rel_l2=1e-4, elem_tol=1e-4, frac=0.9999)
        assert ok, (k, l2, fr)
