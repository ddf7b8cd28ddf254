"""The per-stage pipeline (RenderStep over the C ABI) on the GPU.

* Wave masks: a backward that visits only the entries the forward composited
  for each wave gives the same gradients as the full back-to-front walk, and
  the masks leave the forward untouched.
* The per-stage path agrees with the oracle (same binning) for the forward
  outputs and every parameter gradient, with random output cotangents.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, grad_close, log, np_, random_grads
from helpers import psnr

pytestmark = pytest.mark.gpu

SET_UV = 1 << 8
SET_BLUR = 1 << 9
SET_NDC = 1 << 10

# name -> (workload, num_points, W, H, block_width, settings)
CASES = {
    "w1_small": (1, 300, 96, 64, 16, SET_UV),
    "w1_long_lists": (1, 6000, 160, 128, 16, SET_UV),
    "w1_bw8": (1, 2000, 100, 72, 8, SET_UV),
    "w1_bw5": (1, 1500, 67, 41, 5, SET_UV),
    "w1_blur_ndc": (1, 3000, 128, 96, 16, SET_UV | SET_BLUR | SET_NDC),
    "w3_one_texel": (3, 4000, 160, 120, 16, SET_UV),
}

GRAD_NAMES = ("v_colors", "v_opacity", "v_means", "v_scales", "v_quats", "v_uv0", "v_umap",
              "v_vmap", "v_texture")
FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")


def _scene(case):
    from gstex_cuda_amd import synthetic

    wl, n, W, H, bw, st = CASES[case]
    spec = dataclasses.replace(synthetic.WORKLOADS[wl], block_width=bw, settings=st,
                               sh_degree=-1, texel_budget=0,
                               texel_side=1 if wl == 3 else 4)
    return synthetic.make_scene(wl, seed=10 + sum(map(ord, case)), num_points=n, img_w=W,
                                img_h=H, spec=spec)


def _run(s, masks, cot):
    from gstex_cuda_amd.pipeline import scene_to_step

    step = scene_to_step(s, DEV, use_wave_masks=masks)
    step.forward()
    step.v_img.copy_(torch.from_numpy(cot["img"]).to(DEV))
    step.v_tex.copy_(torch.from_numpy(cot["tex"]).to(DEV))
    step.backward(v_depth=torch.from_numpy(cot["depth"]).to(DEV),
                  v_reg=torch.from_numpy(cot["reg"]).to(DEV),
                  v_alpha=torch.from_numpy(cot["alpha"]).to(DEV),
                  v_normal=torch.from_numpy(cot["normal"]).to(DEV))
    torch.cuda.synchronize()
    fwd = {k: np_(getattr(step, k)) for k in FWD_NAMES}
    grads = {k: np_(v) for k, v in step.grads.items()}
    return step, fwd, grads


@pytest.mark.parametrize("case", list(CASES))
def test_wave_masks_match_full_walk(case):
    s = _scene(case)
    log(f"case {case}: N={s.num_points} {s.img_w}x{s.img_h} bw={s.block_width}")
    cot = random_grads(np.random.default_rng(7), s.img_h, s.img_w, s.texture.shape[1])
    step, fa, ga = _run(s, True, cot)
    _, fb, gb = _run(s, False, cot)
    log(f"  intersects {step.num_intersects}")
    for k in FWD_NAMES:
        assert np.array_equal(fa[k], fb[k]), k
    for k in GRAD_NAMES:
        # float atomics from different waves land in any order
        ok, l2, fr = grad_close(ga[k], gb[k], rel_l2=1e-5, elem_tol=1e-5, frac=1.0)
        assert ok, (k, l2, fr)


@pytest.mark.parametrize("case", ["w1_long_lists", "w1_bw5", "w1_blur_ndc", "w3_one_texel"])
def test_pipeline_matches_oracle(oracle, case):
    O = oracle
    s = _scene(case)
    log(f"case {case}")
    cot = random_grads(np.random.default_rng(8), s.img_h, s.img_w, s.texture.shape[1])
    step, fwd, grads = _run(s, True, cot)
    sc = O.Scene(means=s.means, scales=s.scales, quats=s.quats, colors=s.colors,
                 opacities=s.opacities, uv0=s.uv0, umap=s.umap, vmap=s.vmap, texture=s.texture,
                 texture_dims=s.texture_dims, viewmat=s.viewmat, c2w=s.c2w,
                 background=s.background, fx=s.fx, fy=s.fy, cx=s.cx, cy=s.cy, img_h=s.img_h,
                 img_w=s.img_w, block_width=s.block_width, settings=s.settings,
                 ids_sorted=np_(step.ids[: step.num_intersects]), tile_bins=np_(step.tile_bins))
    f = O.texture_forward(sc, "f64")
    for k in ("out_img", "out_texture", "out_normal"):
        err = np.abs(fwd[k] - f[k])
        assert psnr(fwd[k], f[k]) >= 50.0 or np.mean(err < 1e-4) >= 0.999, k
    g = O.texture_backward(sc, f, cot["img"], cot["depth"], cot["reg"], cot["alpha"], cot["tex"],
                           cot["normal"], "f64")
    res = {k: grad_close(grads[k], g[k].reshape(grads[k].shape)) for k in GRAD_NAMES}
    log("  rel L2 vs f64 oracle: " + " ".join(f"{k}={r[1]:.2e}" for k, r in res.items()))
    for k, (ok, l2, fr) in res.items():
        assert ok, (k, l2, fr)


def test_wave_masks_full_size_1080p_500k():
    """1920x1080 with 500k Gaussians: masked and full walks agree."""
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    s = synthetic.make_scene(2)
    tgt = torch.from_numpy(s.target).to(DEV)
    out = []
    for masks in (True, False):
        step = scene_to_step(s, DEV, use_wave_masks=masks)
        step.step(tgt)
        torch.cuda.synchronize()
        out.append({k: np_(v) for k, v in step.grads.items()})
        log(f"masks={masks}: {step.num_intersects} intersects")
    for k in GRAD_NAMES + ("v_sh_coeffs",):
        ok, l2, fr = grad_close(out[0][k], out[1][k], rel_l2=1e-4, elem_tol=1e-4, frac=0.9999)
        assert ok, (k, l2, fr)


@pytest.mark.parametrize("case", ["w1_small", "w1_bw5", "w1_blur_ndc"])
def test_backward_l2_equals_separate_loss_pass(case):
    # gstex_raster_backward_l2 (the L2 cotangents formed in the backward, as
    # RenderStep.step does for three channels) gives the gradients of
    # gstex_image_loss_grad followed by gstex_raster_backward: the same
    # per-pixel cotangents bit for bit, so the same sums up to the order of
    # the float atomics
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene(case)
    tgt = torch.from_numpy(np.ascontiguousarray(s.target, dtype=np.float32)).to(DEV)
    ref = scene_to_step(s, DEV)
    assert ref._fused_loss_target(tgt) is tgt
    ref.forward()
    ref.l2_loss_grad(tgt)
    ref.backward()
    fused = scene_to_step(s, DEV)
    fused.step(tgt)
    torch.cuda.synchronize()
    for k in FWD_NAMES:
        assert np.array_equal(np_(getattr(fused, k)), np_(getattr(ref, k))), k
    for k in GRAD_NAMES:
        ok, l2, fr = grad_close(np_(fused.grads[k]), np_(ref.grads[k]), rel_l2=1e-5,
                                elem_tol=1e-5, frac=0.9999)
        assert ok, (k, l2, fr)
    # and the raw per-Gaussian gradient records agree to that order
    rec_ref = np_(ref.grad_records)
    rec_fused = np_(fused.grad_records)
    assert np.allclose(rec_fused, rec_ref, rtol=1e-5, atol=1e-7 * np.abs(rec_ref).max())


def test_backward_l2_rejects_other_channel_counts():
    from gstex_cuda_amd import _lib

    lib = _lib.lib()
    dummy = torch.zeros(64, device=DEV)
    p = dummy.data_ptr()
    rc = lib.gstex_raster_backward_l2(32, 32, 16, 4, 0, p, 1, p, 1, p, p, 16, 0, p, p, 1.0, 1.0,
                                      16.0, 16.0, p, p, p, p, p, p, p, p, 1.0, p, p, p, p, p,
                                      p, None, 0, None)
    assert rc != 0 and b"3 channels" in lib.gstex_last_error()


@pytest.mark.parametrize("degree,num_points", [(3, 3001), (1, 700), (0, 300)])
def test_finalize_sh_equals_separate_passes(degree, num_points):
    # gstex_finalize_gradients_sh (RenderStep's finalize with SH colours) gives
    # gstex_finalize_gradients followed by gstex_compute_sh_backward at the
    # view directions means - camera position, bit for bit
    from gstex_cuda_amd import _lib, synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    spec = dataclasses.replace(synthetic.WORKLOADS[1], sh_degree=degree, texel_budget=0,
                               texel_side=4)
    s = synthetic.make_scene(1, seed=70 + degree, num_points=num_points, img_w=96, img_h=64,
                             spec=spec)
    tgt = torch.from_numpy(np.ascontiguousarray(s.target, dtype=np.float32)).to(DEV)
    step = scene_to_step(s, DEV)
    assert step.sh_degree == degree
    step.step(tgt)
    torch.cuda.synchronize()
    got = {k: np_(v).copy() for k, v in step.grads.items()}
    lib = _lib.lib()
    p, cam, n = step.p, step.cam, step.n
    ref = {k: torch.full_like(v, 7.0) for k, v in step.grads.items()}
    st = torch.cuda.current_stream().cuda_stream
    ptr = lambda t: t.data_ptr()
    assert lib.gstex_finalize_gradients(
        n, ptr(step.grad_records), ptr(step.records), ptr(p["quats"]), ptr(p["scales"]),
        ptr(cam.viewmat), ptr(cam.c2w), cam.fx, cam.fy, ptr(ref["v_colors"]),
        ptr(ref["v_opacity"]), ptr(ref["v_means"]), ptr(ref["v_scales"]), ptr(ref["v_quats"]),
        ptr(ref["v_uv0"]), ptr(ref["v_umap"]), ptr(ref["v_vmap"]), st) == 0
    dirs = (p["means"] - cam.c2w[:3, 3]).contiguous()
    assert lib.gstex_compute_sh_backward(n, degree, degree, ptr(dirs), ptr(ref["v_colors"]),
                                         ptr(ref["v_sh_coeffs"]), st) == 0
    torch.cuda.synchronize()
    for k in ("v_colors", "v_opacity", "v_means", "v_scales", "v_quats", "v_uv0", "v_umap",
              "v_vmap", "v_sh_coeffs"):
        assert np.array_equal(got[k], np_(ref[k])), k
