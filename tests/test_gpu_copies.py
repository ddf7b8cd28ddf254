"""Runs of copies of one list entry.

The reference counts a tile for every Gaussian whose screen extent is zero
(a mean at view z <= 0.01) but never writes its slots, so they keep Gaussian 0
and sort to the front of tile 0 (forward.cu:31-32 with get_aabb_2d.py:70-92).
The forward passes over the windows of copies that cannot change any pixel of
a wave, and the backward reads the mask words of 64 windows at once; both
must leave every output as the plain walk computes it.  Three cases: Gaussian
0 culled for tile 0, Gaussian 0 covering tile 0 (its copies composite until
the pixels saturate), and Gaussian 0 evaluated at tile 0 with culling off but
alpha below 1/255 everywhere (every copy skipped).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import (DEV, gpu_backward, gpu_forward, grad_close, log, np_, oracle_with,
                         project_and_bin, random_grads)
from helpers import frac_close

pytestmark = pytest.mark.gpu

FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")


def _scene(kind):
    from gstex_cuda_amd import synthetic

    s = synthetic.make_scene(1, seed=21, num_points=4000, img_w=200, img_h=136)
    # 2600 means behind the camera, mirrored into the image: each is counted
    # in one tile and leaves a zero slot
    s.means[1000:3600, 2] = -s.means[1000:3600, 2]
    if kind == "covering":
        # Gaussian 0 over the top-left corner, facing the camera, opaque
        z = 3.0
        s.means[0] = [(8 - s.cx) / s.fx * z, (8 - s.cy) / s.fy * z, z]
        s.quats[0] = [1.0, 0.0, 0.0, 0.0]
        s.scales[0, :2] = 0.25
        s.opacities[0] = 0.95
    if kind == "skipped":
        # Gaussian 0 over the corner with alpha < 1/255 everywhere
        z = 3.0
        s.means[0] = [(8 - s.cx) / s.fx * z, (8 - s.cy) / s.fy * z, z]
        s.quats[0] = [1.0, 0.0, 0.0, 0.0]
        s.scales[0, :2] = 0.25
        s.opacities[0] = 0.003
    return s


@pytest.mark.parametrize("kind", ["culled", "covering", "skipped"])
def test_runs_of_copies_match_the_plain_walk(oracle, kind):
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene(kind)
    prev = _C.set_raster_culling(kind != "skipped")
    try:
        pb = project_and_bin(s)
        ids, bins = np_(pb["ids"]), np_(pb["bins"])
        run = int(np.argmax(ids != 0))
        log(f"{kind}: I={pb['num_intersects']} copies of Gaussian 0 leading tile 0: {run}, "
            f"tile 0 holds {bins[0, 1] - bins[0, 0]}")
        assert bins[0, 0] == 0 and run > 1500
        g, _ = gpu_forward(s, pb["ids"], pb["bins"])
        sc = oracle_with(oracle, s, pb["ids"], pb["bins"])
        o = oracle.texture_forward(sc, "f32")
        for k in FWD_NAMES:
            if k in ("final_idx", "depth_idx"):
                assert np.mean(np_(g[k]) == o[k]) >= 0.999, k
            else:
                fc = frac_close(np_(g[k]), o[k], 1e-5, 1e-5)
                assert fc >= 0.999, (k, fc)
        t0 = np_(g["final_Ts"])[:16, :16]
        log(f"  tile 0: min T {t0.min():.3e}, last entry {int(np_(g['final_idx'])[:16, :16].max())}")
        if kind == "covering":
            assert t0.min() <= 1e-2  # the copies composited

        grads = random_grads(np.random.default_rng(3), s.img_h, s.img_w, s.texture.shape[1], 1e-3)
        gg = gpu_backward(s, pb["ids"], pb["bins"], g, grads)
        o64 = oracle.texture_forward(sc, "f64")
        og = oracle.texture_backward(sc, o64, grads["img"], grads["depth"], grads["reg"],
                                     grads["alpha"], grads["tex"], grads["normal"], "f64")
        for k in og:
            ok, l2, fr = grad_close(np_(gg[k]), og[k])
            assert ok, (k, l2, fr)

        # the per-stage path: wave masks (written over the passed windows too)
        # against the full back-to-front walk
        tgt = torch.from_numpy(s.target).to(DEV)
        a = scene_to_step(s, DEV, use_wave_masks=True)
        b = scene_to_step(s, DEV, use_wave_masks=False)
        a.step(tgt)
        b.step(tgt)
        torch.cuda.synchronize()
        for k in FWD_NAMES:
            assert torch.equal(getattr(a, k), getattr(b, k)), k
        for k in a.grads:
            ok, l2, fr = grad_close(np_(a.grads[k]), np_(b.grads[k]), rel_l2=1e-5,
                                    elem_tol=1e-5, frac=1.0)
            assert ok, (k, l2, fr)
    finally:
        _C.set_raster_culling(prev)
