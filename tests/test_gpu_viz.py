"""Alpha visualisation mode (settings bit 16 and its fields) on the GPU.

The forward draws splats with alpha 0.99 inside their sigma bound, outlined in
white or cut out near the edge (texture.cu:56-65,192-241); it is compared with
the oracle's restatement.  The backward ignores these bits, as the reference
backward does (texture.cu:381-387 reads only bits 2, 8, 9, 10): given the same
forward outputs its gradients equal those of the same settings without them.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import (DEV, gpu_backward, gpu_forward, grad_close, log, np_, oracle_with,
                         project_and_bin, random_grads)

pytestmark = pytest.mark.gpu

BASE = 1 << 8
VIZ = {
    "bound2.5": BASE | (1 << 16) | (20 << 17),
    "bound3_outline1.5_white": BASE | (1 << 16) | (24 << 17) | (6 << 26) | (1 << 24),
    "bound2_outline1_cut_opac": BASE | (1 << 16) | (16 << 17) | (4 << 26) | (1 << 25),
    "bound2.5_flip_white": BASE | (1 << 15) | (1 << 16) | (20 << 17) | (5 << 26) | (1 << 24),
}


def _scene():
    from gstex_cuda_amd import synthetic

    return synthetic.make_scene(1, seed=83, num_points=1200, img_w=128, img_h=96)


@pytest.mark.parametrize("name", list(VIZ))
def test_visualisation_forward_matches_oracle(oracle, name):
    s = _scene()
    st = VIZ[name]
    pb = project_and_bin(s)
    log(f"case {name} settings={st:#x} I={pb['num_intersects']}")
    fwd, _ = gpu_forward(s, pb["ids"], pb["bins"], settings=st)
    sc = oracle_with(oracle, s, pb["ids"], pb["bins"], settings=st)
    f = oracle.texture_forward(sc, "f32")
    for k in ("out_img", "out_texture", "out_normal", "final_Ts"):
        err = np.abs(np_(fwd[k]) - f[k])
        err = err.reshape(err.shape[0], err.shape[1], -1).max(-1)
        frac = float(np.mean(err < 1e-4))
        log(f"  {k}: {frac:.5f} of pixels within 1e-4")
        # hard 0 / 0.99 alphas and outline distances flip only for pixels on a
        # float32 decision boundary
        assert frac >= 0.999, k
    # the mode draws something, and differently from the plain render
    plain, _ = gpu_forward(s, pb["ids"], pb["bins"], settings=BASE)
    assert not torch.equal(fwd["out_img"], plain["out_img"])


def test_visualisation_backward_ignores_the_bits():
    s = _scene()
    st = VIZ["bound3_outline1.5_white"]
    pb = project_and_bin(s)
    fwd, _ = gpu_forward(s, pb["ids"], pb["bins"], settings=st)
    grads = random_grads(np.random.default_rng(4), s.img_h, s.img_w, s.texture.shape[1])
    g_viz = gpu_backward(s, pb["ids"], pb["bins"], fwd, grads, settings=st)
    g_base = gpu_backward(s, pb["ids"], pb["bins"], fwd, grads, settings=st & 0xffff)
    for k in g_viz:
        ok, l2, fr = grad_close(np_(g_viz[k]), np_(g_base[k]), rel_l2=1e-5, elem_tol=1e-5,
                                frac=1.0)
        assert ok, (k, l2, fr)


def test_visualisation_pipeline_backward_skips_forward_masks():
    # the per-stage path hands the forward's wave masks to the backward; with
    # the visualisation bits they describe other entries and are not used
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene()
    s.settings = VIZ["bound2.5_flip_white"]
    tgt = torch.from_numpy(s.target).to(DEV)
    a = scene_to_step(s, DEV, use_wave_masks=True)
    b = scene_to_step(s, DEV, use_wave_masks=False)
    a.step(tgt)
    b.step(tgt)
    torch.cuda.synchronize()
    assert torch.equal(a.out_img, b.out_img)
    for k in a.grads:
        ok, l2, fr = grad_close(np_(a.grads[k]), np_(b.grads[k]), rel_l2=1e-5, elem_tol=1e-5,
                                frac=1.0)
        assert ok, (k, l2, fr)
