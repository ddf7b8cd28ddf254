"""bfloat16 storage of texels and SH coefficients.

The kernels widen bfloat16 to float32 exactly and compute in float32, so a
render from bfloat16 texels / SH coefficients must equal, bit for bit, the
float32 render of the same values rounded to bfloat16; gradients agree up to
the order of float atomics.  Against the float64 oracle on the unrounded
float32 inputs the bfloat16 render keeps the error the rounding of the inputs
implies (checked as an image PSNR floor).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, grad_close, log, np_
from helpers import psnr

pytestmark = pytest.mark.gpu

FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")
GRAD_NAMES = ("v_colors", "v_opacity", "v_means", "v_scales", "v_quats", "v_uv0", "v_umap",
              "v_vmap", "v_texture", "v_sh_coeffs")


def _bf16_round(a: np.ndarray) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.bfloat16).float().numpy()


def _steps(s):
    """(bf16-stored step, float32 step on the bf16-rounded values)."""
    from gstex_cuda_amd.pipeline import RenderStep, scene_to_step

    ref = scene_to_step(s, DEV)
    p32 = dict(ref.p)
    p32["texture"] = torch.from_numpy(_bf16_round(s.texture)).to(DEV)
    p32["sh_coeffs"] = torch.from_numpy(_bf16_round(s.sh_coeffs)).to(DEV)
    pbf = dict(ref.p)
    pbf["texture"] = torch.from_numpy(s.texture).to(DEV).to(torch.bfloat16)
    pbf["sh_coeffs"] = torch.from_numpy(s.sh_coeffs).to(DEV).to(torch.bfloat16)
    kw = dict(block_width=s.block_width, settings=s.settings, background=ref.background,
              sh_degree=s.sh_degree)
    return RenderStep(pbf, ref.cam, **kw), RenderStep(p32, ref.cam, **kw)


def _run(step, tgt):
    step.step(tgt)
    torch.cuda.synchronize()
    return ({k: np_(getattr(step, k)) for k in FWD_NAMES},
            {k: np_(v) for k, v in step.grads.items()})


@pytest.mark.parametrize("full", [False, True], ids=["small", "1080p_500k"])
def test_bf16_storage_equals_rounded_f32(full):
    from gstex_cuda_amd import synthetic

    s = synthetic.make_scene(2) if full else synthetic.make_scene(2, num_points=4000, img_w=192,
                                                                    img_h=108)
    log(f"N={s.num_points} {s.img_w}x{s.img_h}")
    tgt = torch.from_numpy(s.target).to(DEV)
    sb, sf = _steps(s)
    assert sb.texture_format == 1 and sf.texture_format == 0
    fb, gb = _run(sb, tgt)
    ff, gf = _run(sf, tgt)
    for k in FWD_NAMES:
        assert np.array_equal(fb[k], ff[k]), k
    for k in GRAD_NAMES:
        ok, l2, fr = grad_close(gb[k], gf[k], rel_l2=1e-4, elem_tol=1e-4, frac=0.9999)
        assert ok, (k, l2, fr)


def test_bf16_render_against_f64_oracle(oracle):
    """Image error of bf16 storage vs the f64 oracle on the float32 inputs."""
    from gstex_cuda_amd import synthetic

    s = synthetic.make_scene(2, num_points=3000, img_w=160, img_h=90)
    tgt = torch.from_numpy(s.target).to(DEV)
    sb, _ = _steps(s)
    fb, _ = _run(sb, tgt)
    O = oracle
    colors = O.sh_forward(s.sh_degree, s.sh_degree, s.means - s.c2w[:3, 3], s.sh_coeffs, "f64")
    sc = O.Scene(means=s.means, scales=s.scales, quats=s.quats, colors=colors,
                 opacities=s.opacities, uv0=s.uv0, umap=s.umap, vmap=s.vmap, texture=s.texture,
                 texture_dims=s.texture_dims, viewmat=s.viewmat, c2w=s.c2w,
                 background=s.background, fx=s.fx, fy=s.fy, cx=s.cx, cy=s.cy, img_h=s.img_h,
                 img_w=s.img_w, block_width=s.block_width, settings=s.settings,
                 ids_sorted=np_(sb.ids[: sb.num_intersects]), tile_bins=np_(sb.tile_bins))
    f = O.texture_forward(sc, "f64")
    img = fb["out_img"] + fb["out_texture"][..., :3]
    ref = f["out_img"] + f["out_texture"][..., :3]
    p = psnr(img, ref, peak=max(float(np.abs(ref).max()), 1e-6))
    log(f"bf16 storage PSNR vs f64 oracle: {p:.1f} dB")
    # bf16 keeps 8 significant bits: relative input error <= 2^-9
    assert p >= 40.0, p
