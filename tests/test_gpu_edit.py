"""texture_edit on the GPU against the oracle's restatement of texture_edit.cu."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, log, np_, project_and_bin, to_dev
from helpers import oracle_scene, tile_bounds

pytestmark = pytest.mark.gpu

# name -> (workload, num_points, W, H, block_width, edit settings (bit 0 = blur), channels)
CASES = {
    "w1": (1, 800, 96, 64, 16, 0, 5),
    "w1_blur": (1, 800, 96, 64, 16, 1, 5),
    "w1_bw8": (1, 1500, 80, 56, 8, 0, 5),
    "w2_small": (2, 3000, 160, 90, 16, 0, 5),
    "w3_one_texel": (3, 2000, 160, 120, 16, 0, 5),  # every chart 1x1
    # C = 8: the LDS route's texel index (byte offset / 4C) and flush stride,
    # and the direct route's row stride, at a width other than 5; the mixed
    # 1x1..16x16 charts of workload 2 take both routes (h w 8 <= 1024 or not)
    "w2_c8": (2, 3000, 160, 90, 16, 0, 8),
    "w1_c8_bw8": (1, 1500, 80, 56, 8, 0, 8),
    # charts of 2 or 3 texels a side: most pairs take the shared-quad and
    # 3x3-window routes (wave reduce-scatters over 20 and 45 values)
    "w2_small_charts": (2, 3000, 160, 90, 16, 0, 5),
}


def _case(name):
    from gstex_cuda_amd import synthetic

    wl, n, W, H, bw, st, C = CASES[name]
    spec = dataclasses.replace(synthetic.WORKLOADS[wl], block_width=bw)
    s = synthetic.make_scene(wl, seed=20 + sum(map(ord, name)), num_points=n, img_w=W, img_h=H,
                             spec=spec)
    rng = np.random.default_rng(sum(map(ord, name)))
    if name.endswith("small_charts"):
        hw = rng.integers(2, 4, (s.num_points, 2))
        start = np.concatenate([[0], np.cumsum(hw[:, 0] * hw[:, 1])[:-1]])
        s.texture_dims = np.stack([hw[:, 0], hw[:, 1], start], 1).astype(np.int32)
        s.texture = np.zeros((int((hw[:, 0] * hw[:, 1]).sum()), C), np.float32)
    s.texture = np.zeros((s.texture.shape[0], C), np.float32)
    img = rng.random((H, W, 3)).astype(np.float32)
    al = (rng.random((H, W)) * (rng.random((H, W)) < 0.7)).astype(np.float32)
    lo = rng.uniform(0.0, 4.0, (H, W)).astype(np.float32)
    hi = (lo + rng.uniform(0.5, 20.0, (H, W))).astype(np.float32)
    return s, st, img, al, lo, hi


@pytest.mark.parametrize("name", list(CASES))
def test_texture_edit_matches_oracle(oracle, name):
    from gstex_cuda_amd import cuda as _C

    s, st, img, al, lo, hi = _case(name)
    log(f"case {name}")
    b = project_and_bin(s)
    X, C = s.texture.shape
    tb = tile_bounds(s.img_w, s.img_h, s.block_width)
    if C > 5:
        hw = s.texture_dims[:, 0].astype(np.int64) * s.texture_dims[:, 1]
        log(f"  charts on the LDS route {int((hw * C <= 1024).sum())}, "
            f"direct {int((hw * C > 1024).sum())}")
    out = _C.texture_edit(tb, (s.block_width, s.block_width, 1), (s.img_w, s.img_h, 1),
                          (s.num_points, 1, C), X, to_dev(s.texture_dims), to_dev(img),
                          to_dev(al), to_dev(lo), to_dev(hi), b["ids"], b["bins"],
                          to_dev(s.opacities), to_dev(s.means), to_dev(s.scales), 1.0,
                          to_dev(s.quats), to_dev(s.uv0), to_dev(s.umap), to_dev(s.vmap),
                          to_dev(s.viewmat), to_dev(s.c2w), s.fx, s.fy, s.cx, s.cy, st,
                          to_dev(s.background))
    torch.cuda.synchronize()
    got = np_(out)
    sc = oracle_scene(oracle, s, ids_sorted=np_(b["ids"]), tile_bins=np_(b["bins"]), settings=st)
    ref = oracle.texture_edit(sc, img, al, lo, hi, "f64")
    log(f"  weight total {ref[:, 4].sum():.1f}")
    assert ref[:, 4].sum() > 0
    assert got.shape == (X, C)
    # the edit writes channels 0..4 only; the rest of the output stays zero
    assert not np.any(got[:, 5:]), "channels past 4 were written"
    # float atomics in any order and float32 pair math: relative to each channel's scale
    for k in range(5):
        scale = max(np.abs(ref[:, k]).max(), 1e-12)
        err = np.abs(got[:, k] - ref[:, k]) / scale
        assert np.mean(err < 1e-4) >= 0.999 and err.max() < 5e-2, (k, float(err.max()))
        np.testing.assert_allclose(got[:, k].sum(), ref[:, k].sum(), rtol=2e-4)


def test_texture_edit_driver_and_argument_checks():
    from gstex_cuda_amd.texture_edit import texture_edit

    s, st, img, al, lo, hi = _case("w1")
    b = project_and_bin(s)
    out = texture_edit((s.num_points, 1, 5), to_dev(s.texture_dims), to_dev(img), to_dev(al),
                       to_dev(lo), to_dev(hi), b["centers"], b["extents"], b["depths"],
                       b["num_tiles_hit"], to_dev(s.opacities), to_dev(s.means),
                       to_dev(s.scales), 1.0, to_dev(s.quats), to_dev(s.uv0), to_dev(s.umap),
                       to_dev(s.vmap), to_dev(s.viewmat), to_dev(s.c2w), s.fx, s.fy, s.cx, s.cy,
                       s.img_h, s.img_w, s.block_width, st)
    assert out.shape == (s.texture.shape[0], 5)
    assert float(out[:, 4].sum()) > 0
    with pytest.raises(Exception, match="5 channels"):
        texture_edit((s.num_points, 1, 3), to_dev(s.texture_dims), to_dev(img), to_dev(al),
                     to_dev(lo), to_dev(hi), b["centers"], b["extents"], b["depths"],
                     b["num_tiles_hit"], to_dev(s.opacities), to_dev(s.means), to_dev(s.scales),
                     1.0, to_dev(s.quats), to_dev(s.uv0), to_dev(s.umap), to_dev(s.vmap),
                     to_dev(s.viewmat), to_dev(s.c2w), s.fx, s.fy, s.cx, s.cy, s.img_h, s.img_w,
                     s.block_width, st)
