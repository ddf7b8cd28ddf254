"""The large workloads at full size: 1600x1200 with 1M untextured Gaussians
and the regularisers, 16 orbit views of mixed resolution over 1M Gaussians,
and 8 views at 1920x1080 over 3M Gaussians with bfloat16 storage.

At these sizes the oracle cannot render whole frames in a test, so each case
checks the binning exactly where the oracle's sort finishes in seconds, renders
a sample of tiles (random ones plus the busiest) with the oracle on the GPU's
own sorted lists (float32 for the forward; float64 for gradients and image
PSNR), and asserts size-independent properties of the whole
frame: with unit colours the image equals 1 - T, T stays in (1e-4, 1], the
distortion is non-negative, and the graph replay equals the eager batch.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import (DEV, gpu_backward, gpu_forward, grad_close, log, np_, oracle_with,
                         project_and_bin, to_dev)
from helpers import frac_close, psnr, tile_bounds, view_depths

pytestmark = pytest.mark.gpu

FWD_KEYS = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts")


def _pick_tiles(bins, nt, k_random=12, k_busy=4, seed=0):
    counts = np.diff(np_(bins), axis=1)[:, 0]
    rng = np.random.default_rng(seed)
    busy = np.argsort(counts, kind="stable")[-k_busy:]
    rest = np.setdiff1d(np.arange(nt), busy)
    pick = np.concatenate([rng.choice(rest, k_random, replace=False), busy])
    return np.sort(pick), counts


def _tile_pixels(pick, tiles, H, W, bw=16):
    sel = np.zeros((H, W), bool)
    for t in pick:
        ty, tx = divmod(int(t), tiles[0])
        sel[ty * bw:(ty + 1) * bw, tx * bw:(tx + 1) * bw] = True
    return sel


def _masked_bins(bins, pick):
    """Tile ranges with every tile outside ``pick`` emptied."""
    b = bins.clone()
    keep = torch.zeros(b.shape[0], dtype=torch.bool, device=b.device)
    keep[torch.from_numpy(pick).to(b.device)] = True
    b[~keep] = 0
    return b


def _check_binning(oracle, s, pb):
    c, e = oracle.get_aabb_2d(s.means, s.scales, 1.0, s.quats, s.viewmat, s.fx, s.fy, s.cx, s.cy)
    d = view_depths(s.means, s.viewmat)
    nt = oracle.num_tiles_hit(c, e, s.img_h, s.img_w, s.block_width)
    assert int(nt.sum()) == pb["num_intersects"]
    assert np.array_equal(nt, np_(pb["num_tiles_hit"]))
    _, _, _, ids, bins = oracle.bin_and_sort(c, e, d, nt, tile_bounds(s.img_w, s.img_h, 16), 16)
    assert np.array_equal(ids, np_(pb["ids"]))
    assert np.array_equal(bins, np_(pb["bins"]))


def _frame_invariants(s, pb):
    """Unit colours: out_img = 1 - T; T in (1e-4, 1]; distortion >= 0."""
    ones = np.ones_like(s.colors)
    g, _ = gpu_forward(s, pb["ids"], pb["bins"], colors=ones)
    T = np_(g["final_Ts"])
    np.testing.assert_allclose(np_(g["out_img"]), np.repeat((1 - T)[..., None], 3, -1), atol=3e-5)
    assert np.all((T > 1e-4) & (T <= 1.0))
    assert np.all(np_(g["out_reg"]) >= -1e-6)
    return T


def _sampled_forward_backward(oracle, s, pb, colors, cot, label, fwd_tol=1e-4):
    """Forward of a sample of tiles against the float32 oracle and backward
    against the float64 oracle, both sides walking the GPU's sorted lists with
    the other tiles emptied."""
    tiles = tile_bounds(s.img_w, s.img_h, 16)
    pick, counts = _pick_tiles(pb["bins"], tiles[0] * tiles[1])
    mb = _masked_bins(pb["bins"], pick)
    g, _ = gpu_forward(s, pb["ids"], mb, colors=colors)
    sc = oracle_with(oracle, s, pb["ids"], mb, colors=colors)
    o32 = oracle.texture_forward(sc, "f32")
    sel = _tile_pixels(pick, tiles, s.img_h, s.img_w)
    log(f"{label}: busiest tile {int(counts.max())} entries, sampled "
        f"{int(counts[pick].sum())} entries over {len(pick)} tiles")
    for k in FWD_KEYS:
        fc = frac_close(np_(g[k])[sel], o32[k][sel], fwd_tol, fwd_tol)
        log(f"  {k}: {fc:.5f} within {fwd_tol:g}")
        assert fc >= 0.999, k
    o = oracle.texture_forward(sc, "f64")
    grads = cot(g)
    gg = gpu_backward(s, pb["ids"], mb, g, grads, colors=colors)
    og = oracle.texture_backward(sc, o, grads["img"], grads["depth"], grads["reg"],
                                 grads["alpha"], grads["tex"], grads["normal"], "f64")
    bad = []
    for k in og:
        ok, l2, fr = grad_close(np_(gg[k]), og[k])
        log(f"  {k}: relL2 {l2:.2e} frac {fr:.4f}")
        if not ok:
            bad.append((k, l2, fr))
    assert not bad, bad
    return g, o, sel


def _regularised_cotangents(s, reg_weight=100.0, normal_weight=0.05):
    """Cotangents of L2(out_img + out_texture, target) + 100 mean(out_reg)
    + 0.05 mean(n_x^2 + n_y^2 + (1 - n_z)^2) at the GPU forward."""
    H, W = s.img_h, s.img_w
    P = H * W

    def cot(g):
        img = np_(g["out_img"]) + np_(g["out_texture"])[..., :3]
        diff = (2.0 / (P * 3) * (img - s.target)).astype(np.float32)
        n = np_(g["out_normal"])
        vn = np.empty_like(n)
        w = normal_weight / P
        vn[..., 0] = 2 * w * n[..., 0]
        vn[..., 1] = 2 * w * n[..., 1]
        vn[..., 2] = -2 * w * (1 - n[..., 2])
        z = np.zeros((H, W), np.float32)
        tex = np.zeros((H, W, s.texture.shape[1]), np.float32)
        tex[..., :3] = diff
        return dict(img=diff, depth=z, reg=np.full((H, W), reg_weight / P, np.float32),
                    alpha=z, tex=tex, normal=vn.astype(np.float32))

    return cot


# ------------------------------------------------------------------ 1600x1200, 1M, regularisers
def test_untextured_1m_regularised_full_size(oracle):
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    s = synthetic.make_scene(3)
    log(f"untextured N={s.num_points} {s.img_w}x{s.img_h} X={s.texture.shape[0]}")
    pb = project_and_bin(s)
    log(f"  intersects {pb['num_intersects']}")
    _check_binning(oracle, s, pb)
    _frame_invariants(s, pb)
    colors = np.random.default_rng(3).uniform(0, 1, s.colors.shape).astype(np.float32)
    _sampled_forward_backward(oracle, s, pb, colors, _regularised_cotangents(s), "untextured")

    # the whole-frame regularised step: finite gradients, and the distortion
    # and normal cotangents move the gradients
    step = scene_to_step(s, DEV)
    tgt = to_dev(s.target)
    step.regularised_step(tgt, 100.0, 0.05)
    reg = {k: v.clone() for k, v in step.grads.items()}
    step.step(tgt)
    torch.cuda.synchronize()
    for k, v in reg.items():
        assert torch.isfinite(v).all(), k
    assert not torch.equal(reg["v_means"], step.grads["v_means"])
    assert not torch.equal(reg["v_quats"], step.grads["v_quats"])


# ------------------------------------------------------------------ 16 orbit views, 1M
def test_orbit_16_views_1m_full_size(oracle):
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_multiview

    s = synthetic.make_scene(4)
    views = synthetic.workload_views(4, s)
    log(f"orbit N={s.num_points} X={s.texture.shape[0]} views="
        + ",".join(f"{v.img_w}x{v.img_h}" for v in views))
    colors = np.random.default_rng(4).uniform(0, 1, s.colors.shape).astype(np.float32)
    # exact binning, frame invariants and sampled tiles on the view with the
    # most Gaussians behind the camera and on the largest frame
    behind = [int((view_depths(s.means, v.viewmat) < 0).sum()) for v in views]
    largest = max(range(len(views)), key=lambda i: views[i].img_w * views[i].img_h)
    for i in sorted({int(np.argmax(behind)), largest}):
        vs = synthetic.view_scene(s, views[i])
        pb = project_and_bin(vs)
        log(f"  view {i} {vs.img_w}x{vs.img_h} behind={behind[i]} I={pb['num_intersects']}")
        _check_binning(oracle, vs, pb)
        _frame_invariants(vs, pb)
        _sampled_forward_backward(oracle, vs, pb, colors, _regularised_cotangents(vs, 0.0, 0.0),
                                  f"orbit view {i}")

    # the whole batch as one graph: capacities hold and replay equals the
    # eager batch (the batch gradient as the sum of single-view steps is
    # checked at a smaller size in test_gpu_graph.py)
    mv = scene_to_multiview(s, views, DEV)
    mv.step()
    torch.cuda.synchronize()
    eager = {k: v.clone() for k, v in mv.grads.items()}
    counts = [int(v.intersect_total()) for v in mv.views]
    log(f"  per-view intersects min {min(counts)} max {max(counts)} total {sum(counts)}")
    assert min(counts) * 4 < max(counts)  # uneven work per view
    mv.capture()
    mv.replay_checked()
    for k, v in eager.items():
        assert torch.isfinite(v).all(), k
        ok, l2, fr = grad_close(np_(mv.grads[k]), np_(v), rel_l2=1e-4, elem_tol=1e-4, frac=0.999)
        log(f"  replay {k}: relL2 {l2:.2e}")
        assert ok, (k, l2, fr)


# ------------------------------------------------------------------ 8 x 1080p, 3M, bf16
def _bf16_round(a: np.ndarray) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch.bfloat16).float().numpy()


def test_bf16_8x1080p_3m_full_size(oracle):
    """bf16 texels and SH coefficients, f32 compositing and accumulation, the
    float64 ray/plane solve where |n.ray| < 0.05.  Image PSNR and gradient
    relative error are measured against the float64 oracle render."""
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_multiview

    s = synthetic.make_scene(5)
    views = synthetic.workload_views(5, s)
    log(f"bf16 N={s.num_points} X={s.texture.shape[0]} views={len(views)}")
    prev = _C.set_precise_edge_on(True)
    try:
        mv = scene_to_multiview(s, views, DEV, bf16=True)
        mv.capture()
        mv.replay_checked()
        for k, v in mv.grads.items():
            assert torch.isfinite(v.float()).all(), k
        v0 = mv.views[0]
        log(f"  view 0 intersects {v0.intersect_total()}")
        vs = synthetic.view_scene(s, views[0])
        tiles = tile_bounds(vs.img_w, vs.img_h, 16)
        ids = v0.ids[: v0.num_intersects]
        pick, counts = _pick_tiles(v0.tile_bins, tiles[0] * tiles[1], seed=5)
        sel = _tile_pixels(pick, tiles, vs.img_h, vs.img_w)
        img = np_(v0.out_img) + np_(v0.out_texture)[..., :3]
        # the f64 oracle on the stored (bf16-rounded) values and on the
        # original float32 values
        tex_r, sh_r = _bf16_round(s.texture), _bf16_round(s.sh_coeffs)
        dirs = vs.means - vs.c2w[:3, 3]
        for name, tex, sh in (("stored", tex_r, sh_r), ("float32", s.texture, s.sh_coeffs)):
            cols = oracle.sh_forward(s.sh_degree, s.sh_degree, dirs, sh, "f64")
            sc = oracle_with(oracle, vs, ids, v0.tile_bins, colors=cols)
            sc.texture = tex
            sc.tile_mask = np.zeros(tiles[0] * tiles[1], np.uint8)
            sc.tile_mask[pick] = 1
            o = oracle.texture_forward(sc, "f64")
            ref = o["out_img"] + o["out_texture"][..., :3]
            p = psnr(img[sel], ref[sel], peak=max(float(np.abs(ref[sel]).max()), 1e-6))
            log(f"  PSNR vs f64 oracle ({name} inputs) over {len(pick)} tiles: {p:.1f} dB")
            assert p >= 50.0, (name, p)

        # gradients of the sampled tiles against the f64 oracle on the stored
        # values (the f32 path on bf16-rounded texels equals the bf16 path:
        # bfloat16 widens to float32 exactly)
        mb = _masked_bins(v0.tile_bins, pick)
        cols = np_(v0.colors)
        vr = synthetic.view_scene(s, views[0])
        vr.texture = tex_r
        g, _ = gpu_forward(vr, ids, mb, colors=cols)
        sc = oracle_with(oracle, vr, ids, mb, colors=cols)
        o = oracle.texture_forward(sc, "f64")
        cot = _regularised_cotangents(vr, 0.0, 0.0)(g)
        gg = gpu_backward(vr, ids, mb, g, cot, colors=cols)
        og = oracle.texture_backward(sc, o, cot["img"], cot["depth"], cot["reg"], cot["alpha"],
                                     cot["tex"], cot["normal"], "f64")
        bad = []
        for k in og:
            ok, l2, fr = grad_close(np_(gg[k]), og[k])
            log(f"  grad {k}: relL2 vs f64 oracle {l2:.2e}")
            if not ok:
                bad.append((k, l2, fr))
        assert not bad, bad
    finally:
        _C.set_precise_edge_on(prev)
