"""Capacity-mode binning, graph capture of a whole step, and multi-view batches.

* A capacity-mode step (intersect total kept on the device, no host sync)
  equals the eager step; an undersized capacity is reported, not overrun.
* A captured step replayed several times -- with parameter updates and other
  eager GPU work between replays -- matches an eager step on the same
  parameters after every replay.
* A multi-view batch (shared parameters, bfloat16 storage optional) sums the
  per-view gradients of single-view steps, eagerly and as one replayed graph.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, grad_close, log, np_

pytestmark = pytest.mark.gpu

FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")


def _scene(full=False):
    from gstex_cuda_amd import synthetic

    if full:
        return synthetic.make_scene(2)
    return synthetic.make_scene(2, num_points=6000, img_w=240, img_h=136)


def _fwd(step):
    return {k: np_(getattr(step, k)) for k in FWD_NAMES}


def _grads(step_or_grads):
    g = step_or_grads if isinstance(step_or_grads, dict) else step_or_grads.grads
    return {k: np_(v) for k, v in g.items()}


def _assert_grads(a, b, rel=1e-4):
    for k in a:
        ok, l2, fr = grad_close(a[k], b[k], rel_l2=rel, elem_tol=rel, frac=0.9999)
        assert ok, (k, l2, fr)


@pytest.mark.parametrize("full", [False, True], ids=["small", "1080p_500k"])
def test_capacity_mode_equals_eager(full):
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene(full)
    tgt = torch.from_numpy(s.target).to(DEV)
    eager = scene_to_step(s, DEV)
    eager.step(tgt)
    torch.cuda.synchronize()
    m = eager.num_intersects
    capped = scene_to_step(s, DEV, capacity=int(m * 1.1) + 100)
    capped.step(tgt)
    torch.cuda.synchronize()
    log(f"m={m} cap={capped.cap} device total={capped.intersect_total()}")
    assert capped.intersect_total() == m and capped.capacity_ok()
    fa, fb = _fwd(eager), _fwd(capped)
    for k in FWD_NAMES:
        assert np.array_equal(fa[k], fb[k]), k
    _assert_grads(_grads(eager), _grads(capped))


def test_capacity_overflow_is_reported():
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene()
    tgt = torch.from_numpy(s.target).to(DEV)
    small = scene_to_step(s, DEV, capacity=1000)
    small.step(tgt)
    torch.cuda.synchronize()
    assert small.intersect_total() > 1000 and not small.capacity_ok()
    assert np.isfinite(np_(small.out_img)).all()


def test_graph_replays_match_eager():
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene()
    tgt = torch.from_numpy(s.target).to(DEV)
    probe = scene_to_step(s, DEV)
    probe.step(tgt)
    torch.cuda.synchronize()
    cap = int(probe.num_intersects * 1.3) + 1000
    g_step = scene_to_step(s, DEV, capacity=cap)
    g_step.capture(tgt)
    eager = scene_to_step(s, DEV)
    other = scene_to_step(s, DEV)  # unrelated eager work between replays
    for r in range(3):
        g_step.replay()
        other.step(tgt)
        eager.step(tgt)
        torch.cuda.synchronize()
        assert g_step.capacity_ok()
        fa, fb = _fwd(eager), _fwd(g_step)
        for k in FWD_NAMES:
            assert np.array_equal(fa[k], fb[k]), (r, k)
        _assert_grads(_grads(eager), _grads(g_step))
        log(f"replay {r} ok")
        # in-place parameter update seen by the next replay (shared tensors)
        with torch.no_grad():
            g_step.p["opacities"].mul_(0.97)
            eager.p["opacities"].copy_(g_step.p["opacities"])


@pytest.mark.parametrize("bf16,streams", [(False, 4), (True, 4), (False, 1)],
                         ids=["f32", "bf16", "f32_one_stream"])
def test_multiview_batch_sums_single_views(bf16, streams):
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import RenderStep, scene_to_multiview

    s = _scene()
    views = synthetic.make_views(s, 4, resolutions=[(240, 136), (160, 120), (200, 100)])
    mv = scene_to_multiview(s, views, DEV, bf16=bf16, streams=streams)
    assert sorted(i for i, _ in mv.assign) == list(range(len(views)))
    mv.step()
    torch.cuda.synchronize()
    assert mv.capacity_ok()
    want = None
    for v, tgt in zip(mv.views, mv.targets):
        single = RenderStep(v.p, v.cam, block_width=v.bw, settings=v.settings,
                            background=v.background, sh_degree=v.sh_degree)
        single.step(tgt)
        torch.cuda.synchronize()
        g = _grads(single)
        want = g if want is None else {k: want[k] + g[k] for k in want}
    _assert_grads(want, _grads(mv.grads))
    mv.capture()
    for r in range(2):
        mv.replay_checked()
        torch.cuda.synchronize()
        _assert_grads(want, _grads(mv.grads))
        log(f"multiview replay {r} ok")


def test_autotune_picks_a_config_and_caches_it(tmp_path):
    from gstex_cuda_amd import autotune
    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd.pipeline import scene_to_step

    s = _scene()
    tgt = torch.from_numpy(s.target).to(DEV)
    base = scene_to_step(s, DEV)
    path = str(tmp_path / "tune.json")
    cfg = autotune.tune(base.p, base.cam, tgt, settings=base.settings,
                        background=base.background, sh_degree=base.sh_degree,
                        cache_path=path)
    log(f"tuned: {cfg} {autotune.last_timings()}")
    assert cfg.block_width == 16
    again = autotune.tune(base.p, base.cam, tgt, settings=base.settings,
                          background=base.background, sh_degree=base.sh_degree,
                          cache_path=path)
    assert again == cfg
    prev = _C.set_raster_culling(True)
    try:
        step = autotune.tuned_step(base.p, base.cam, tgt, settings=base.settings,
                                   background=base.background, sh_degree=base.sh_degree)
        step.step(tgt)
        base.step(tgt)
        torch.cuda.synchronize()
        for k in FWD_NAMES:
            assert np.array_equal(np_(getattr(step, k)), np_(getattr(base, k))), k
    finally:
        _C.set_raster_culling(prev)
