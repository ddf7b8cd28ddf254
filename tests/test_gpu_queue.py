"""The multi-view raster pass (gstex_views_raster_forward / _backward): one
forward and one backward launch over every tile of every view, taken by
persistent waves from one queue ordered longest list first.

Each tile is rendered by the same per-wave code as the per-view kernels, so
every view's forward outputs must equal the per-view batch bit for bit; the
gradients are the same sums in another order (atomics), so they must agree to
float32 rounding.  Covered: views of mixed resolution with uneven lists, the
general kernels (blur + NDC), bfloat16 storage, and graph capture replayed
several times with eager work in between.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, grad_close, log, np_

pytestmark = pytest.mark.gpu

FWD_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")


def _batch(settings=None, bf16=False, n=6000, resolutions=None, queue=True, streams=4):
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_multiview

    s = synthetic.make_scene(2, num_points=n, img_w=240, img_h=136)
    if settings is not None:
        s.settings = settings
    views = synthetic.make_views(s, 5, seed=3, resolutions=resolutions)
    return s, scene_to_multiview(s, views, DEV, bf16=bf16, queue=queue, streams=streams)


def _compare(a, b, rel=2e-5):
    for i, (va, vb) in enumerate(zip(a.views, b.views)):
        for k in FWD_NAMES:
            assert np.array_equal(np_(getattr(va, k)), np_(getattr(vb, k))), (i, k)
    for k in a.grads:
        ok, l2, fr = grad_close(np_(a.grads[k].float()), np_(b.grads[k].float()), rel_l2=rel,
                                elem_tol=rel, frac=0.9999)
        log(f"  {k}: relL2 {l2:.2e}")
        assert ok, (k, l2, fr)


@pytest.mark.parametrize("case", ["plain", "mixed_res", "blur_ndc", "bf16"])
def test_queue_pass_equals_per_view_batch(case):
    kw = {}
    if case == "mixed_res":
        kw["resolutions"] = [(96, 64), (240, 136), (160, 160), (200, 90)]
    if case == "blur_ndc":
        kw["settings"] = (1 << 8) | (1 << 9) | (1 << 10)
    if case == "bf16":
        kw["bf16"] = True
    _, ref = _batch(queue=False, **kw)
    _, q = _batch(queue=True, **kw)
    ref.step()
    q.step()
    torch.cuda.synchronize()
    log(case, "intersects", [int(v.intersect_total()) for v in q.views])
    assert q.capacity_ok()
    _compare(q, ref)


def test_queue_pass_graph_replays():
    _, ref = _batch(queue=False)
    _, q = _batch(queue=True)
    ref.step()
    q.capture()
    for r in range(3):
        q.replay_checked()
        # eager work on the same stream between replays
        x = torch.randn(1 << 20, device=DEV)
        torch.cuda.synchronize()
        assert float((x * x).sum()) > 0
        _compare(q, ref)
        log(f"replay {r} matches")


def test_queue_pass_one_stream():
    _, ref = _batch(queue=False, streams=1)
    _, q = _batch(queue=True, streams=1)
    ref.step()
    q.step()
    torch.cuda.synchronize()
    _compare(q, ref)
