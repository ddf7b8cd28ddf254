"""Charts that do not lie inside the texture.

The reference indexes texels start + i*w + j with no check, so a chart whose
texture_dims row points outside the (X, C) texture reads and writes out of
bounds.  Here such a chart reads as zero and receives no texel gradient.  The
test compares a scene with broken charts (past the end, negative start, zero
height, one texel short) against the same scene whose broken Gaussians use an
appended 1x1 all-zero chart instead: every forward output must be identical,
and every gradient must match (the zero chart's own texel gradient aside).
Settings include the uv gradient (bit 8), so tap differences are covered too.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from gpu_helpers import (DEV, gpu_backward, gpu_forward, grad_close, log, np_, project_and_bin,
                         random_grads)

pytestmark = pytest.mark.gpu


def _scenes():
    from gstex_cuda_amd import synthetic

    s = synthetic.make_scene(1, seed=71, num_points=1500, img_w=128, img_h=96)
    X = s.texture.shape[0]
    rng = np.random.default_rng(5)
    bad = rng.choice(s.num_points, 120, replace=False)
    broken = s.texture_dims.copy()
    for k, g in enumerate(bad):
        kind = k % 4
        if kind == 0:
            broken[g] = (2, 3, X + 7)        # past the end
        elif kind == 1:
            broken[g] = (2, 2, -3)           # negative start
        elif kind == 2:
            broken[g] = (0, 4, 0)            # zero height
        else:
            broken[g] = (2, 2, X - 3)        # one texel short
    fixed = s.texture_dims.copy()
    fixed[bad] = (1, 1, X)
    a = dataclasses.replace(s, texture_dims=broken)
    b = dataclasses.replace(s, texture_dims=fixed, texture=np.concatenate(
        [s.texture, np.zeros((1, s.texture.shape[1]), np.float32)]))
    return a, b, X


def test_charts_outside_texture_read_zero():
    a, b, X = _scenes()
    pb = project_and_bin(a)
    log(f"intersects {pb['num_intersects']}")
    fa, _ = gpu_forward(a, pb["ids"], pb["bins"])
    fb, _ = gpu_forward(b, pb["ids"], pb["bins"])
    for k in fa:
        assert np.array_equal(np_(fa[k]), np_(fb[k])), k
    assert np.abs(np_(fa["out_texture"])).max() > 0
    grads = random_grads(np.random.default_rng(2), a.img_h, a.img_w, a.texture.shape[1])
    ga = gpu_backward(a, pb["ids"], pb["bins"], fa, grads)
    gb = gpu_backward(b, pb["ids"], pb["bins"], fb, grads)
    for k in ga:
        x, y = np_(ga[k]), np_(gb[k])
        if k == "v_texture":
            assert x.shape[0] == X and np.isfinite(x).all()
            y = y[:X]
        ok, l2, fr = grad_close(x, y, rel_l2=1e-5, elem_tol=1e-5, frac=1.0)
        assert ok, (k, l2, fr)
