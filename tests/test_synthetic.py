"""The seeded workload generators: camera frames and texel budgets."""
from __future__ import annotations

import numpy as np

from gstex_cuda_amd import synthetic


def test_orbit_views_look_at_the_origin():
    views = synthetic.make_orbit_views(16, radius=8.0, seed=4)
    again = synthetic.make_orbit_views(16, radius=8.0, seed=4)
    assert len(views) == 16
    res = set()
    for v, w in zip(views, again):
        assert np.array_equal(v.viewmat, w.viewmat) and np.array_equal(v.target, w.target)
        R = v.c2w[:3, :3].astype(np.float64)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-6)
        assert abs(np.linalg.det(R) - 1.0) < 1e-5
        np.testing.assert_allclose(v.viewmat.astype(np.float64) @ v.c2w, np.eye(4), atol=1e-5)
        eye = v.c2w[:3, 3].astype(np.float64)
        assert 8.0 * 0.95 - 1e-4 <= np.linalg.norm(eye) <= 8.0 * 1.05 + 1e-4
        # the origin lies in front, within 0.1 * sqrt(3) of the optical axis
        o = v.viewmat.astype(np.float64) @ np.array([0.0, 0.0, 0.0, 1.0])
        assert o[2] > 7.0
        assert np.hypot(o[0], o[1]) <= 0.1 * np.sqrt(3) + 1e-5
        assert v.fx == v.fy == float(v.img_h)
        assert (v.cx, v.cy) == (v.img_w / 2.0, v.img_h / 2.0)
        assert v.target.shape == (v.img_h, v.img_w, 3)
        res.add((v.img_w, v.img_h))
    assert res <= set(synthetic.ORBIT_RESOLUTIONS) and len(res) >= 3
    # the directions spread over the sphere
    dirs = np.stack([v.c2w[:3, 3] / np.linalg.norm(v.c2w[:3, 3]) for v in views])
    assert dirs[:, 1].max() > 0.8 and dirs[:, 1].min() < -0.8


def test_texel_budget_rule():
    # side = clamp(round(sqrt(budget * area_i / sum(area))), 1, 16)
    spec = synthetic.SceneSpec(num_points=2000, img_w=320, img_h=180, focal=180.0,
                               z_range=(2.0, 20.0), texel_budget=20_000)
    s = synthetic.make_scene(seed=9, spec=spec)
    side = s.texture_dims[:, 0]
    assert np.array_equal(side, s.texture_dims[:, 1])
    assert side.min() >= 1 and side.max() <= 16
    start = np.concatenate([[0], np.cumsum(side.astype(np.int64) ** 2)[:-1]])
    assert np.array_equal(s.texture_dims[:, 2], start)
    assert s.texture.shape[0] == int((side.astype(np.int64) ** 2).sum())
    # a larger projected area never gets a smaller grid
    a1, a2, _ = synthetic.quat_axes_np(s.quats)
    area = synthetic._projected_area(s.means, s.scales[:, :2].astype(np.float64), a1, a2,
                                     spec.focal)
    order = np.argsort(area, kind="stable")
    assert np.all(np.diff(side[order]) >= 0)


def test_view_scene_shares_parameters():
    s = synthetic.make_scene(1, num_points=50)
    v = synthetic.make_orbit_views(1, seed=1)[0]
    vs = synthetic.view_scene(s, v)
    assert vs.means is s.means and vs.texture is s.texture
    assert vs.viewmat is v.viewmat and (vs.img_w, vs.img_h) == (v.img_w, v.img_h)
