"""The reference's import path resolves unchanged (CPU test, no GPU calls).

Every module and name example.py:9-14 and the reference's tests
(tests/test_sample.py:4-5, tests/test_sh.py:12-13) import is importable from
``gstex_cuda``; the pure-torch render path it deliberately leaves out raises
with a documented message.
"""
from __future__ import annotations

import importlib

import numpy as np
import pytest
import torch

IMPORTED = {
    "gstex_cuda.texture": ["texture_gaussians"],
    "gstex_cuda.get_aabb_2d": ["get_aabb_2d", "get_num_tiles_hit_2d", "project_points",
                               "merge_aabbs"],
    "gstex_cuda._torch_impl": ["normalized_quat_to_rotmat", "quat_to_rotmat", "project_pix"],
    "gstex_cuda.timer": ["Timer"],
    "gstex_cuda.texture_sample": ["texture_sample"],
    "gstex_cuda.sh": ["spherical_harmonics", "num_sh_bases", "deg_from_sh"],
    "gstex_cuda.texture_edit": ["texture_edit"],
    "gstex_cuda.utils": ["get_tile_bin_edges", "compute_cumulative_intersects",
                         "map_gaussian_to_intersects", "bin_and_sort_gaussians"],
    # cuda/__init__.py:14-26 of the reference
    "gstex_cuda.cuda": ["compute_sh_forward", "compute_sh_backward",
                        "map_gaussian_to_intersects", "get_tile_bin_edges", "get_aabb_2d",
                        "texture_forward", "texture_backward", "texture_sample_forward",
                        "texture_sample_backward", "texture_edit"],
}

ABSENT = {
    "gstex_cuda._torch_impl": ["texture_forward", "sample_texture", "compute_sh_color",
                               "eval_sh_bases"],
    "gstex_cuda.get_aabb_2d": ["get_aabb_2d_torch"],
}


@pytest.mark.parametrize("module", sorted(IMPORTED))
def test_reference_names_import(module):
    m = importlib.import_module(module)
    for name in IMPORTED[module]:
        assert callable(getattr(m, name)), f"{module}.{name}"


@pytest.mark.parametrize("module", sorted(ABSENT))
def test_torch_render_path_is_documented_absent(module):
    m = importlib.import_module(module)
    for name in ABSENT[module]:
        with pytest.raises(AttributeError, match="pure-torch"):
            getattr(m, name)
        with pytest.raises(ImportError):
            exec(f"from {module} import {name}", {})


def test_same_objects_as_the_package():
    import gstex_cuda.cuda
    import gstex_cuda.texture
    import gstex_cuda_amd.cuda
    import gstex_cuda_amd.texture

    assert gstex_cuda.texture.texture_gaussians is gstex_cuda_amd.texture.texture_gaussians
    assert gstex_cuda.cuda.texture_forward is gstex_cuda_amd.cuda.texture_forward


def test_use_torch_impl_raises():
    from gstex_cuda.texture import texture_gaussians

    z = torch.zeros
    with pytest.raises(NotImplementedError):
        texture_gaussians((1, 1, 3), z((1, 3), dtype=torch.int32), z(1, 2), z(1, 2), z(1),
                          z(1, dtype=torch.int32), z(1, 3), z(1, 1), z(1, 3), z(1, 3), 1.0,
                          z(1, 4), z(1, 1, 2), z(1, 1, 3), z(1, 1, 3), z(1, 3), torch.eye(4),
                          torch.eye(4), 1.0, 1.0, 0.0, 0.0, 16, 16, 16, 0,
                          use_torch_impl=True)


def test_rotmat_helpers():
    from scipy.spatial.transform import Rotation

    from gstex_cuda._torch_impl import normalized_quat_to_rotmat, project_pix, quat_to_rotmat

    rng = np.random.default_rng(3)
    q = rng.normal(size=(64, 4))
    qn = q / np.linalg.norm(q, axis=1, keepdims=True)
    # scipy takes (x, y, z, w); the reference layout is (w, x, y, z)
    want = Rotation.from_quat(qn[:, [1, 2, 3, 0]]).as_matrix()
    got = normalized_quat_to_rotmat(torch.from_numpy(qn)).numpy()
    np.testing.assert_allclose(got, want, atol=1e-12)
    got2 = quat_to_rotmat(torch.from_numpy(q * 3.0)).numpy()
    np.testing.assert_allclose(got2, want, atol=1e-12)
    # 90 degrees about z: ax1 = (0, 1, 0)
    c = np.cos(np.pi / 4)
    r = normalized_quat_to_rotmat(torch.tensor([c, 0.0, 0.0, c], dtype=torch.float64)).numpy()
    np.testing.assert_allclose(r[:, 0], [0.0, 1.0, 0.0], atol=1e-15)
    p = project_pix((100.0, 50.0), torch.tensor([[1.0, -2.0, 4.0]], dtype=torch.float64),
                    (10.0, 20.0)).numpy()
    rw = 1.0 / (4.0 + 1e-6)
    np.testing.assert_allclose(p, [[100.0 * rw + 10.0, -100.0 * rw + 20.0]], rtol=1e-12)


def test_timer_host_mode(capsys):
    from gstex_cuda.timer import Timer

    t = Timer(cuda_support=False)
    t.start("a")
    t.stop()
    t.start("b")
    t.stop()
    t.lap()
    t.dump()
    out = capsys.readouterr().out.splitlines()
    assert [line.split("\t")[0] for line in out] == ["a", "b", "TOTAL"]
    assert t.laps == 1
    assert Timer(disabled=True).dump() is None
