"""The C-ABI library (CPU only, no kernel launches): it loads, exports every
symbol declared in include/gstex_hip.h, and rejects bad arguments before any
HIP call with the reference's error messages."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gstex_hip.h")
LIB = os.path.join(ROOT, "gstex_cuda_amd", "libgstex_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gstex_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    syms = set(declared_symbols())
    # one replacement per native entry point of the reference (ext.cpp:8-25)
    for name in ("gstex_compute_sh_forward", "gstex_compute_sh_backward",
                 "gstex_map_gaussian_to_intersects", "gstex_get_tile_bin_edges",
                 "gstex_texture_forward", "gstex_texture_backward", "gstex_get_aabb_2d",
                 "gstex_texture_sample_forward", "gstex_texture_sample_backward",
                 "gstex_texture_edit"):
        assert name in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (make -C gstex_cuda_amd/csrc)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b(gstex_[a-z0-9_]+)\b", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    handle = ctypes.CDLL(LIB)
    for s in declared_symbols():
        getattr(handle, s)


def test_python_binding_covers_the_header():
    from gstex_cuda_amd import _lib

    assert set(_lib.exported_symbols()) == set(declared_symbols())
    assert _lib.lib().gstex_abi_version() == 1


def _call_forward(**over):
    from gstex_cuda_amd import _lib

    args = dict(tiles_x=2, tiles_y=2, bw=16, W=32, H=32, charts=1, probs=1, C=3, ws=1 << 20)
    args.update(over)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    rc = _lib.lib().gstex_texture_forward(
        args["tiles_x"], args["tiles_y"], args["bw"], args["W"], args["H"], args["charts"],
        args["probs"], args["C"], fake, fake, 4, fake, fake, fake, fake, fake, 1, 1.0, fake, fake,
        fake, fake, fake, 1, fake, fake, 1.0, 1.0, 1.0, 1.0, args.get("settings", 0), fake,
        fake, fake, fake, fake, fake, fake, fake, fake, fake, fake, args["ws"], None)
    return rc, _lib.lib().gstex_last_error().decode()


@pytest.mark.parametrize("over,msg", [
    (dict(bw=1), "block_width must be between 2 and 16"),
    (dict(bw=17), "block_width must be between 2 and 16"),
    (dict(tiles_x=3), "tile_bounds"),
    (dict(probs=2), "num_probs"),
    (dict(C=65), "texture channels"),
    (dict(ws=8), "workspace too small"),
    (dict(settings=1 << 30), "settings bits 30 and 31"),
])
def test_forward_rejects_bad_arguments(over, msg):
    rc, err = _call_forward(**over)
    assert rc == 1 and msg in err


def test_sh_rejects_degrees_to_use_above_degree():
    from gstex_cuda_amd import _lib

    fake = ctypes.c_void_p(0x1000)
    rc = _lib.lib().gstex_compute_sh_forward(4, 2, 3, fake, fake, fake, None)
    assert rc == 1 and "degrees_to_use" in _lib.lib().gstex_last_error().decode()


def test_python_api_requires_cuda_tensors():
    from gstex_cuda_amd import cuda as _C

    m = torch.zeros(4, 3)
    with pytest.raises(RuntimeError, match="means must be a CUDA tensor"):
        _C.get_aabb_2d(m, m, 1.0, torch.zeros(4, 4), torch.eye(4), 1.0, 1.0, 0.0, 0.0)
    with pytest.raises(RuntimeError, match="viewdirs must be a CUDA tensor"):
        _C.compute_sh_forward(4, 0, 0, m, torch.zeros(4, 1, 3))


def test_texture_gaussians_argument_checks():
    from gstex_cuda_amd.texture import texture_gaussians

    z = torch.zeros(2, 3)
    with pytest.raises(AssertionError, match="block_width"):
        texture_gaussians((2, 1, 3), torch.zeros(2, 3, dtype=torch.int32), z[:, :2], z[:, :2],
                          z[:, 0], torch.zeros(2, dtype=torch.int32), z, z[:, :1], z, z, 1.0,
                          torch.zeros(2, 4), z, z, z, z, torch.eye(4), torch.eye(4), 1, 1, 0, 0,
                          8, 8, 1, 0)


def test_synthetic_scenes_are_deterministic():
    from gstex_cuda_amd.synthetic import make_scene

    a = make_scene(2, num_points=2000)
    b = make_scene(2, num_points=2000)
    for k in ("means", "scales", "quats", "texture", "texture_dims", "sh_coeffs", "target"):
        assert np.array_equal(getattr(a, k), getattr(b, k))
    assert a.texture_dims[:, 0].min() >= 1 and a.texture_dims[:, 0].max() <= 16
    assert np.array_equal(a.texture_dims[1:, 2],
                          np.cumsum(a.texture_dims[:, 0] * a.texture_dims[:, 1])[:-1])


def test_binning_rejects_tile_grids_beyond_the_packing():
    """Tile coordinates travel as 16-bit halves in the binning emission and tile
    indices as ints, so grids of 65536 tiles on a side (or above 2^31 - 1
    tiles) are refused before any launch."""
    from gstex_cuda_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    for tx, ty in ((65536, 1), (1, 70000), (50000, 50000)):
        rc = L.gstex_bin_gaussians(4, 8, fake, fake, fake, fake, tx, ty, 2, fake, fake, fake,
                                   1 << 30, None)
        assert rc == 1 and "tile grid" in L.gstex_last_error().decode(), (tx, ty)
        rc = L.gstex_bin_gaussians_capped(4, 8, fake, fake, fake, fake, tx, ty, 2, fake, fake,
                                          fake, fake, 1 << 30, None)
        assert rc == 1 and "tile grid" in L.gstex_last_error().decode(), (tx, ty)
        rc = L.gstex_map_gaussian_to_intersects(4, 8, fake, fake, fake, fake, tx, ty, 2, 0,
                                                fake, fake, None)
        assert rc == 1 and "tile grid" in L.gstex_last_error().decode(), (tx, ty)
        rc = L.gstex_get_tile_bin_edges(8, fake, tx, ty, fake, None)
        assert rc == 1 and "tile grid" in L.gstex_last_error().decode(), (tx, ty)
    # a raster launch over such a grid (width 2 * 65536 at block width 2)
    rc, err = _call_forward(tiles_x=65536, tiles_y=1, bw=2, W=131072, H=2)
    assert rc == 1 and "tile grid" in err
