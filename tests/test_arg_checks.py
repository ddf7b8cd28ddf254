"""Argument validation of the Python entry points, on the CPU: the dtype and
shape checks run before any kernel, so they are exercised here with the
device check switched off (no GPU call is made)."""
from __future__ import annotations

import pytest
import torch

import gstex_cuda_amd.cuda as G


@pytest.fixture
def no_device_check(monkeypatch):
    monkeypatch.setattr(G, "_check", lambda t, name: None)


def _sample_args(q=6, C=3, X=20):
    dims = torch.tensor([[2, 2, 0]] * q, dtype=torch.int32)
    uvs = torch.rand(q, 2)
    tex = torch.rand(X, C)
    return (q, 1, C), dims, uvs, tex


def test_sample_checks_accept_valid(no_device_check):
    info, dims, uvs, tex = _sample_args()
    assert G._sample_checks(info, dims, uvs, tex) == 6
    assert G._sample_checks(info, dims, uvs, tex, torch.rand(6, 3)) == 6


@pytest.mark.parametrize("bad", ["bf16_texture", "f64_uvs", "i64_dims", "short_dims",
                                 "dims_cols", "texture_width", "v_shape", "v_dtype", "uvs_cols"])
def test_sample_checks_reject(no_device_check, bad):
    info, dims, uvs, tex = _sample_args()
    v = torch.rand(6, 3)
    if bad == "bf16_texture":
        tex = tex.to(torch.bfloat16)
    elif bad == "f64_uvs":
        uvs = uvs.double()
    elif bad == "i64_dims":
        dims = dims.long()
    elif bad == "short_dims":
        dims = dims[:5]
    elif bad == "dims_cols":
        dims = dims[:, :2].contiguous()
    elif bad == "texture_width":
        tex = torch.rand(20, 2)
    elif bad == "v_shape":
        v = torch.rand(6, 4)
    elif bad == "v_dtype":
        v = v.to(torch.bfloat16)
    elif bad == "uvs_cols":
        uvs = torch.rand(6, 3)
    with pytest.raises(RuntimeError):
        G._sample_checks(info, dims, uvs, tex, v)


def test_texture_width_must_match_channels():
    G._texture_shape(torch.zeros(5, 3), 3)
    G._texture_shape(torch.zeros(0, 3), 3)
    with pytest.raises(RuntimeError, match="texture_info"):
        G._texture_shape(torch.zeros(5, 2), 3)
    with pytest.raises(RuntimeError, match="texture_info"):
        G._texture_shape(torch.zeros(15), 3)
