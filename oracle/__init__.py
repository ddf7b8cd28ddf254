"""CPU oracle for the textured 2DGS rasteriser — TEST INFRASTRUCTURE ONLY.

ctypes glue over ``oracle/_build/libgstex_oracle.so`` (built from
``gstex_oracle.c`` by ``oracle/Makefile``).  Every function restates the
reference semantics cited in the C source.  ``precision`` selects the FP32
restatement (reference operation order) or the FP64 evaluation.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, as the checker; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgstex_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _in(a, precision: str) -> np.ndarray:
    """Input array in the precision's input type (float32 / float64)."""
    return np.ascontiguousarray(np.asarray(a, dtype=_real(precision)[0]))


def _real(precision: str):
    if precision == "f32":
        return np.float32, "_f32"
    if precision == "f64":
        return np.float64, "_f64"
    raise ValueError(precision)


class _Scene(C.Structure):
    _fields_ = [
        ("img_w", C.c_int), ("img_h", C.c_int), ("block_width", C.c_int),
        ("channels", C.c_int), ("settings", C.c_int),
        ("glob_scale", C.c_double), ("fx", C.c_double), ("fy", C.c_double),
        ("cx", C.c_double), ("cy", C.c_double),
        ("means", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
        ("colors", C.c_void_p), ("opacities", C.c_void_p), ("uv0", C.c_void_p),
        ("umap", C.c_void_p), ("vmap", C.c_void_p), ("texture", C.c_void_p),
        ("viewmat", C.c_void_p), ("c2w", C.c_void_p), ("background", C.c_void_p),
        ("texture_dims", C.c_void_p), ("ids_sorted", C.c_void_p), ("tile_bins", C.c_void_p),
        ("tile_mask", C.c_void_p),
    ]


@dataclass
class Scene:
    """All inputs of one texture_forward / texture_backward call (numpy, float32/int32)."""

    means: np.ndarray
    scales: np.ndarray
    quats: np.ndarray
    colors: np.ndarray
    opacities: np.ndarray
    uv0: np.ndarray
    umap: np.ndarray
    vmap: np.ndarray
    texture: np.ndarray
    texture_dims: np.ndarray
    viewmat: np.ndarray
    c2w: np.ndarray
    background: np.ndarray
    fx: float
    fy: float
    cx: float
    cy: float
    img_h: int
    img_w: int
    block_width: int = 16
    settings: int = 1 << 8
    glob_scale: float = 1.0
    ids_sorted: Optional[np.ndarray] = None
    tile_bins: Optional[np.ndarray] = None
    tile_mask: Optional[np.ndarray] = None  # uint8 per tile; None = every tile
    _keep: list = field(default_factory=list, repr=False)

    _FLOAT_FIELDS = ("means", "scales", "quats", "colors", "opacities", "uv0", "umap", "vmap",
                     "texture", "viewmat", "c2w", "background")

    def normalise(self) -> "Scene":
        self.texture_dims = _i32(self.texture_dims)
        if self.ids_sorted is not None:
            self.ids_sorted = _i32(self.ids_sorted)
        if self.tile_bins is not None:
            self.tile_bins = _i32(self.tile_bins)
        if self.tile_mask is not None:
            self.tile_mask = np.ascontiguousarray(self.tile_mask, np.uint8)
        return self

    @property
    def channels(self) -> int:
        return int(self.texture.shape[1]) if self.texture.ndim == 2 else 0

    def _struct(self, precision: str) -> _Scene:
        """ctypes view of the scene; float inputs converted to the precision's type
        (the converted arrays are kept alive on ``self._keep``)."""
        self.normalise()
        s = _Scene()
        s.img_w, s.img_h, s.block_width = self.img_w, self.img_h, self.block_width
        s.channels, s.settings = self.channels, self.settings
        s.glob_scale, s.fx, s.fy, s.cx, s.cy = self.glob_scale, self.fx, self.fy, self.cx, self.cy
        self._keep = []
        for name in self._FLOAT_FIELDS:
            arr = _in(getattr(self, name), precision)
            self._keep.append(arr)
            setattr(s, name, _p(arr))
        for name in ("texture_dims", "ids_sorted", "tile_bins", "tile_mask"):
            setattr(s, name, _p(getattr(self, name)))
        return s


# --------------------------------------------------------------------------- projection
def get_aabb_2d(means, scales, glob_scale, quats, viewmat, fx, fy, cx, cy, precision="f32"):
    """get_aabb_2d.cu:11-89 -> (centers (N,2), extents (N,2))."""
    real, sfx = _real(precision)
    means, scales, quats, viewmat = (_in(means, precision), _in(scales, precision),
                                     _in(quats, precision), _in(viewmat, precision))
    n = means.shape[0]
    centers = np.zeros((n, 2), real)
    extents = np.zeros((n, 2), real)
    getattr(lib(), "orc_get_aabb_2d" + sfx)(
        C.c_int(n), _p(means), _p(scales), C.c_double(glob_scale), _p(quats), _p(viewmat),
        C.c_double(fx), C.c_double(fy), C.c_double(cx), C.c_double(cy), _p(centers), _p(extents))
    return centers, extents


def num_tiles_hit(centers, extents, img_h, img_w, block_width, wrapped=False):
    """get_aabb_2d.py:70-92 in float32 -> (N,) int32; ``wrapped``: tile count of
    the wrapped rectangle (helpers.cuh:53-73,94-111) instead."""
    centers, extents = _f32(centers), _f32(extents)
    n = centers.shape[0]
    out = np.zeros(n, np.int32)
    if wrapped:
        lib().orc_num_tiles_hit_wrapped(C.c_int(n), _p(centers), _p(extents),
                                        C.c_int(block_width), _p(out))
        return out
    lib().orc_num_tiles_hit(C.c_int(n), _p(centers), _p(extents), C.c_int(img_h), C.c_int(img_w),
                            C.c_int(block_width), _p(out))
    return out


# --------------------------------------------------------------------------- binning
def map_gaussian_to_intersects(centers, extents, depths, cum_tiles_hit, tile_bounds, block_width,
                               num_intersects=None, wrapped=False):
    """forward.cu:13-71 -> (isect_ids int64, gaussian_ids int32), zero-initialised;
    ``wrapped``: forward.cu:34-62 with writes bounded by each Gaussian's window."""
    centers, extents, depths = _f32(centers), _f32(extents), _f32(depths)
    cum = _i32(cum_tiles_hit)
    n = centers.shape[0]
    m = int(cum[-1]) if num_intersects is None else int(num_intersects)
    keys = np.zeros(m, np.int64)
    ids = np.zeros(m, np.int32)
    fn = (lib().orc_map_gaussian_to_intersects_wrapped if wrapped
          else lib().orc_map_gaussian_to_intersects)
    fn(C.c_int(n), C.c_int(m), _p(centers), _p(extents), _p(depths), _p(cum),
       C.c_int(tile_bounds[0]), C.c_int(tile_bounds[1]), C.c_int(block_width), _p(keys), _p(ids))
    return keys, ids


def sort_pairs(keys, ids):
    """utils.py:159-160 with a stable order -> (keys_sorted, ids_sorted)."""
    keys = np.ascontiguousarray(keys, np.int64)
    ids = _i32(ids)
    ko = np.empty_like(keys)
    io = np.empty_like(ids)
    lib().orc_sort_pairs(C.c_int(keys.shape[0]), _p(keys), _p(ids), _p(ko), _p(io))
    return ko, io


def tile_bin_edges(keys_sorted, num_tiles):
    """forward.cu:76-98 -> (T, 2) int32."""
    keys_sorted = np.ascontiguousarray(keys_sorted, np.int64)
    bins = np.zeros((num_tiles, 2), np.int32)
    lib().orc_tile_bin_edges(C.c_int(keys_sorted.shape[0]), _p(keys_sorted), C.c_int(num_tiles),
                             _p(bins))
    return bins


def bin_and_sort(centers, extents, depths, num_tiles_hit_arr, tile_bounds, block_width):
    """utils.py:40-59,106-162 -> (keys, ids, keys_sorted, ids_sorted, tile_bins)."""
    cum = np.cumsum(np.asarray(num_tiles_hit_arr, np.int64)).astype(np.int32)
    keys, ids = map_gaussian_to_intersects(centers, extents, depths, cum, tile_bounds, block_width)
    ks, is_ = sort_pairs(keys, ids)
    bins = tile_bin_edges(ks, tile_bounds[0] * tile_bounds[1])
    return keys, ids, ks, is_, bins


# --------------------------------------------------------------------------- SH
def num_sh_bases(degree: int) -> int:
    return int(lib().orc_num_sh_bases(C.c_int(degree)))


def sh_forward(degree, degrees_to_use, viewdirs, coeffs, precision="f32"):
    real, sfx = _real(precision)
    viewdirs, coeffs = _in(viewdirs, precision), _in(coeffs, precision)
    n = viewdirs.shape[0]
    out = np.zeros((n, 3), real)
    getattr(lib(), "orc_sh_forward" + sfx)(C.c_int(n), C.c_int(degree), C.c_int(degrees_to_use),
                                           _p(viewdirs), _p(coeffs), _p(out))
    return out


def sh_backward(degree, degrees_to_use, viewdirs, v_colors, precision="f32"):
    real, sfx = _real(precision)
    viewdirs, v_colors = _in(viewdirs, precision), _in(v_colors, precision)
    n = viewdirs.shape[0]
    out = np.zeros((n, num_sh_bases(degree), 3), real)
    getattr(lib(), "orc_sh_backward" + sfx)(C.c_int(n), C.c_int(degree), C.c_int(degrees_to_use),
                                            _p(viewdirs), _p(v_colors), _p(out))
    return out


# --------------------------------------------------------------------------- sampling
def texture_sample(texture_dims, uvs, texture, precision="f32"):
    real, sfx = _real(precision)
    texture_dims, uvs, texture = _i32(texture_dims), _in(uvs, precision), _in(texture, precision)
    q, ch = uvs.shape[0], texture.shape[1]
    out = np.zeros((q, ch), real)
    getattr(lib(), "orc_texture_sample" + sfx)(C.c_int(q), C.c_int(ch), _p(texture_dims), _p(uvs),
                                               _p(texture), _p(out))
    return out


def texture_sample_vjp(texture_dims, uvs, v_out, num_texels, precision="f32"):
    real, sfx = _real(precision)
    texture_dims, uvs, v_out = _i32(texture_dims), _in(uvs, precision), _in(v_out, precision)
    q, ch = v_out.shape
    out = np.zeros((num_texels, ch), real)
    getattr(lib(), "orc_texture_sample_vjp" + sfx)(C.c_int(q), C.c_int(ch), _p(texture_dims),
                                                   _p(uvs), _p(v_out), _p(out))
    return out


# --------------------------------------------------------------------------- raster
def texture_forward(scene: Scene, precision="f32"):
    """texture.cu:11-329 -> dict of the nine reference outputs."""
    real, sfx = _real(precision)
    st = scene._struct(precision)
    H, W, ch = scene.img_h, scene.img_w, scene.channels
    out = dict(
        out_img=np.zeros((H, W, 3), real), out_depth=np.zeros((H, W), real),
        out_reg=np.zeros((H, W), real), out_texture=np.zeros((H, W, ch), real),
        out_normal=np.zeros((H, W, 3), real), final_Ts=np.zeros((H, W), real),
        final_idx=np.zeros((H, W), np.int32), depth_idx=np.zeros((H, W), np.int32),
        out_reg_s=np.zeros((H, W, 3), real))
    getattr(lib(), "orc_texture_forward" + sfx)(
        C.byref(st), _p(out["out_img"]), _p(out["out_depth"]), _p(out["out_reg"]),
        _p(out["out_texture"]), _p(out["out_normal"]), _p(out["final_Ts"]), _p(out["final_idx"]),
        _p(out["depth_idx"]), _p(out["out_reg_s"]))
    return out


def texture_edit(scene: Scene, updated_img, updated_alpha, depth_lower, depth_upper,
                 precision="f32"):
    """texture_edit.cu:11-227 -> (X, C) accumulated edit channels (r*a, g*a, b*a, a, 1).

    ``scene.texture`` only supplies the shape (X, C), C >= 5; settings bit 0 is
    the blur switch of the edit kernel.
    """
    real, sfx = _real(precision)
    st = scene._struct(precision)
    X, ch = scene.texture.shape
    if ch < 5:
        raise ValueError("texture_edit writes 5 channels; the texture needs C >= 5")
    out = np.zeros((X, ch), real)
    ui, ua, dl, du = (_in(x, precision) for x in (updated_img, updated_alpha, depth_lower,
                                                  depth_upper))
    getattr(lib(), "orc_texture_edit" + sfx)(C.byref(st), _p(ui), _p(ua), _p(dl), _p(du),
                                             _p(out))
    return out


def texture_backward(scene: Scene, fwd: dict, v_img, v_depth, v_reg, v_alpha, v_texture_out,
                     v_normal, precision="f32"):
    """texture.cu:331-760 -> dict of the nine reference gradients.

    ``fwd`` holds final_Ts, final_idx, depth_idx and out_reg_s from a forward
    of the same precision.
    """
    real, sfx = _real(precision)
    st = scene._struct(precision)
    n = scene.means.shape[0]
    X, ch = scene.texture.shape
    g = dict(v_colors=np.zeros((n, 3), real), v_opacity=np.zeros((n, 1), real),
             v_means=np.zeros((n, 3), real), v_scales=np.zeros((n, 3), real),
             v_quats=np.zeros((n, 4), real), v_uv0=np.zeros((n, 1, 2), real),
             v_umap=np.zeros((n, 1, 3), real), v_vmap=np.zeros((n, 1, 3), real),
             v_texture=np.zeros((X, ch), real))
    fT = np.ascontiguousarray(fwd["final_Ts"], real)
    fS = np.ascontiguousarray(fwd["out_reg_s"], real)
    fI, dI = _i32(fwd["final_idx"]), _i32(fwd["depth_idx"])
    vi, vd, vr, va, vt, vn = (_in(x, precision) for x in (v_img, v_depth, v_reg, v_alpha,
                                                          v_texture_out, v_normal))
    getattr(lib(), "orc_texture_backward" + sfx)(
        C.byref(st), _p(fT), _p(fI), _p(dI), _p(fS), _p(vi), _p(vd), _p(vr), _p(va), _p(vt), _p(vn),
        _p(g["v_colors"]), _p(g["v_opacity"]), _p(g["v_means"]), _p(g["v_scales"]),
        _p(g["v_quats"]), _p(g["v_uv0"]), _p(g["v_umap"]), _p(g["v_vmap"]), _p(g["v_texture"]))
    return g
