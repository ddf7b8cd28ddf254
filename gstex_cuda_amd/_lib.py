"""ctypes binding of ``libgstex_hip.so`` (the C ABI declared in include/gstex_hip.h).

The library is built in-tree by ``gstex_cuda_amd/csrc/Makefile`` (see
``gstex_cuda_amd.build``).  Loading fails loudly: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgstex_hip.so")

_lock = threading.Lock()
_lib = None

P = C.c_void_p
I = C.c_int
F = C.c_float
Z = C.c_size_t
D = C.c_double

# name -> (restype, argtypes)
_SIGNATURES = {
    "gstex_last_error": (C.c_char_p, []),
    "gstex_preprocess_gaussians": (I, [I, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "gstex_preprocess_gaussians_backward": (I, [I] + [P] * 17 + [P]),
    "gstex_adam_step": (I, [Z, P, P, P, P, D, D, D, D, I, I, P, I, P]),
    "gstex_abi_version": (I, []),
    "gstex_get_aabb_2d": (I, [I, P, P, F, P, P, F, F, F, F, P, P, P]),
    "gstex_num_tiles_hit": (I, [I, P, P, I, I, I, P, P]),
    "gstex_num_tiles_hit_wrapped": (I, [I, P, P, I, P, P]),
    "gstex_project_gaussians": (I, [I, P, P, F, P, P, F, F, F, F, I, I, I, P, P, P, P, P]),
    "gstex_project_gaussians_total": (I, [I, P, P, F, P, P, F, F, F, F, I, I, I, P, P, P, P, P,
                                          P]),
    "gstex_project_points": (I, [I, P, P, F, F, F, F, I, P, P, P]),
    "gstex_compute_sh_forward": (I, [I, I, I, P, P, P, P]),
    "gstex_compute_sh_forward_bf16": (I, [I, I, I, P, P, P, P]),
    "gstex_compute_sh_forward_means": (I, [I, I, I, P, P, P, I, P, P]),
    "gstex_compute_sh_backward": (I, [I, I, I, P, P, P, P]),
    "gstex_compute_sh_backward_views": (I, [I, I, I, I, P, P, P, P]),
    "gstex_map_gaussian_to_intersects": (I, [I, I, P, P, P, P, I, I, I, I, P, P, P]),
    "gstex_sort_intersects_workspace_size": (Z, [I]),
    "gstex_sort_intersects": (I, [I, I, P, P, P, P, P, Z, P]),
    "gstex_sort_pairs_u32_workspace_size": (Z, [I]),
    "gstex_count_intersects": (I, [I, P, P, P]),
    "gstex_sort_pairs_u32": (I, [I, I, P, P, P, P, P, Z, P]),
    "gstex_get_tile_bin_edges": (I, [I, P, I, I, P, P]),
    "gstex_bin_gaussians_workspace_size": (Z, [I, I]),
    "gstex_bin_gaussians": (I, [I, I, P, P, P, P, I, I, I, P, P, P, Z, P]),
    "gstex_bin_gaussians_depth_workspace_size": (Z, [I]),
    "gstex_bin_gaussians_tiles_workspace_size": (Z, [I, I]),
    "gstex_bin_gaussians_depth": (I, [I, P, P, P, P, I, I, I, P, P, Z, P]),
    "gstex_bin_gaussians_tiles": (I, [I, I, P, I, I, P, Z, P, P, P, Z, P]),
    "gstex_bin_gaussians_capped": (I, [I, I, P, P, P, P, I, I, I, P, P, P, P, Z, P]),
    "gstex_texture_forward_workspace_size": (Z, [I]),
    "gstex_texture_forward": (I, [I, I, I, I, I, I, I, I,  # tiles, bw, img, texture_info
                                  P, P, I, P,  # dims, ids, num_ids, bins
                                  P, P, P, P, I, F, P, P, P, P, P, I,  # params .. texture, X
                                  P, P, F, F, F, F, I, P,  # viewmat, c2w, intrinsics, settings, bg
                                  P, P, P, P, P, P, P, P, P,  # nine outputs
                                  P, Z, P]),
    "gstex_texture_forward_noalloc": (I, [I, I, I, I, I, I, I, I,
                                          P, P, I, P,
                                          P, P, P, P, I, F, P, P, P, P, P, I,
                                          P, P, F, F, F, F, I, P,
                                          P, P, P, P, P, P, P, P, P,
                                          P]),  # stream
    "gstex_texture_backward_workspace_size": (Z, [I]),
    "gstex_texture_backward": (I, [I, I, I, I, I, I,  # H, W, bw, texture_info
                                   P, P, I, P,
                                   P, P, P, P, I, F, P, P, P, P, P, I,
                                   P, P, F, F, F, F, I, P,
                                   P, P, P, P,  # final_Ts, final_idx, depth_idx, final_s
                                   P, P, P, P, P, P,  # six output gradients
                                   P, P, P, P, P, P, P, P, P,  # nine parameter gradients
                                   P, Z, P]),
    "gstex_splat_records_size": (Z, [I]),
    "gstex_build_splat_records": (I, [I, P, P, F, P, P, P, P, P, P, P, P, P, F, F, F, F, P, P]),
    "gstex_set_raster_culling": (I, [I]),
    "gstex_set_precise_edge_on": (I, [I]),
    "gstex_raster_wave_masks_size": (Z, [I, I, I]),
    "gstex_raster_forward": (I, [I, I, I, I, I, P, I, P, I, P, P, I, I, P, P, F, F, F, F, P,
                                 P, P, P, P, P, P, P, P, P,  # nine outputs
                                 P, Z, P]),  # wave_masks, bytes, stream
    "gstex_raster_backward_l2": (I, [I, I, I, I, I, P, I, P, I, P, P, I, I, P, P, F, F, F, F, P,
                                     P, P, P, P,  # final_Ts, final_idx, depth_idx, final_s
                                     P, P, P, F,  # out_img, out_texture, target, scale
                                     P, P, P, P,  # v_output_depth, _reg, _alpha, _normal
                                     P, P, P, Z, P]),  # grad_records, v_texture, masks, bytes, stream
    "gstex_raster_backward": (I, [I, I, I, I, I, P, I, P, I, P, P, I, I, P, P, F, F, F, F, P,
                                  P, P, P, P,  # final_Ts, final_idx, depth_idx, final_s
                                  P, P, P, P, P, P,  # six output gradients
                                  P, P, P, Z, P]),  # grad_records, v_texture, masks, bytes, stream
    "gstex_finalize_gradients": (I, [I, P, P, P, P, P, P, F, F, P, P, P, P, P, P, P, P, P]),
    "gstex_finalize_gradients_sh": (I, [I, P, P, P, P, P, P, F, F, P, I, I,  # .., means, degrees
                                        P, P, P, P, P, P, P, P, P, P]),  # 9 gradients, stream
    "gstex_finalize_gradients_views": (I, [I, I, P, P, Z, P, P, P, P, P, P, P, P, P, P, P, P, P,
                                           P, P]),
    "gstex_image_loss_grad": (I, [I, I, I, P, P, P, F, P, F, P, P, P, P]),
    "gstex_fill_f32": (I, [P, Z, F, P]),
    "gstex_texture_edit_workspace_size": (Z, [I]),
    "gstex_texture_edit": (I, [I, I, I, I, I, I, I, I, I,  # tiles, bw, img, texture_info, X
                               P, P, P, P, P, P, I, P,  # dims, edits, ids, num_ids, bins
                               P, P, P, I, F, P, P, P, P,  # opac, means, scales, N, g, quats, uv
                               P, P, F, F, F, F, I, P,  # viewmat, c2w, intrinsics, settings, bg
                               P, P, Z, P]),  # out, workspace, bytes, stream
    "gstex_views_table_size": (Z, [I, I]),
    "gstex_views_setup": (I, [I, P, I, I, I, I, P, I, I, P, P, P, Z, P]),
    "gstex_views_release": (I, [P]),
    "gstex_views_raster_forward": (I, [P, P]),
    "gstex_views_raster_backward": (I, [P, P]),
    "gstex_texture_sample_forward": (I, [I, I, I, P, P, I, P, I, P, P]),
    "gstex_texture_sample_backward_workspace_size": (Z, [I]),
    "gstex_texture_sample_backward": (I, [I, I, I, P, P, I, P, I, P, P, P, Z, P]),
}


class View(C.Structure):
    """``gstex_view`` of include/gstex_hip.h: one view of a multi-view raster pass."""

    _fields_ = [("img_width", I), ("img_height", I), ("fx", F), ("fy", F), ("cx", F),
                ("cy", F), ("viewmat", P), ("c2w", P), ("records", P),
                ("gaussian_ids_sorted", P), ("num_ids", I), ("tile_bins", P),
                ("wave_masks", P), ("wave_masks_bytes", Z),
                ("out_img", P), ("out_depth", P), ("out_reg", P), ("out_texture", P),
                ("out_normal", P), ("final_Ts", P), ("final_idx", P), ("depth_idx", P),
                ("out_reg_s", P), ("v_img", P), ("v_depth", P), ("v_reg", P), ("v_alpha", P),
                ("v_texture_out", P), ("v_normal", P), ("grad_records", P)]


class GstexError(RuntimeError):
    """A C-ABI call returned a nonzero status."""


def lib():
    """Load ``libgstex_hip.so`` (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `make -C gstex_cuda_amd/csrc` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            handle = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def call(name: str, *args):
    """Invoke ``name`` and raise GstexError with the library's message on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().gstex_last_error().decode(errors="replace")
        raise GstexError(msg or f"{name} failed with status {rc}")
    return rc


def exported_symbols():
    return list(_SIGNATURES)
