"""One-view forward + backward of the textured-2DGS path on the per-stage C ABI.

``RenderStep`` runs what a GStex training iteration runs for one view
(example.py:121-187,256-258) without autograd bookkeeping: SH colours,
projection with tile counts, depth-first binning, splat records, raster
forward, the L2 photometric loss gradient, raster backward, gradient finalize
and the SH adjoint.  Buffers are allocated once and reused; every kernel runs
on torch's current stream, and a kernel's duration can be read from HIP events
recorded around it.
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass
from typing import Dict, Optional

import torch

from . import _lib
from ._lib import call


def _p(t: Optional[torch.Tensor]):
    return None if t is None or t.numel() == 0 else t.data_ptr()


@dataclass
class Camera:
    viewmat: torch.Tensor
    c2w: torch.Tensor
    fx: float
    fy: float
    cx: float
    cy: float
    img_h: int
    img_w: int


# hip_runtime_api.h: an event created with hipEventDisableSystemFence skips
# the system-scope fence when it is recorded, i.e. the cache write-back and
# invalidation that otherwise leave the GPU idle for a few microseconds
# between the kernels the event brackets.  Such events only time: nothing
# waits on them and no host read depends on them.
_HIP_EVENT_DISABLE_SYSTEM_FENCE = 0x20000000


def _hip_fn(name: str, *argtypes):
    """A HIP runtime entry, resolved through the library's own dependency
    (the runtime its kernels run on)."""
    fn = getattr(_lib.lib(), name)
    fn.restype = ctypes.c_int
    fn.argtypes = list(argtypes)
    return fn


class StageTimer:
    """HIP-event pairs per stage, recorded on the current stream (timing-only
    events, without the system-scope fence)."""

    def __init__(self):
        self.events: Dict[str, list] = {}
        vp = ctypes.c_void_p
        self._create = _hip_fn("hipEventCreateWithFlags", ctypes.POINTER(vp), ctypes.c_uint)
        self._record = _hip_fn("hipEventRecord", vp, vp)
        self._sync = _hip_fn("hipEventSynchronize", vp)
        self._elapsed = _hip_fn("hipEventElapsedTime", ctypes.POINTER(ctypes.c_float), vp, vp)
        self._destroy = _hip_fn("hipEventDestroy", vp)

    def _mark(self) -> ctypes.c_void_p:
        ev = ctypes.c_void_p()
        if self._create(ctypes.byref(ev), _HIP_EVENT_DISABLE_SYSTEM_FENCE) != 0:
            raise RuntimeError("hipEventCreateWithFlags failed")
        if self._record(ev, torch.cuda.current_stream().cuda_stream) != 0:
            self._destroy(ev)
            raise RuntimeError("hipEventRecord failed")
        return ev

    def start(self, name: str):
        self.events.setdefault(name, []).append([self._mark(), None])

    def stop(self, name: str):
        self.events[name][-1][1] = self._mark()

    def mean_ms(self, name: str) -> float:
        pairs = self.events.get(name, [])
        if not pairs:
            return float("nan")
        total = 0.0
        for a, b in pairs:
            ms = ctypes.c_float()
            if self._sync(b) != 0 or self._elapsed(ctypes.byref(ms), a, b) != 0:
                raise RuntimeError("hipEventElapsedTime failed")
            total += ms.value
        return total / len(pairs)

    def clear(self):
        for pairs in self.events.values():
            for pair in pairs:
                for ev in pair:
                    if ev is not None:
                        self._destroy(ev)
        self.events.clear()

    def __del__(self):
        if sys.is_finalizing():  # the runtime may already be shutting down
            return
        try:
            self.clear()
        except Exception:
            pass


class RenderStep:
    """Forward + backward of one view with reusable device buffers.

    ``params`` (device tensors, reference layouts): means (N,3), scales (N,3),
    quats (N,4), opacities (N,1), uv0 (N,1,2), umap (N,1,3), vmap (N,1,3),
    texture (X,C) float32 or bfloat16, texture_dims (N,3) int32, and either
    colors (N,3) or sh_coeffs (N,D,3) float32 or bfloat16 with ``sh_degree``.
    ``culling`` (None: the library's process-wide setting) switches the raster
    kernels' screen-box culling for this step's launches only.
    """

    def __init__(self, params: Dict[str, torch.Tensor], cam: Camera, block_width: int = 16,
                 settings: int = 1 << 8, background: Optional[torch.Tensor] = None,
                 sh_degree: int = -1, use_wave_masks: bool = True,
                 capacity: Optional[int] = None, culling: Optional[bool] = None,
                 overlap: bool = True):
        self.p = params
        self.culling = culling
        self.cam = cam
        self.bw = block_width
        self.settings = settings
        self.sh_degree = sh_degree
        dev = params["means"].device
        self.dev = dev
        n = params["means"].shape[0]
        self.n = n
        self.C = params["texture"].shape[1]
        self.X = params["texture"].shape[0]
        # bfloat16 texels / SH coefficients are read as stored; all arithmetic
        # and every gradient (v_texture, v_sh_coeffs) is float32
        self.texture_format = 1 if params["texture"].dtype == torch.bfloat16 else 0
        H, W = cam.img_h, cam.img_w
        self.tiles = ((W + block_width - 1) // block_width, (H + block_width - 1) // block_width)
        f32 = dict(dtype=torch.float32, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        self.background = (torch.zeros(3, **f32) if background is None else background)
        # per-Gaussian
        self.centers = torch.empty((n, 2), **f32)
        self.extents = torch.empty((n, 2), **f32)
        self.depths = torch.empty((n,), **f32)
        self.tiles_hit = torch.empty((n,), **i32)
        self.colors = params.get("colors")
        if sh_degree >= 0:
            self.colors = torch.empty((n, 3), **f32)
        self.records = torch.empty((max(_lib.lib().gstex_splat_records_size(n), 1),),
                                   dtype=torch.uint8, device=dev)
        self.grad_records = torch.empty((n, 32), **f32)
        self.tile_bins = torch.empty((self.tiles[0] * self.tiles[1], 2), **i32)
        self.ids = torch.empty((0,), **i32)
        self.bin_ws = torch.empty((1,), dtype=torch.uint8, device=dev)
        self.wave_masks = torch.empty((1,), dtype=torch.uint8, device=dev)
        self.use_wave_masks = use_wave_masks
        # per-pixel
        self.out_img = torch.empty((H, W, 3), **f32)
        self.out_depth = torch.empty((H, W), **f32)
        self.out_reg = torch.empty((H, W), **f32)
        self.out_texture = torch.empty((H, W, self.C), **f32)
        self.out_normal = torch.empty((H, W, 3), **f32)
        self.final_Ts = torch.empty((H, W), **f32)
        self.final_idx = torch.empty((H, W), **i32)
        self.depth_idx = torch.empty((H, W), **i32)
        self.out_reg_s = torch.empty((H, W, 3), **f32)
        self.v_img = torch.empty((H, W, 3), **f32)
        self.v_tex = torch.empty((H, W, self.C), **f32)
        self.v_zero1 = torch.zeros((H, W), **f32)
        self.v_zero3 = torch.zeros((H, W, 3), **f32)
        # a batch that shares the texel gradient or runs the SH adjoint for all
        # of its views at once turns these off
        self.zero_texture_grad = True
        self.sh_backward = True
        self.finalize = True
        self.v_reg = None  # regulariser cotangents, allocated on first use
        self.v_normal = None
        self._reg_weight = None
        # gradients (reference layouts)
        self.grads = dict(
            v_colors=torch.empty((n, 3), **f32), v_opacity=torch.empty((n, 1), **f32),
            v_means=torch.empty((n, 3), **f32), v_scales=torch.empty((n, 3), **f32),
            v_quats=torch.empty((n, 4), **f32), v_uv0=torch.empty((n, 1, 2), **f32),
            v_umap=torch.empty((n, 1, 3), **f32), v_vmap=torch.empty((n, 1, 3), **f32),
            v_texture=torch.empty((self.X, self.C), **f32))
        if sh_degree >= 0:
            self.grads["v_sh_coeffs"] = torch.empty(params["sh_coeffs"].shape, **f32)
            self._viewdirs = torch.empty((n, 3), **f32)
        self.num_intersects = 0
        # Capacity mode: the intersect list has a fixed capacity and its length
        # stays on the device, so a step makes no host synchronisation and can
        # be captured in a graph (see capture()).
        self.cap = capacity
        self.m_dev = torch.zeros((1,), **i32)
        self.graph = None
        # With overlap, the work that needs neither the projection nor the
        # binning (SH colours, splat records, zeroing the gradient buffers)
        # runs on a second stream (default priority) beside them; the raster
        # launches wait for it through events.
        self._side = torch.cuda.Stream(dev) if overlap else None
        self._ev_records = torch.cuda.Event() if overlap else None
        self._ev_zeroed = torch.cuda.Event() if overlap else None
        self._zeroed = False       # zeroing queued on the side stream, not yet waited for
        self._zero_ready = None    # the stream that waited for it
        if capacity is not None:
            self._ensure_bins(int(capacity))
        else:
            # exact size: binning runs as a depth phase (queued before the host
            # waits for the intersect total) and a tile phase
            self.depth_ws = torch.empty(
                (max(_lib.lib().gstex_bin_gaussians_depth_workspace_size(n), 1),),
                dtype=torch.uint8, device=dev)
            self.m_host = torch.zeros((1,), dtype=torch.int32, pin_memory=True)
            self._ev_m = torch.cuda.Event()
            # the total's copy to the host runs on a stream of its own
            # (default priority), so the main stream goes on with binning
            self._copy_stream = torch.cuda.Stream(dev)
            self._ev_projected = torch.cuda.Event()

    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _ensure_bins(self, m: int):
        if self.ids.numel() < m:
            self.ids = torch.empty((int(m * 1.25) + 1024,), dtype=torch.int32, device=self.dev)
        if self.cap is None:
            need = _lib.lib().gstex_bin_gaussians_tiles_workspace_size(self.n, m)
        else:
            need = _lib.lib().gstex_bin_gaussians_workspace_size(self.n, m)
        if self.bin_ws.numel() < need:
            self.bin_ws = torch.empty((int(need * 1.25) + 4096,), dtype=torch.uint8,
                                      device=self.dev)
        need = _lib.lib().gstex_raster_wave_masks_size(m, self.tiles[0], self.tiles[1])
        if self.wave_masks.numel() < need:
            self.wave_masks = torch.empty((int(need * 1.25) + 4096,), dtype=torch.uint8,
                                          device=self.dev)

    def _raster(self, name: str, *args):
        """One raster launch, under this step's culling choice when it has one."""
        if self.culling is None:
            return call(name, *args)
        prev = _lib.lib().gstex_set_raster_culling(int(bool(self.culling)))
        try:
            return call(name, *args)
        finally:
            _lib.lib().gstex_set_raster_culling(prev)

    def _masks(self):
        if not self.use_wave_masks:
            return None, 0
        return _p(self.wave_masks), self.wave_masks.numel()

    def forward(self, timer: Optional[StageTimer] = None):
        self.forward_pre()
        self.forward_raster(timer)

    def _zero_grads(self, s):
        call("gstex_fill_f32", _p(self.grad_records), self.grad_records.numel(), 0.0, s)
        if self.zero_texture_grad:
            call("gstex_fill_f32", _p(self.grads["v_texture"]), self.grads["v_texture"].numel(),
                 0.0, s)

    def _sh_in_finalize(self) -> bool:
        """True when the SH adjoint runs inside the gradient finalize, which
        forms the view directions itself (no direction buffer needed)."""
        return (0 <= self.sh_degree <= 3 and self.sh_backward
                and self.grads["v_sh_coeffs"].data_ptr() % 16 == 0)

    def _colors_and_records(self, s):
        p, cam, n = self.p, self.cam, self.n
        if self.sh_degree >= 0:
            bf16 = p["sh_coeffs"].dtype == torch.bfloat16
            if self._sh_in_finalize():
                # directions formed in the SH kernel: no pass of their own
                call("gstex_compute_sh_forward_means", n, self.sh_degree, self.sh_degree,
                     _p(p["means"]), _p(cam.c2w), _p(p["sh_coeffs"]), int(bf16),
                     _p(self.colors), s)
            else:
                torch.sub(p["means"], cam.c2w[:3, 3], out=self._viewdirs)
                fn = "gstex_compute_sh_forward_bf16" if bf16 else "gstex_compute_sh_forward"
                call(fn, n, self.sh_degree, self.sh_degree, _p(self._viewdirs),
                     _p(p["sh_coeffs"]), _p(self.colors), s)
        call("gstex_build_splat_records", n, _p(p["means"]), _p(p["scales"]), 1.0,
             _p(p["quats"]), _p(self.colors), _p(p["opacities"]), _p(p["uv0"]), _p(p["umap"]),
             _p(p["vmap"]), _p(p["texture_dims"]), _p(cam.viewmat), _p(cam.c2w), cam.fx, cam.fy,
             cam.cx, cam.cy, _p(self.records), s)

    def forward_pre(self):
        """SH colours, projection, binning and splat records (everything of
        the forward before the raster launch)."""
        p, cam, s = self.p, self.cam, self._stream()
        n = self.n
        main = torch.cuda.current_stream(self.dev)
        side = self._side
        if side is not None:
            # SH colours, splat records and the gradient zeroing beside the
            # projection and binning; the side stream first waits for
            # everything already queued (parameter updates, the last step)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                self._colors_and_records(side.cuda_stream)
                self._ev_records.record(side)
                self._zero_grads(side.cuda_stream)
                self._ev_zeroed.record(side)
            self._zeroed = True
        proj = (_p(p["means"]), _p(p["scales"]), 1.0, _p(p["quats"]), _p(cam.viewmat), cam.fx,
                cam.fy, cam.cx, cam.cy, cam.img_h, cam.img_w, self.bw, _p(self.centers),
                _p(self.extents), _p(self.depths), _p(self.tiles_hit))
        if self.cap is not None:
            call("gstex_project_gaussians", n, *proj, s)
        else:
            # the intersect total comes out of the projection launch itself
            call("gstex_project_gaussians_total", n, *proj, _p(self.m_dev), s)
        if self.cap is not None:
            m = int(self.cap)
            call("gstex_bin_gaussians_capped", n, m, _p(self.centers), _p(self.extents),
                 _p(self.depths), _p(self.tiles_hit), self.tiles[0], self.tiles[1], self.bw,
                 _p(self.ids), _p(self.tile_bins), _p(self.m_dev), _p(self.bin_ws),
                 self.bin_ws.numel(), s)
        else:
            # the depth phase needs no total: it is queued right after the
            # projection, and the host's wait for the total's copy overlaps it
            self._ev_projected.record(main)
            cs = self._copy_stream
            cs.wait_event(self._ev_projected)
            with torch.cuda.stream(cs):
                self.m_host.copy_(self.m_dev, non_blocking=True)
                self._ev_m.record(cs)
            call("gstex_bin_gaussians_depth", n, _p(self.centers), _p(self.extents),
                 _p(self.depths), _p(self.tiles_hit), self.tiles[0], self.tiles[1], self.bw,
                 _p(self.tile_bins), _p(self.depth_ws), self.depth_ws.numel(), s)
            self._ev_m.synchronize()
            m = int(self.m_host[0])
            self.num_intersects = m
            self._ensure_bins(m)
            if m > 0:
                call("gstex_bin_gaussians_tiles", n, m, _p(self.m_dev), self.tiles[0],
                     self.tiles[1], _p(self.depth_ws), self.depth_ws.numel(), _p(self.ids),
                     _p(self.tile_bins), _p(self.bin_ws), self.bin_ws.numel(), s)
        self._list_len = m
        if side is None:
            self._colors_and_records(s)
        else:
            # The raster forward needs the records.  The zeroing, queued after
            # them on the side stream, ends long before the binning does, so
            # one wait here covers both (each wait is a barrier packet that
            # the main stream processes between kernels).
            main.wait_event(self._ev_zeroed)
            self._zeroed = False
            self._zero_ready = main

    def forward_raster(self, timer: Optional[StageTimer] = None):
        p, cam, s = self.p, self.cam, self._stream()
        n, m = self.n, self._list_len
        if timer:
            timer.start("raster_forward")
        self._raster("gstex_raster_forward", cam.img_w, cam.img_h, self.bw, self.C, self.settings,
             _p(self.records), n, _p(self.ids), m, _p(self.tile_bins), _p(p["texture"]), self.X,
             self.texture_format, _p(cam.viewmat), _p(cam.c2w), cam.fx, cam.fy, cam.cx, cam.cy,
             _p(self.background),
             _p(self.out_img), _p(self.out_depth), _p(self.out_reg), _p(self.out_texture),
             _p(self.out_normal), _p(self.final_Ts), _p(self.final_idx), _p(self.depth_idx),
             _p(self.out_reg_s), *self._masks(), s)
        if timer:
            timer.stop("raster_forward")

    def l2_loss_grad(self, target: torch.Tensor, v_normal: Optional[torch.Tensor] = None,
                     normal_scale: float = 0.0):
        """d/d(img, texture) of mean((out_img + out_texture[..., :3] - target)^2),
        in one kernel; with ``v_normal``, also the normal term's cotangent
        ``normal_scale * n`` minus ``normal_scale`` in z."""
        H, W = self.cam.img_h, self.cam.img_w
        if target.dtype != torch.float32 or not target.is_contiguous():
            target = target.to(torch.float32).contiguous()
        call("gstex_image_loss_grad", H * W, self.C, 1, _p(self.out_img), _p(self.out_texture),
             _p(target), 2.0 / target.numel(), _p(self.out_normal if v_normal is not None
                                                   else None),
             normal_scale, _p(self.v_img), _p(self.v_tex), _p(v_normal), self._stream())

    def regularised_loss_grad(self, target: torch.Tensor, reg_weight: float = 100.0,
                              normal_weight: float = 0.05):
        """Cotangents of mean((out_img + out_texture - target)^2)
        + reg_weight * mean(out_reg)
        + normal_weight * mean(n_x^2 + n_y^2 + (1 - n_z)^2),
        the distortion and normal terms of example.py:199-207 with weights;
        returns (v_reg, v_normal) for ``backward``."""
        H, W = self.cam.img_h, self.cam.img_w
        if self.v_reg is None or self._reg_weight != reg_weight:
            if self.v_reg is None:
                self.v_reg = torch.empty_like(self.v_zero1)
                self.v_normal = torch.empty_like(self.v_zero3)
            self.v_reg.fill_(reg_weight / (H * W))  # constant: refilled on change only
            self._reg_weight = reg_weight
        self.l2_loss_grad(target, self.v_normal, 2.0 * normal_weight / (H * W))
        return self.v_reg, self.v_normal

    def regularised_step(self, target: torch.Tensor, reg_weight: float = 100.0,
                         normal_weight: float = 0.05, timer: Optional[StageTimer] = None):
        """Forward, gradient of the L2 loss plus the distortion and normal
        regularisers (``regularised_loss_grad``), backward."""
        self.forward(timer)
        v_reg, v_normal = self.regularised_loss_grad(target, reg_weight, normal_weight)
        self.backward(timer, v_reg=v_reg, v_normal=v_normal)

    def backward(self, timer: Optional[StageTimer] = None, v_depth=None, v_reg=None,
                 v_alpha=None, v_normal=None, loss_target: Optional[torch.Tensor] = None):
        """The raster backward and the parameter gradients.  The image and
        texture cotangents are ``v_img`` / ``v_tex``, or with ``loss_target``
        (three channels, float32, contiguous, H x W x 3) those of
        ``l2_loss_grad(loss_target)``, formed inside the raster backward
        (``v_img`` / ``v_tex`` are then not written)."""
        p, cam, s, g = self.p, self.cam, self._stream(), self.grads
        n, m = self.n, self._list_len
        cur = torch.cuda.current_stream(self.dev)
        if self._zero_ready is not None and self._zero_ready == cur:
            pass  # zeroed on the side stream; this stream already waited for it
        elif self._zeroed or self._zero_ready is not None:
            # zeroed on the side stream during the forward
            cur.wait_event(self._ev_zeroed)
        else:
            self._zero_grads(s)
        self._zeroed = False
        self._zero_ready = None
        v_depth = self.v_zero1 if v_depth is None else v_depth
        v_reg = self.v_zero1 if v_reg is None else v_reg
        v_alpha = self.v_zero1 if v_alpha is None else v_alpha
        v_normal = self.v_zero3 if v_normal is None else v_normal
        if timer:
            timer.start("raster_backward")
        head = (cam.img_w, cam.img_h, self.bw, self.C, self.settings,
                _p(self.records), n, _p(self.ids), m, _p(self.tile_bins), _p(p["texture"]),
                self.X, self.texture_format, _p(cam.viewmat), _p(cam.c2w), cam.fx, cam.fy,
                cam.cx, cam.cy, _p(self.background),
                _p(self.final_Ts), _p(self.final_idx), _p(self.depth_idx), _p(self.out_reg_s))
        if loss_target is None:
            self._raster("gstex_raster_backward", *head, _p(self.v_img), _p(v_depth), _p(v_reg),
                         _p(v_alpha), _p(self.v_tex), _p(v_normal), _p(self.grad_records),
                         _p(g["v_texture"]), *self._masks(), s)
        else:
            self._raster("gstex_raster_backward_l2", *head, _p(self.out_img),
                         _p(self.out_texture), _p(loss_target), 2.0 / loss_target.numel(),
                         _p(v_depth), _p(v_reg), _p(v_alpha), _p(v_normal),
                         _p(self.grad_records), _p(g["v_texture"]), *self._masks(), s)
        if timer:
            timer.stop("raster_backward")
        if not self.finalize:
            return
        fin = (n, _p(self.grad_records), _p(self.records), _p(p["quats"]), _p(p["scales"]),
               _p(cam.viewmat), _p(cam.c2w), cam.fx, cam.fy)
        grads = (_p(g["v_colors"]), _p(g["v_opacity"]), _p(g["v_means"]), _p(g["v_scales"]),
                 _p(g["v_quats"]), _p(g["v_uv0"]), _p(g["v_umap"]), _p(g["v_vmap"]))
        sh = self.sh_degree >= 0 and self.sh_backward
        if sh and self._sh_in_finalize():
            # the SH adjoint in the finalize pass (view directions from the means)
            call("gstex_finalize_gradients_sh", *fin, _p(p["means"]), self.sh_degree,
                 self.sh_degree, *grads, _p(g["v_sh_coeffs"]), s)
            return
        call("gstex_finalize_gradients", *fin, *grads, s)
        if sh:
            call("gstex_compute_sh_backward", n, self.sh_degree, self.sh_degree,
                 _p(self._viewdirs), _p(g["v_colors"]), _p(g["v_sh_coeffs"]), s)

    def step(self, target: torch.Tensor, timer: Optional[StageTimer] = None):
        """Forward, L2 loss gradient against ``target``, backward.  With three
        channels (float32 target of the image's shape) the loss gradient is
        formed inside the raster backward, and ``v_img`` / ``v_tex`` are not
        written; call ``forward``, ``l2_loss_grad`` and ``backward`` to keep
        them."""
        self.forward(timer)
        fused = self._fused_loss_target(target)
        if fused is None:
            self.l2_loss_grad(target)
        self.backward(timer, loss_target=fused)

    def _fused_loss_target(self, target: torch.Tensor) -> Optional[torch.Tensor]:
        """target when the raster backward can form the L2 cotangents itself."""
        if (self.C == 3 and target.dtype == torch.float32 and target.is_contiguous()
                and target.device == self.out_img.device
                and tuple(target.shape) == (self.cam.img_h, self.cam.img_w, 3)):
            return target
        return None

    def intersect_total(self) -> int:
        """Intersect count of the last step (a host synchronisation)."""
        if self.cap is None:
            return self.num_intersects
        self.num_intersects = int(self.m_dev.item())
        return self.num_intersects

    def capacity_ok(self) -> bool:
        """True when the last step's intersects fitted the capacity (synchronises)."""
        return self.cap is None or self.intersect_total() <= self.cap

    def capture(self, target: torch.Tensor):
        """Capture one step against ``target`` (a tensor that stays in place) as a
        graph; ``replay()`` then reruns it with no host work.  Needs capacity mode."""
        if self.cap is None:
            raise RuntimeError("graph capture needs a RenderStep with a capacity")
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            for _ in range(2):
                self.step(target)
        torch.cuda.current_stream(self.dev).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.step(target)
        return self.graph

    def replay(self):
        self.graph.replay()


def scene_to_step(scene, device, **kw) -> RenderStep:
    """RenderStep for a ``synthetic.Scene``."""
    t = scene.to_torch(device)
    params = dict(means=t["means"], scales=t["scales"], quats=t["quats"],
                  opacities=t["opacities"], uv0=t["uv0"], umap=t["umap"], vmap=t["vmap"],
                  texture=t["texture"], texture_dims=t["texture_dims"])
    if scene.sh_coeffs is not None:
        params["sh_coeffs"] = t["sh_coeffs"]
    else:
        params["colors"] = t["colors"]
    cam = Camera(viewmat=t["viewmat"], c2w=t["c2w"], fx=scene.fx, fy=scene.fy, cx=scene.cx,
                 cy=scene.cy, img_h=scene.img_h, img_w=scene.img_w)
    return RenderStep(params, cam, block_width=scene.block_width, settings=scene.settings,
                      background=t["background"], sh_degree=scene.sh_degree, **kw)


class MultiViewStep:
    """Forward + backward of several views of one scene, gradients summed.

    Each view is a capacity-mode ``RenderStep`` over the shared parameter
    tensors (texels / SH coefficients may be bfloat16).  ``step()`` renders every
    view, applies the L2 loss against its target and backpropagates; the
    per-view gradients are summed into ``grads``.  ``capture()`` records the
    whole batch as one graph, so ``replay()`` costs one launch from the host.
    The capacities come from one eager pass per view (``capacity_factor`` of
    the measured intersect count); ``replay_checked()`` verifies them.

    The views run on ``streams`` HIP streams (default priority) forked from
    and joined back into the caller's stream, so the tiles of one view fill
    the compute units another view's longest tiles leave idle; in a captured
    graph they become parallel branches.  With ``balance`` the views are
    timed once and placed longest first on the least loaded stream; otherwise
    they are dealt round-robin.

    Gradient passes run once per batch, not once per view: every view's
    backward adds its texel gradients into the one ``grads["v_texture"]``
    (zeroed once per batch); the views' gradient records and splat records sit
    in stacks that one ``gstex_finalize_gradients_views`` turns into the summed
    per-Gaussian gradients; with SH, the views' directions and colour
    gradients sit in (V, N, 3) stacks for one ``gstex_compute_sh_backward_views``.
    """

    def __init__(self, params: Dict[str, torch.Tensor], cams, targets, block_width: int = 16,
                 settings: int = 1 << 8, background: Optional[torch.Tensor] = None,
                 sh_degree: int = -1, capacity_factor: float = 1.1,
                 use_wave_masks: bool = True, streams: int = 4, balance: bool = True,
                 queue: bool = False):
        self.targets = list(targets)
        # the views' concurrency comes from the batch's streams, not from
        # each view's own side stream
        kw = dict(block_width=block_width, settings=settings, background=background,
                  sh_degree=sh_degree, use_wave_masks=use_wave_masks, overlap=False)
        self.views = []
        for cam, tgt in zip(cams, self.targets):
            probe = RenderStep(params, cam, **kw)
            probe.forward()
            cap = int(probe.num_intersects * capacity_factor) + 4096
            del probe
            self.views.append(RenderStep(params, cam, capacity=cap, **kw))
        self.grads = {k: torch.zeros_like(v) for k, v in self.views[0].grads.items()}
        self.graph = None
        dev = self.views[0].dev
        n, nv = self.views[0].n, len(self.views)
        f32 = dict(dtype=torch.float32, device=dev)
        self.sh_degree = sh_degree
        self.params = params
        # per-view splat records and gradient records, stacked
        rec_size = self.views[0].records.numel()
        self._rec_stride = (rec_size + 255) // 256 * 256
        self._records = torch.empty((nv * self._rec_stride,), dtype=torch.uint8, device=dev)
        self._grad_records = torch.empty((nv, n, 32), **f32)
        self._viewmats = torch.stack([v.cam.viewmat.to(**f32) for v in self.views]).contiguous()
        self._c2ws = torch.stack([v.cam.c2w.to(**f32) for v in self.views]).contiguous()
        self._focals = torch.tensor([[v.cam.fx, v.cam.fy] for v in self.views], **f32)
        self._v_colors = None
        if sh_degree >= 0:
            self._viewdirs = torch.empty((nv, n, 3), **f32)
            self._v_colors = torch.empty((nv, n, 3), **f32)
        for i, v in enumerate(self.views):
            v.records = self._records[i * self._rec_stride:i * self._rec_stride + rec_size]
            v.grad_records = self._grad_records[i]
            v.grads = {"v_texture": self.grads["v_texture"]}
            v.zero_texture_grad = False
            v.finalize = False
            v.sh_backward = False
            if sh_degree >= 0:
                v._viewdirs = self._viewdirs[i]
        self.queue = bool(queue)
        self._table = None
        if self.queue:
            self._setup_queue(block_width, settings)
        k = min(max(int(streams), 1), len(self.views))
        self.streams = [torch.cuda.Stream(dev) for _ in range(k)] if k > 1 else []
        # views -> streams: longest first onto the least loaded stream, by the
        # time of one eager step per view (tiles of very uneven length make
        # equal-sized views cost very different times)
        self.assign = [(i, i % max(k, 1)) for i in range(nv)]
        if k > 1 and balance:
            cost = []
            for v, tgt in zip(self.views, self.targets):
                v.step(tgt)  # warm
                t0 = torch.cuda.Event(enable_timing=True)
                t1 = torch.cuda.Event(enable_timing=True)
                t0.record()
                v.step(tgt)
                t1.record()
                t1.synchronize()
                cost.append(t0.elapsed_time(t1))
            load = [0.0] * k
            self.assign = []
            for i in sorted(range(nv), key=lambda j: -cost[j]):
                st = min(range(k), key=lambda j: load[j])
                load[st] += cost[i]
                self.assign.append((i, st))
            self.view_ms = cost

    def _setup_queue(self, block_width: int, settings: int):
        """The launch table of the multi-view raster pass (gstex_views_setup)."""
        lib = _lib.lib()
        v0 = self.views[0]
        arr = (_lib.View * len(self.views))()
        total = 0
        for i, v in enumerate(self.views):
            cam = v.cam
            total += v.tiles[0] * v.tiles[1]
            masks, mbytes = v._masks()
            arr[i] = _lib.View(
                cam.img_w, cam.img_h, cam.fx, cam.fy, cam.cx, cam.cy, _p(cam.viewmat),
                _p(cam.c2w), _p(v.records), _p(v.ids), int(v.cap), _p(v.tile_bins), masks,
                mbytes, _p(v.out_img), _p(v.out_depth), _p(v.out_reg), _p(v.out_texture),
                _p(v.out_normal), _p(v.final_Ts), _p(v.final_idx), _p(v.depth_idx),
                _p(v.out_reg_s), _p(v.v_img), _p(v.v_zero1), _p(v.v_zero1), _p(v.v_zero1),
                _p(v.v_tex), _p(v.v_zero3), _p(v.grad_records))
        self._views_arr = arr
        nbytes = lib.gstex_views_table_size(len(self.views), total)
        self._table = torch.empty((nbytes + 256,), dtype=torch.uint8, device=v0.dev)
        base = self._table.data_ptr()
        self._table_ptr = (base + 255) // 256 * 256
        tex = self.params["texture"]
        call("gstex_views_setup", len(self.views), arr, block_width, v0.C, settings, v0.n,
             _p(tex), v0.X, v0.texture_format, _p(v0.background), _p(self.grads["v_texture"]),
             self._table_ptr, nbytes, v0._stream())

    def __del__(self):
        if getattr(self, "_table", None) is not None:
            try:
                _lib.lib().gstex_views_release(self._table_ptr)
            except Exception:
                pass

    def _step_queue(self):
        """One batch through the multi-view raster pass: the views' SH,
        projection, binning and records (on the streams), one forward launch
        over every view's tiles, the loss cotangents, one backward launch."""
        dev = self.views[0].dev
        main = torch.cuda.current_stream(dev)
        if self.streams:
            for st in self.streams:
                st.wait_stream(main)
            for i, st in self.assign:
                with torch.cuda.stream(self.streams[st]):
                    self.views[i].forward_pre()
            for st in self.streams:
                main.wait_stream(st)
        else:
            for v in self.views:
                v.forward_pre()
        call("gstex_views_raster_forward", self._table_ptr, main.cuda_stream)
        for v, tgt in zip(self.views, self.targets):
            v.l2_loss_grad(tgt)
        call("gstex_fill_f32", _p(self._grad_records), self._grad_records.numel(), 0.0,
             main.cuda_stream)
        call("gstex_views_raster_backward", self._table_ptr, main.cuda_stream)

    def step(self):
        self.grads["v_texture"].zero_()
        if self.queue:
            self._step_queue()
        elif self.streams:
            main = torch.cuda.current_stream(self.views[0].dev)
            for st in self.streams:
                st.wait_stream(main)
            for i, st in self.assign:
                with torch.cuda.stream(self.streams[st]):
                    self.views[i].step(self.targets[i])
            for st in self.streams:
                main.wait_stream(st)
        else:
            for v, tgt in zip(self.views, self.targets):
                v.step(tgt)
        self._batch_gradients()

    def _batch_gradients(self):
        v0, g, p = self.views[0], self.grads, self.params
        call("gstex_finalize_gradients_views", v0.n, len(self.views), _p(self._grad_records),
             _p(self._records), self._rec_stride, _p(p["quats"]), _p(p["scales"]),
             _p(self._viewmats), _p(self._c2ws), _p(self._focals), _p(g["v_colors"]),
             _p(g["v_opacity"]), _p(g["v_means"]), _p(g["v_scales"]), _p(g["v_quats"]),
             _p(g["v_uv0"]), _p(g["v_umap"]), _p(g["v_vmap"]), _p(self._v_colors),
             v0._stream())
        if self.sh_degree >= 0:
            call("gstex_compute_sh_backward_views", v0.n, len(self.views), self.sh_degree,
                 self.sh_degree, _p(self._viewdirs), _p(self._v_colors),
                 _p(g["v_sh_coeffs"]), v0._stream())

    def capacity_ok(self) -> bool:
        return all(v.capacity_ok() for v in self.views)

    def capture(self):
        dev = self.views[0].dev
        side = torch.cuda.Stream(dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            for _ in range(2):
                self.step()
        torch.cuda.current_stream(dev).wait_stream(side)
        if not self.capacity_ok():
            raise RuntimeError("an intersect list overflowed its capacity; raise capacity_factor")
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.step()
        return self.graph

    def replay(self):
        self.graph.replay()

    def replay_checked(self):
        """Replay and confirm every view's intersects fitted (synchronises)."""
        self.graph.replay()
        if not self.capacity_ok():
            raise RuntimeError("an intersect list overflowed its capacity; recapture larger")


def scene_to_multiview(scene, views, device, bf16: bool = False, **kw) -> MultiViewStep:
    """A MultiViewStep for a ``synthetic.Scene`` and ``synthetic.make_views`` cameras;
    ``bf16`` stores texels and SH coefficients as bfloat16."""
    t = scene.to_torch(device)
    params = dict(means=t["means"], scales=t["scales"], quats=t["quats"],
                  opacities=t["opacities"], uv0=t["uv0"], umap=t["umap"], vmap=t["vmap"],
                  texture=t["texture"], texture_dims=t["texture_dims"])
    if bf16:
        params["texture"] = params["texture"].to(torch.bfloat16)
    if scene.sh_coeffs is not None:
        params["sh_coeffs"] = t["sh_coeffs"].to(torch.bfloat16) if bf16 else t["sh_coeffs"]
    else:
        params["colors"] = t["colors"]
    cams, targets = [], []
    for v in views:
        cams.append(Camera(viewmat=torch.from_numpy(v.viewmat).to(device),
                           c2w=torch.from_numpy(v.c2w).to(device), fx=v.fx, fy=v.fy, cx=v.cx,
                           cy=v.cy, img_h=v.img_h, img_w=v.img_w))
        targets.append(torch.from_numpy(v.target).to(device))
    return MultiViewStep(params, cams, targets, block_width=scene.block_width,
                         settings=scene.settings, background=t["background"],
                         sh_degree=scene.sh_degree, **kw)
