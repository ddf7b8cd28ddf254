"""Training step of a GStex scene on the gfx950 raster path.

Mirrors the reference's example training loop (``example.py:119-295``):

* raw parameters are preprocessed into the renderer's inputs
  (``example.py:126-143``): ``scales.xy = exp(log_scales.xy)`` with
  ``scales.z = 1e-5 * mean(scales.xy)`` (detached), ``quats = q / |q|``,
  ``umap / vmap`` from the mapping (uv0, log uv scale, theta) and the rotation
  axes, and sigmoids of the colour, opacity and texel logits;
* the view is rendered and scored with the example loss
  (``example.py:188-210``): ``mse(out_texture[..., :3], target) +
  mean(out_reg) + mean(n_x^2 + n_y^2 + (1 - n_z)^2)`` (weights adjustable);
* every raw parameter takes a ``torch.optim.Adam`` step
  (``example.py:231-233,277``; the mapping only when ``train_mapping``, as
  ``example.py:117`` freezes it).

Preprocessing and its adjoint are one kernel each
(``gstex_preprocess_gaussians[_backward]``); Adam is one fused kernel per
tensor (``gstex_adam_step``).  The texel step chains the sigmoid itself and
writes ``sigmoid(logits)`` straight into the texel buffer the next render
reads (float32 or bfloat16).
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Optional, Sequence

import torch

from . import _lib
from ._lib import call
from .pipeline import Camera, RenderStep, _p

GRAD_PLAIN = 0
GRAD_SIGMOID = 1


def _stream(t: torch.Tensor):
    return torch.cuda.current_stream(t.device).cuda_stream


def _check(name: str, t: torch.Tensor, shape: Sequence[int], dtype=torch.float32):
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


def preprocess_gaussians(log_scales: torch.Tensor, quats: torch.Tensor, mapping: torch.Tensor,
                         rgb_logits: Optional[torch.Tensor], opacity_logits: torch.Tensor,
                         out: Optional[Dict[str, torch.Tensor]] = None) -> Dict[str, torch.Tensor]:
    """Render inputs from raw parameters (``example.py:126-143``).

    Returns (or fills ``out`` with) ``scales`` (N,3), ``quats`` (N,4), ``uv0``
    (N,1,2), ``umap``, ``vmap`` (N,1,3), ``opacities`` (N,1) and, when
    ``rgb_logits`` is given, ``colors`` (N,3).
    """
    n = log_scales.shape[0]
    _check("log_scales", log_scales, (n, 3))
    _check("quats", quats, (n, 4))
    _check("mapping", mapping, (n, 1, 4))
    _check("opacity_logits", opacity_logits, (n, 1))
    if rgb_logits is not None:
        _check("rgb_logits", rgb_logits, (n, 3))
    if out is None:
        f32 = dict(dtype=torch.float32, device=log_scales.device)
        out = dict(scales=torch.empty((n, 3), **f32), quats=torch.empty((n, 4), **f32),
                   uv0=torch.empty((n, 1, 2), **f32), umap=torch.empty((n, 1, 3), **f32),
                   vmap=torch.empty((n, 1, 3), **f32), opacities=torch.empty((n, 1), **f32))
        if rgb_logits is not None:
            out["colors"] = torch.empty((n, 3), **f32)
    call("gstex_preprocess_gaussians", n, _p(log_scales), _p(quats), _p(mapping),
         _p(rgb_logits), _p(opacity_logits), _p(out["scales"]), _p(out["quats"]), _p(out["uv0"]),
         _p(out["umap"]), _p(out["vmap"]), _p(out.get("colors")), _p(out["opacities"]),
         _stream(log_scales))
    return out


def preprocess_gaussians_backward(log_scales: torch.Tensor, quats: torch.Tensor,
                                  mapping: torch.Tensor, rgb_logits: Optional[torch.Tensor],
                                  opacity_logits: torch.Tensor, grads: Dict[str, torch.Tensor],
                                  train_mapping: bool = True,
                                  out: Optional[Dict[str, torch.Tensor]] = None
                                  ) -> Dict[str, torch.Tensor]:
    """Adjoint of :func:`preprocess_gaussians`.

    ``grads`` holds the render-input gradients under the names the raster
    backward uses (``v_scales``, ``v_quats``, ``v_uv0``, ``v_umap``, ``v_vmap``,
    ``v_colors``, ``v_opacity``).  Returns ``log_scales``, ``quats``,
    ``opacity_logits`` and, as applicable, ``mapping`` and ``rgb_logits``
    gradients.
    """
    n = log_scales.shape[0]
    if out is None:
        f32 = dict(dtype=torch.float32, device=log_scales.device)
        out = dict(log_scales=torch.empty((n, 3), **f32), quats=torch.empty((n, 4), **f32),
                   opacity_logits=torch.empty((n, 1), **f32))
        if train_mapping:
            out["mapping"] = torch.empty((n, 1, 4), **f32)
        if rgb_logits is not None:
            out["rgb_logits"] = torch.empty((n, 3), **f32)
    call("gstex_preprocess_gaussians_backward", n, _p(log_scales), _p(quats), _p(mapping),
         _p(rgb_logits), _p(opacity_logits), _p(grads["v_scales"]), _p(grads["v_quats"]),
         _p(grads["v_uv0"]), _p(grads["v_umap"]), _p(grads["v_vmap"]),
         _p(grads.get("v_colors") if rgb_logits is not None else None), _p(grads["v_opacity"]),
         _p(out["log_scales"]), _p(out["quats"]), _p(out.get("mapping")),
         _p(out.get("rgb_logits")), _p(out["opacity_logits"]), _stream(log_scales))
    return out


def adam_step(param: torch.Tensor, grad: torch.Tensor, exp_avg: torch.Tensor,
              exp_avg_sq: torch.Tensor, step: int, lr: float, betas=(0.9, 0.999),
              eps: float = 1e-8, sigmoid_grad: bool = False,
              values: Optional[torch.Tensor] = None) -> None:
    """One in-place ``torch.optim.Adam`` step for one float32 tensor.

    ``sigmoid_grad``: ``grad`` is the gradient of ``sigmoid(param)``.
    ``values`` (float32 or bfloat16, same number of elements) receives
    ``sigmoid`` of the updated parameter.
    """
    for name, t in (("param", param), ("grad", grad), ("exp_avg", exp_avg),
                    ("exp_avg_sq", exp_avg_sq)):
        _check(name, t, param.shape)
    fmt = 0
    if values is not None:
        if values.numel() != param.numel() or not values.is_contiguous():
            raise RuntimeError("values must be contiguous with as many elements as param")
        if values.dtype == torch.bfloat16:
            fmt = 1
        elif values.dtype != torch.float32:
            raise RuntimeError("values must be float32 or bfloat16")
    call("gstex_adam_step", param.numel(), _p(param), _p(grad), _p(exp_avg), _p(exp_avg_sq),
         float(lr), float(betas[0]), float(betas[1]), float(eps), int(step),
         GRAD_SIGMOID if sigmoid_grad else GRAD_PLAIN, _p(values), fmt, _stream(param))


class FusedAdam:
    """Adam over a fixed list of float32 tensors, one fused kernel each.

    ``step(grads)`` takes the gradients in the order of ``params``.  Per-tensor
    options: ``sigmoid_grad`` and ``values`` as in :func:`adam_step`.
    """

    def __init__(self, params: Sequence[torch.Tensor], lr: float = 1e-3, betas=(0.9, 0.999),
                 eps: float = 1e-8, sigmoid_grad: Optional[Sequence[bool]] = None,
                 values: Optional[Sequence[Optional[torch.Tensor]]] = None):
        self.params = list(params)
        self.lr, self.betas, self.eps = lr, tuple(betas), eps
        self.exp_avg = [torch.zeros_like(p) for p in self.params]
        self.exp_avg_sq = [torch.zeros_like(p) for p in self.params]
        self.sigmoid_grad = list(sigmoid_grad) if sigmoid_grad else [False] * len(self.params)
        self.values = list(values) if values else [None] * len(self.params)
        self.steps = 0

    def step(self, grads: Iterable[torch.Tensor]) -> None:
        self.steps += 1
        for k, g in enumerate(grads):
            adam_step(self.params[k], g, self.exp_avg[k], self.exp_avg_sq[k], self.steps,
                      self.lr, self.betas, self.eps, self.sigmoid_grad[k], self.values[k])


def raw_from_scene(scene, device) -> tuple:
    """Raw parameters (``TrainStep`` layout) and texture_dims for a
    ``synthetic.Scene``: log scales, the scene's quaternions, logits of its
    opacities and texels (clamped to (1e-4, 1 - 1e-4)), a mapping with the
    scene's uv0, an isotropic uv scale of 1 / (6 sqrt(s_x s_y)) and theta 0, and
    its SH coefficients or colour logits."""
    import numpy as np

    f32 = np.float32
    sc = scene.scales.astype(np.float64)
    n = sc.shape[0]
    mapping = np.zeros((n, 1, 4), f32)
    mapping[:, 0, :2] = scene.uv0.reshape(n, 2)
    mapping[:, 0, 2] = np.log(1.0 / (6.0 * np.sqrt(sc[:, 0] * sc[:, 1])))

    def logit(x):
        x = np.clip(np.asarray(x, np.float64), 1e-4, 1 - 1e-4)
        return np.log(x / (1 - x)).astype(f32)

    raw = dict(means=scene.means, log_scales=np.log(np.maximum(sc, 1e-12)).astype(f32),
               quats=scene.quats, mapping=mapping, opacity_logits=logit(scene.opacities),
               texture_logits=logit(scene.texture))
    if scene.sh_coeffs is not None:
        raw["sh_coeffs"] = scene.sh_coeffs
    else:
        raw["rgb_logits"] = logit(scene.colors)
    raw = {k: torch.from_numpy(np.ascontiguousarray(v, dtype=f32)).to(device)
           for k, v in raw.items()}
    return raw, torch.from_numpy(scene.texture_dims).to(device)


class TrainStep:
    """One training iteration of ``example.py``'s trainer on the native path.

    ``raw`` (float32 device tensors): ``means`` (N,3), ``log_scales`` (N,3),
    ``quats`` (N,4), ``mapping`` (N,1,4), ``opacity_logits`` (N,1),
    ``texture_logits`` (X,C) and either ``rgb_logits`` (N,3) or ``sh_coeffs``
    (N,D,3) with ``sh_degree``.  ``texture_dims`` (N,3) int32 is constant.
    ``step()`` preprocesses, renders, applies the example loss gradient,
    backpropagates and takes one Adam step on every trainable tensor, all on
    the device with no host synchronisation (unless ``capacity`` is None, when
    the intersect count is read as in the reference).
    """

    def __init__(self, raw: Dict[str, torch.Tensor], texture_dims: torch.Tensor, cam: Camera,
                 target: torch.Tensor, lr: float = 1e-2, betas=(0.9, 0.999), eps: float = 1e-8,
                 train_mapping: bool = False, texture_bf16: bool = False,
                 block_width: int = 16, settings: int = 1 << 8,
                 background: Optional[torch.Tensor] = None, sh_degree: int = -1,
                 reg_weight: float = 1.0, normal_weight: float = 1.0,
                 capacity: Optional[int] = None):
        self.raw = raw
        self.cam = cam
        self.target = target
        self.train_mapping = train_mapping
        self.reg_weight, self.normal_weight = reg_weight, normal_weight
        use_rgb = "rgb_logits" in raw
        self.use_rgb = use_rgb
        self.pre = preprocess_gaussians(raw["log_scales"], raw["quats"], raw["mapping"],
                                        raw.get("rgb_logits"), raw["opacity_logits"])
        tex_values = torch.sigmoid(raw["texture_logits"])
        if texture_bf16:
            tex_values = tex_values.to(torch.bfloat16)
        params = dict(means=raw["means"], scales=self.pre["scales"], quats=self.pre["quats"],
                      opacities=self.pre["opacities"], uv0=self.pre["uv0"],
                      umap=self.pre["umap"], vmap=self.pre["vmap"], texture=tex_values,
                      texture_dims=texture_dims)
        if use_rgb:
            params["colors"] = self.pre["colors"]
        else:
            params["sh_coeffs"] = raw["sh_coeffs"]
        self.render = RenderStep(params, cam, block_width=block_width, settings=settings,
                                 background=background, sh_degree=-1 if use_rgb else sh_degree,
                                 capacity=capacity)
        H, W = cam.img_h, cam.img_w
        dev = raw["means"].device
        self.v_reg = torch.full((H, W), reg_weight / (H * W), dtype=torch.float32, device=dev)
        self.v_normal = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        self.raw_grads = None
        order = ["means", "log_scales", "quats", "opacity_logits",
                 "rgb_logits" if use_rgb else "sh_coeffs", "texture_logits"]
        if train_mapping:
            order.append("mapping")
        self.order = order
        self.adam = FusedAdam([raw[k] for k in order], lr, betas, eps,
                              sigmoid_grad=[k == "texture_logits" for k in order],
                              values=[tex_values if k == "texture_logits" else None
                                      for k in order])

    def loss_grad(self):
        """Cotangents of the example loss (``example.py:188-210``)."""
        r, H, W = self.render, self.cam.img_h, self.cam.img_w
        # texture MSE (the example renders out_texture only) and the normal term
        call("gstex_image_loss_grad", H * W, r.C, 0, None, _p(r.out_texture), _p(self.target),
             2.0 / (H * W * 3), _p(r.out_normal), 2.0 * self.normal_weight / (H * W),
             _p(r.v_img), _p(r.v_tex), _p(self.v_normal), _stream(self.target))

    def loss(self) -> float:
        """The example loss of the last forward (a host synchronisation)."""
        r = self.render
        gt = torch.mean((r.out_texture[..., :3] - self.target) ** 2)
        n = r.out_normal
        normal = torch.mean(n[..., 0] ** 2 + n[..., 1] ** 2 + (1 - n[..., 2]) ** 2)
        return float(gt + self.reg_weight * torch.mean(r.out_reg) + self.normal_weight * normal)

    def step(self):
        raw = self.raw
        preprocess_gaussians(raw["log_scales"], raw["quats"], raw["mapping"],
                             raw.get("rgb_logits"), raw["opacity_logits"], out=self.pre)
        r = self.render
        r.forward()
        self.loss_grad()
        r.backward(v_reg=self.v_reg, v_normal=self.v_normal)
        self.raw_grads = preprocess_gaussians_backward(
            raw["log_scales"], raw["quats"], raw["mapping"], raw.get("rgb_logits"),
            raw["opacity_logits"], r.grads, self.train_mapping, out=self.raw_grads)
        g = r.grads
        by_name = dict(means=g["v_means"], log_scales=self.raw_grads["log_scales"],
                       quats=self.raw_grads["quats"],
                       opacity_logits=self.raw_grads["opacity_logits"],
                       texture_logits=g["v_texture"])
        if self.use_rgb:
            by_name["rgb_logits"] = self.raw_grads["rgb_logits"]
        else:
            by_name["sh_coeffs"] = g["v_sh_coeffs"]
        if self.train_mapping:
            by_name["mapping"] = self.raw_grads["mapping"]
        self.adam.step([by_name[k] for k in self.order])
