"""Drop-in for ``gstex_cuda.texture_sample`` (texture_sample.py:1-81)."""
from __future__ import annotations

from torch import Tensor
from torch.autograd import Function

from . import cuda as _C


def texture_sample(texture_info, texture_dims: Tensor, texture: Tensor, uvs: Tensor,
                   use_torch_impl: bool = False) -> Tensor:
    """Bilinear queries into the jagged texel tensor -> (Q, C).

    Differentiable w.r.t. ``texture`` (the reference returns no gradient here,
    texture_sample.py:70-78); uvs get no gradient, as in the reference.
    """
    if use_torch_impl:
        raise NotImplementedError("gstex_cuda_amd has no torch implementation path")
    return _TextureSample.apply(texture_info, texture_dims.contiguous(), texture.contiguous(),
                                uvs.contiguous())


class _TextureSample(Function):
    @staticmethod
    def forward(ctx, texture_info, texture_dims, texture, uvs):
        ctx.texture_info = texture_info
        ctx.save_for_backward(texture_dims, texture, uvs)
        return _C.texture_sample_forward(texture_info, texture_dims, uvs, texture)

    @staticmethod
    def backward(ctx, v_out):
        texture_dims, texture, uvs = ctx.saved_tensors
        v_texture = None
        if ctx.needs_input_grad[2]:
            v_texture = _C.texture_sample_backward(ctx.texture_info, texture_dims, uvs, texture,
                                                   v_out.contiguous())
        return None, None, v_texture, None
