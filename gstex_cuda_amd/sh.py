"""Drop-in for ``gstex_cuda.sh`` (sh.py:1-96): real SH colours, differentiable
w.r.t. the coefficients."""
from __future__ import annotations

from torch import Tensor
from torch.autograd import Function

from . import cuda as _C


def num_sh_bases(degree: int) -> int:
    return {0: 1, 1: 4, 2: 9, 3: 16}.get(int(degree), 25)


def deg_from_sh(num_bases: int) -> int:
    table = {1: 0, 4: 1, 9: 2, 16: 3, 25: 4}
    if num_bases not in table:
        raise AssertionError("Invalid number of SH bases")
    return table[num_bases]


def spherical_harmonics(degrees_to_use: int, viewdirs: Tensor, coeffs: Tensor) -> Tensor:
    """Colours (N, 3) from coeffs (N, D, 3) along viewdirs (N, 3) (sh.py:36-57)."""
    assert coeffs.shape[-2] >= num_sh_bases(degrees_to_use)
    return _SphericalHarmonics.apply(degrees_to_use, viewdirs.contiguous(), coeffs.contiguous())


class _SphericalHarmonics(Function):
    @staticmethod
    def forward(ctx, degrees_to_use: int, viewdirs: Tensor, coeffs: Tensor):
        num_points = coeffs.shape[0]
        degree = deg_from_sh(coeffs.shape[-2])
        ctx.degrees_to_use = degrees_to_use
        ctx.degree = degree
        ctx.save_for_backward(viewdirs)
        return _C.compute_sh_forward(num_points, degree, degrees_to_use, viewdirs, coeffs)

    @staticmethod
    def backward(ctx, v_colors: Tensor):
        viewdirs = ctx.saved_tensors[0]
        v_coeffs = _C.compute_sh_backward(v_colors.shape[0], ctx.degree, ctx.degrees_to_use,
                                          viewdirs, v_colors.contiguous())
        return None, None, v_coeffs
