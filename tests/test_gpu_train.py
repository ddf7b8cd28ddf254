"""Training-step glue against plain PyTorch float32.

* gstex_preprocess_gaussians equals example.py:126-143 written in torch;
* its adjoint equals torch autograd through the same formulas;
* gstex_adam_step equals torch.optim.Adam over several steps (plain and
  sigmoid-chained, float32 and bfloat16 value outputs);
* TrainStep, three iterations on a small scene, follows example.py's training
  loop composed from torch preprocessing, the native autograd
  texture_gaussians (itself checked against the oracle) and torch.optim.Adam.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from gpu_helpers import DEV, log, np_

pytestmark = pytest.mark.gpu


def _rotmat(q):
    # _torch_impl.py:116-133 (normalised quaternion, w x y z)
    w, x, y, z = torch.unbind(q, dim=-1)
    m = torch.stack([1 - 2 * (y ** 2 + z ** 2), 2 * (x * y - w * z), 2 * (x * z + w * y),
                     2 * (x * y + w * z), 1 - 2 * (x ** 2 + z ** 2), 2 * (y * z - w * x),
                     2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x ** 2 + y ** 2)], -1)
    return m.reshape(q.shape[:-1] + (3, 3))


def _torch_pre(raw):
    # example.py:126-143
    scales = torch.zeros_like(raw["log_scales"])
    scales[:, :2] = torch.exp(raw["log_scales"][:, :2])
    scales[:, -1] = 1e-5 * torch.mean(scales[:, :-1], dim=-1).detach()
    quats = raw["quats"] / raw["quats"].norm(dim=-1, keepdim=True)
    rs = _rotmat(quats)
    uv0 = raw["mapping"][:, :, :2]
    uvscale = torch.exp(raw["mapping"][:, :, None, 2])
    theta = raw["mapping"][:, :, None, 3]
    ax1, ax2 = rs[:, None, :, 0], rs[:, None, :, 1]
    umap = uvscale * (ax1 * torch.cos(theta) + ax2 * torch.sin(theta))
    vmap = uvscale * (-ax1 * torch.sin(theta) + ax2 * torch.cos(theta))
    out = dict(scales=scales, quats=quats, uv0=uv0, umap=umap.reshape(-1, 1, 3),
               vmap=vmap.reshape(-1, 1, 3), opacities=torch.sigmoid(raw["opacity_logits"]))
    if "rgb_logits" in raw:
        out["colors"] = torch.sigmoid(raw["rgb_logits"])
    return out


def _raw(n, seed, th=3, tw=3, spread=16.0):
    """example.py:70-113 initialisation (seeded, CPU generator)."""
    g = torch.Generator().manual_seed(seed)
    means = torch.rand(n, 3, generator=g) - 0.5
    means[:, :2] *= spread
    log_scales = 0.5 * math.log(1 / n) * torch.rand(n, 3, generator=g)
    u, v, w = (torch.rand(n, 1, generator=g) for _ in range(3))
    quats = torch.cat([torch.sqrt(1 - u) * torch.sin(2 * math.pi * v),
                       torch.sqrt(1 - u) * torch.cos(2 * math.pi * v),
                       torch.sqrt(u) * torch.sin(2 * math.pi * w),
                       torch.sqrt(u) * torch.cos(2 * math.pi * w)], -1)
    mapping = torch.zeros(n, 1, 4)
    mapping[:, :, :2] = 0.5
    scale = 0.25 * math.sqrt(1 / float(torch.sum(torch.exp(log_scales[:, 0] + log_scales[:, 1]))))
    mapping[:, :, 2] = math.log(scale)
    mapping[:, :, 3] = 2 * math.pi * torch.rand(n, 1, generator=g)
    raw = dict(means=means, log_scales=log_scales, quats=quats, mapping=mapping,
               rgb_logits=torch.rand(n, 3, generator=g) * 4 - 2,
               opacity_logits=torch.randn(n, 1, generator=g),
               texture_logits=torch.rand(n * th * tw, 3, generator=g) * 4 - 2)
    dims = torch.zeros(n, 3, dtype=torch.int32)
    dims[:, 0], dims[:, 1] = th, tw
    dims[:, 2] = torch.cumsum(dims[:, 0] * dims[:, 1], 0) - dims[:, 0] * dims[:, 1]
    return {k: t.to(DEV).contiguous() for k, t in raw.items()}, dims.to(DEV)


def test_preprocess_matches_torch():
    from gstex_cuda_amd.train import preprocess_gaussians

    raw, _ = _raw(5000, 1)
    raw["mapping"][:, :, 3] = torch.linspace(-7.0, 7.0, 5000, device=DEV)[:, None]
    got = preprocess_gaussians(raw["log_scales"], raw["quats"], raw["mapping"], raw["rgb_logits"],
                               raw["opacity_logits"])
    want = _torch_pre(raw)
    for k in want:
        torch.testing.assert_close(got[k], want[k].reshape(got[k].shape), rtol=2e-6, atol=2e-7,
                                   msg=k)


def test_preprocess_backward_matches_autograd():
    from gstex_cuda_amd.train import preprocess_gaussians_backward

    raw, _ = _raw(5000, 2)
    leaf = {k: t.clone().requires_grad_(True) for k, t in raw.items()}
    pre = _torch_pre(leaf)
    g = torch.Generator(device=DEV).manual_seed(3)
    up = {k: torch.randn(t.shape, generator=g, device=DEV) for k, t in pre.items()}
    sum(torch.sum(pre[k] * up[k]) for k in pre).backward()
    grads = dict(v_scales=up["scales"], v_quats=up["quats"], v_uv0=up["uv0"].contiguous(),
                 v_umap=up["umap"], v_vmap=up["vmap"], v_colors=up["colors"],
                 v_opacity=up["opacities"])
    got = preprocess_gaussians_backward(raw["log_scales"], raw["quats"], raw["mapping"],
                                        raw["rgb_logits"], raw["opacity_logits"], grads)
    for k in ("log_scales", "quats", "mapping", "rgb_logits", "opacity_logits"):
        torch.testing.assert_close(got[k], leaf[k].grad, rtol=1e-5, atol=1e-6, msg=k)
    # frozen mapping: no mapping gradient is written
    got2 = preprocess_gaussians_backward(raw["log_scales"], raw["quats"], raw["mapping"],
                                         raw["rgb_logits"], raw["opacity_logits"], grads,
                                         train_mapping=False)
    assert "mapping" not in got2
    torch.testing.assert_close(got2["quats"], got["quats"], rtol=0, atol=0)


@pytest.mark.parametrize("sigmoid", [False, True], ids=["plain", "sigmoid"])
@pytest.mark.parametrize("fmt", ["f32", "bf16"])
def test_adam_matches_torch(sigmoid, fmt):
    from gstex_cuda_amd.train import adam_step

    n = 20000
    g = torch.Generator(device=DEV).manual_seed(5)
    p0 = torch.randn(n, generator=g, device=DEV)
    ours, m, v = p0.clone(), torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2)
    values = torch.empty(n, device=DEV, dtype=torch.bfloat16 if fmt == "bf16" else torch.float32)
    for step in range(1, 6):
        up = torch.randn(n, generator=g, device=DEV) * (10.0 ** (step - 3))
        up[:100] = 0.0  # parameters without gradient
        opt.zero_grad()
        ((torch.sigmoid(ref) if sigmoid else ref) * up).sum().backward()
        opt.step()
        adam_step(ours, up, m, v, step, 1e-2, sigmoid_grad=sigmoid,
                  values=values if sigmoid else None)
        st = opt.state[ref]
        torch.testing.assert_close(ours, ref.detach(), rtol=1e-6, atol=1e-6)
        torch.testing.assert_close(m, st["exp_avg"], rtol=1e-6, atol=1e-9)
        torch.testing.assert_close(v, st["exp_avg_sq"], rtol=1e-6, atol=1e-12)
        if sigmoid:
            want = torch.sigmoid(ref.detach())
            if fmt == "bf16":
                # round-to-nearest-even of float32 sigmoids within 1 ulp of torch's
                diff = (values.view(torch.int16).int() - want.to(torch.bfloat16).view(
                    torch.int16).int()).abs()
                assert int(diff.max()) <= 1
            else:
                torch.testing.assert_close(values, want, rtol=2e-6, atol=1e-7)


def _camera(W, H, dev):
    from gstex_cuda_amd.pipeline import Camera

    viewmat = torch.tensor([[1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 1.0, 8.0], [0, 0, 0, 1.0]],
                           device=dev)
    focal = 0.5 * W / math.tan(0.25 * math.pi)
    return Camera(viewmat=viewmat, c2w=torch.linalg.inv(viewmat).contiguous(), fx=focal,
                  fy=focal, cx=W / 2, cy=H / 2, img_h=H, img_w=W), focal


def _reference_loop(raw, dims, target, cam, focal, iters, lr):
    """example.py:119-278 with this package's native ops in place of gstex_cuda's."""
    from gstex_cuda_amd.get_aabb_2d import get_aabb_2d, get_num_tiles_hit_2d, project_points
    from gstex_cuda_amd.texture import texture_gaussians

    leaf = {k: t.clone().requires_grad_(k != "mapping") for k, t in raw.items()}
    opt = torch.optim.Adam([leaf[k] for k in ("rgb_logits", "means", "log_scales",
                                              "opacity_logits", "quats", "mapping",
                                              "texture_logits")], lr)
    n = raw["means"].shape[0]
    H, W = cam.img_h, cam.img_w
    intrins = (focal, focal, W / 2, H / 2)
    losses = []
    for _ in range(iters):
        opt.zero_grad()
        pre = _torch_pre(leaf)
        _, depths = project_points(leaf["means"], cam.viewmat[:3, :], intrins)
        centers, extents = get_aabb_2d(leaf["means"], pre["scales"], 1, pre["quats"],
                                       cam.viewmat, intrins)
        nt = get_num_tiles_hit_2d(centers, extents, H, W, 16)
        out = texture_gaussians((n, 1, 3), dims, centers, extents, depths, nt, pre["colors"],
                                pre["opacities"], leaf["means"], pre["scales"], 1, pre["quats"],
                                pre["uv0"], pre["umap"], pre["vmap"],
                                torch.sigmoid(leaf["texture_logits"]), cam.viewmat, cam.c2w,
                                focal, focal, W / 2, H / 2, H, W, 16, 1 << 8,
                                torch.zeros(3, device=DEV))
        img, reg, normal = out[4][:, :, :3], out[2], out[5]
        loss = (torch.mean((img - target) ** 2) + torch.mean(reg) +
                torch.mean(normal[:, :, 0] ** 2 + normal[:, :, 1] ** 2 +
                           (1 - normal[:, :, 2]) ** 2))
        losses.append(float(loss.detach()))
        loss.backward()
        opt.step()
    return {k: t.detach() for k, t in leaf.items()}, losses


def test_train_step_follows_example_loop():
    from gstex_cuda_amd.train import TrainStep

    n, W, H, iters, lr = 400, 96, 80, 3, 1e-2
    raw, dims = _raw(n, 7)
    cam, focal = _camera(W, H, DEV)
    target = torch.rand(H, W, 3, generator=torch.Generator().manual_seed(8)).to(DEV)
    want, want_losses = _reference_loop(raw, dims, target, cam, focal, iters, lr)
    mine = {k: t.clone() for k, t in raw.items()}
    ts = TrainStep(mine, dims, cam, target, lr=lr)
    losses = []
    for _ in range(iters):
        ts.step()
        losses.append(ts.loss())
    log(f"losses ours {losses} reference {want_losses}")
    # ts.loss() reads the forward of the same iteration as the reference's loss
    np.testing.assert_allclose(losses, want_losses, rtol=1e-5)
    for k in raw:
        a, b = np_(mine[k]).ravel(), np_(want[k]).ravel()
        close = np.abs(a - b) <= 1e-5 + 1e-4 * np.abs(b)
        log(f"  {k}: {close.mean():.5f} of elements within tolerance, max diff "
            f"{np.abs(a - b).max():.2e}")
        # Adam's first steps are sign-like, so a gradient within float noise of
        # zero may step the other way: require all but a few elements to agree
        assert close.mean() >= 0.999, k
    assert not torch.equal(mine["means"], raw["means"])


@pytest.mark.parametrize("channels,with_img", [(3, 1), (5, 1), (1, 1), (3, 0), (7, 0)])
def test_image_loss_grad_matches_torch_steps(channels, with_img):
    # gstex_image_loss_grad equals the float32 torch sequence it replaces:
    # add, subtract, multiply (and the normal term's multiply and subtract)
    from gstex_cuda_amd._lib import call

    g = torch.Generator().manual_seed(channels * 10 + with_img)
    H, W = 37, 53
    img = torch.rand((H, W, 3), generator=g).to(DEV)
    tex = torch.rand((H, W, channels), generator=g).to(DEV)
    tgt = torch.rand((H, W, 3), generator=g).to(DEV)
    nrm = torch.randn((H, W, 3), generator=g).to(DEV)
    scale, ns = 2.0 / (H * W * 3), 2.0 * 0.05 / (H * W)
    v_img = torch.full((H, W, 3), float("nan"), device=DEV)
    v_tex = torch.full((H, W, channels), float("nan"), device=DEV)
    v_n = torch.full((H, W, 3), float("nan"), device=DEV)
    call("gstex_image_loss_grad", H * W, channels, with_img, img.data_ptr(), tex.data_ptr(),
         tgt.data_ptr(), scale, nrm.data_ptr(), ns, v_img.data_ptr(), v_tex.data_ptr(),
         v_n.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    t3 = torch.zeros((H, W, 3), device=DEV)
    t3[..., :min(channels, 3)] = tex[..., :3]
    want = ((img + t3) if with_img else t3).sub(tgt).mul(scale)
    assert torch.equal(v_img, want if with_img else torch.zeros_like(want))
    assert torch.equal(v_tex[..., :min(channels, 3)], want[..., :min(channels, 3)])
    if channels > 3:
        assert torch.equal(v_tex[..., 3:], torch.zeros_like(v_tex[..., 3:]))
    wn = torch.mul(nrm, ns)
    wn[..., 2].sub_(ns)
    assert torch.equal(v_n, wn)
