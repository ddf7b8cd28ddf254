"""gstex_texture_forward_noalloc: the forward with its splat records in the
library's fixed per-device scratch array instead of a caller workspace.

Through the C ABI it must equal gstex_texture_forward (workspace from the
caller) bit for bit on every output, on two calls in a row, the second with
fewer Gaussians and a smaller image (the scratch keeps the first call's
records, which the second must not read), and reject more Gaussians than the
scratch holds before touching the device."""
import ctypes

import numpy as np
import pytest
import torch

OUT_NAMES = ("out_img", "out_depth", "out_reg", "out_texture", "out_normal", "final_Ts",
             "final_idx", "depth_idx", "out_reg_s")
SCRATCH_BYTES = 96 << 20
MAX_POINTS = (SCRATCH_BYTES - 256) // 156


def test_noalloc_rejects_more_points_than_the_scratch_holds():
    from gstex_cuda_amd import _lib

    fake = ctypes.c_void_p(0x1000)  # never dereferenced: the size check fails first
    rc = _lib.lib().gstex_texture_forward_noalloc(
        2, 2, 16, 32, 32, 1, 1, 3, fake, fake, 4, fake, fake, fake, fake, fake, MAX_POINTS + 1,
        1.0, fake, fake, fake, fake, fake, 1, fake, fake, 1.0, 1.0, 1.0, 1.0, 0, fake,
        fake, fake, fake, fake, fake, fake, fake, fake, fake, None)
    err = _lib.lib().gstex_last_error().decode()
    assert rc == 1, rc
    assert f"holds at most {MAX_POINTS} Gaussians" in err, err


def _outputs(H, W, C, dev):
    f = dict(dtype=torch.float32, device=dev)
    i = dict(dtype=torch.int32, device=dev)
    # filled with a pattern so that an output left unwritten cannot match
    return dict(out_img=torch.full((H, W, 3), 7.0, **f), out_depth=torch.full((H, W), 7.0, **f),
                out_reg=torch.full((H, W), 7.0, **f), out_texture=torch.full((H, W, C), 7.0, **f),
                out_normal=torch.full((H, W, 3), 7.0, **f), final_Ts=torch.full((H, W), 7.0, **f),
                final_idx=torch.full((H, W), 7, **i), depth_idx=torch.full((H, W), 7, **i),
                out_reg_s=torch.full((H, W, 3), 7.0, **f))


def _forward(name, s, pb, r, outs, ws=None):
    from gstex_cuda_amd._lib import call

    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    bw = s.block_width
    tx, ty = (s.img_w + bw - 1) // bw, (s.img_h + bw - 1) // bw
    C = s.texture.shape[1]
    args = [tx, ty, bw, s.img_w, s.img_h, s.num_points, 1, C, p(r["texture_dims"]),
            p(pb["ids"]), int(pb["ids"].numel()), p(pb["bins"]), p(r["colors"]),
            p(r["opacities"]), p(r["means"]), p(r["scales"]), s.num_points, 1.0, p(r["quats"]),
            p(r["uv0"]), p(r["umap"]), p(r["vmap"]), p(r["texture"]), int(s.texture.shape[0]),
            p(r["viewmat"]), p(r["c2w"]), s.fx, s.fy, s.cx, s.cy, s.settings, p(r["background"])]
    args += [p(outs[k]) for k in OUT_NAMES]
    if ws is not None:
        args += [p(ws), ws.numel()]
    args.append(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    call(name, *args)


@pytest.mark.gpu
def test_noalloc_forward_equals_workspace_forward_on_two_calls():
    from gpu_helpers import DEV, log, project_and_bin, raster_inputs

    from gstex_cuda_amd import _lib, synthetic

    cases = [synthetic.make_scene(1, seed=21, num_points=600, img_w=112, img_h=80),
             synthetic.make_scene(1, seed=22, num_points=150, img_w=48, img_h=64)]
    for k, s in enumerate(cases):
        log(f"noalloc call {k}: N={s.num_points} {s.img_w}x{s.img_h}")
        pb = project_and_bin(s)
        r = raster_inputs(s)
        C = s.texture.shape[1]
        nbytes = _lib.lib().gstex_texture_forward_workspace_size(s.num_points)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=DEV)
        ref = _outputs(s.img_h, s.img_w, C, DEV)
        got = _outputs(s.img_h, s.img_w, C, DEV)
        _forward("gstex_texture_forward", s, pb, r, ref, ws)
        _forward("gstex_texture_forward_noalloc", s, pb, r, got)
        torch.cuda.synchronize()
        assert pb["num_intersects"] > 0
        for name in OUT_NAMES:
            assert torch.equal(got[name], ref[name]), (k, name)
        img = ref["out_img"].cpu().numpy()
        assert np.isfinite(img).all() and float(img.max()) > 0.0
