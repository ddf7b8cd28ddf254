"""Benchmark: textured-2DGS forward + backward iterations per second on one GPU.

Workload (default): one 1920x1080 view, 500k textured Gaussians, SH degree 3,
2M texels in area-proportional 1x1..16x16 grids, FP32, L2 loss against a
seeded random target.  One step = SH colours, projection, binning (depth sort
+ tile sort + bin edges), splat records, raster forward, loss gradient, raster
backward, gradient finalize, SH adjoint; inputs are resident in HBM before the
timed region.  Prints one JSON line.

Roofline: the slower of the two raster kernels, timed live with HIP events on
the stream it runs on, against its algorithmic bytes per launch, counted per
unit of work (I intersects, P pixels, T tiles, X texels, C channels, b bytes
per texel value):
  forward   I*108 + P*68 + T*8 + X*C*b      (id + 104-byte Gaussian record per
  backward  I*108 + P*72 + T*8 + 2*X*C*b     intersect; per-pixel outputs / inputs)
``record_model`` beside it counts this build's own 128-byte records and
buffers.  ``step`` is the whole iteration against
  N*(68 + 6*D*b + 24 + 88) + I*(12 + 24 + 8 + 216) + P*(68 + 72) + T*8 + 3*X*C*b.
``traffic`` is the measured HBM bytes per launch of that kernel, read from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
profiles/hbm_traffic_<workload tag>.md.

variants.bf16_batched_graph: views/s of an 8-view batch of the same scene
(seeded cameras around the original, same resolution) with bfloat16 texels and
SH coefficients, the whole batch captured as one graph and replayed; vs_plain
is its views/s over the plain single-view iter/s of the headline view, and
vs_plain_same_cameras over the plain FP32 path rendering the same 8 cameras
one view at a time.

cpu_baseline: the PyTorch CPU path (oracle/torch_path.py: the same semantics,
autograd backward) on the host's cores: the 256x256 / 1k-Gaussian workload
whole (3 warm-up, median of 20), and the benched 1080p frame as full-frame
preprocessing (SH, projection, binning) plus the render fwd+bwd of a random
sample of tiles scaled by its share of intersects.  The same leg checks the
GPU results against the float64 C oracle over sampled tiles, outside every
timed region: image PSNR and per-gradient relative L2 (``vs_f64_oracle`` of
the headline, bf16 batch and 3M bf16 lines).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+bwd iter/s, 1080p 500k textured Gaussians; BF16+batched vs FP32 plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--num-points", type=int, default=None)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=15.0,
                    help="target CPU seconds for the oracle sample (0 disables)")
    ap.add_argument("--batched-views", type=int, default=8,
                    help="also time a batch of this many views with bfloat16 texels/SH, "
                         "captured as one graph (0 disables)")
    ap.add_argument("--train-steps", type=int, default=10,
                    help="timed steps of the training-step variant (0 = skip)")
    ap.add_argument("--config-steps", type=int, default=5,
                    help="timed steps of each of the 1600x1200 regularised, 16-view orbit and "
                         "8-view 3M bf16 workloads (0 = skip)")
    ap.add_argument("--edit-sample", type=int, default=1,
                    help="also time texture_edit on this view and texture_sample on 16M "
                         "queries (0 = skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the timed region, run one device-to-device copy of this many MiB "
                         "(a kernel of known traffic for scaling FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def raster_bytes(kind, I, T, P, X, C, b=4):
    """Algorithmic bytes of one raster launch, per unit of work: an id and a
    104-byte Gaussian record per intersect, the per-pixel outputs (forward,
    68 B) or inputs (backward, 72 B), a bin per tile, the texels read (and in
    the backward their gradients written)."""
    if kind == "raster_forward":
        return I * 108 + P * 68 + T * 8 + X * C * b
    return I * 108 + P * 72 + T * 8 + 2 * X * C * b


def raster_bytes_record_model(kind, I, T, P, X, C, N):
    """The same launch counted with this build's buffers: a 4-byte id and a
    128-byte splat record per intersect, its own per-pixel buffers and the
    128-byte gradient record per Gaussian."""
    if kind == "raster_forward":
        return I * (4 + 128) + T * 8 + P * (56 + 4 * C) + X * C * 4
    return I * (4 + 128) + T * 8 + P * (60 + 4 * C) + 2 * X * C * 4 + N * 128


def step_bytes(N, I, P, T, X, C, D, b=4):
    """Algorithmic bytes of one whole fwd+bwd iteration: projection in/out
    (68), SH coefficients read and their gradient written (2 * 3 D b), colour
    and its gradient (24), parameter gradients (88) per Gaussian; key + id
    (12), one sort pass read + write (24), a bin edge (8) and the id + record
    read in both passes (216) per intersect; forward outputs (68) and backward
    inputs (72) per pixel; a bin per tile; texels read twice and their
    gradient written once."""
    return (N * (68 + 2 * 3 * D * b + 24 + 88) + I * (12 + 24 + 8 + 2 * 108)
            + P * (68 + 72) + T * 8 + 3 * X * C * b)


# file-name tag of each workload's committed rocprofv3 summaries
PROFILE_TAGS = {1: "256px_1k", 2: "1080p_500k", 3: "1600x1200_1m"}


def measured_traffic(workload: int, kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary for this workload (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{PROFILE_TAGS.get(workload, '')}.md")
    if not os.path.exists(path):
        return None, None
    key = f"k_{kernel}_traffic_bytes_per_launch:"
    with open(path) as fh:
        for line in fh:
            if line.startswith(key):
                return int(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def measured_valu(workload: int, kernel: str):
    """(lane utilisation, VALU issue share) of ``kernel`` from the committed
    rocprofv3 SQ counter summary for this workload, or (None, None)."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_{PROFILE_TAGS.get(workload, '')}.md")
    if not os.path.exists(path):
        return None, None
    key = f"- `k_{kernel}`: lane utilisation "
    with open(path) as fh:
        for line in fh:
            if line.startswith(key):
                util, share = line[len(key):].split(", VALU issue share ")
                return float(util), float(share)
    return None, None


def _pick_tiles(bins, k=16, seed=0):
    counts = bins[:, 1] - bins[:, 0]
    nz = np.nonzero(counts > 0)[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(nz, min(k, len(nz)), replace=False))


def quality_job(step, target, seed=0, k_tiles=16):
    """GPU side of the float64-oracle check of an exact-size RenderStep over
    ``k_tiles`` random non-empty tiles: the step's image (out_img +
    out_texture), its L2-loss cotangents, and the gradients of those tiles
    alone (the other tiles' bins emptied before the backward).  Returns host
    arrays; the oracle side runs in the cpu_baseline leg."""
    import torch

    step.forward()
    step.l2_loss_grad(target)
    bins = step.tile_bins.cpu().numpy()
    pick = _pick_tiles(bins, k_tiles, seed)
    keep = torch.zeros(bins.shape[0], dtype=torch.bool, device=step.dev)
    keep[torch.from_numpy(pick).to(step.dev)] = True
    step.tile_bins[~keep] = 0
    step.backward()
    torch.cuda.synchronize()
    return dict(pick=pick, bins=bins, ids=step.ids[: step.num_intersects].cpu().numpy(),
                img=(step.out_img + step.out_texture[..., :3]).cpu().numpy(),
                v_img=step.v_img.cpu().numpy(), v_tex=step.v_tex.cpu().numpy(),
                colors=step.colors.float().cpu().numpy(),
                grads={k: v.float().cpu().numpy() for k, v in step.grads.items()
                       if k != "v_sh_coeffs"})


def _view_step(mv, i, **kw):
    """An exact-size single-view RenderStep over a batch's parameter tensors
    (same storage formats) and its view i's camera."""
    from gstex_cuda_amd.pipeline import RenderStep

    v = mv.views[i]
    return RenderStep(mv.params, v.cam, block_width=v.bw, settings=v.settings,
                      background=v.background, sh_degree=v.sh_degree, **kw)


def plain_same_cameras(scene, views, dev, steps, warmup):
    """Views/s of the plain FP32 single-view path (exact-size steps, one view
    at a time) over the given cameras."""
    import torch

    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    total = 0.0
    for v in views:
        vs = synthetic.view_scene(scene, v)
        st = scene_to_step(vs, dev)
        tgt = torch.from_numpy(vs.target).to(dev)
        total += _timed(lambda: st.step(tgt), steps, warmup) / steps
        del st
    torch.cuda.empty_cache()
    return len(views) / total


def batched_variant(scene, dev, num_views, steps, warmup):
    """Views/s of a multi-view batch of the same scene (seeded cameras around the
    original one, same resolution), bfloat16 texels and SH coefficients, the
    whole batch (SH, projection, binning, raster forward, loss, raster
    backward, finalize, SH adjoint per view, gradient sum) captured as one
    graph and replayed."""
    import torch

    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_multiview

    views = synthetic.make_views(scene, num_views)
    mv = scene_to_multiview(scene, views, dev, bf16=True)
    mv.capture()
    for _ in range(warmup):
        mv.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mv.replay()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ok = mv.capacity_ok()
    ms = 1e3 * elapsed / steps
    out = {"views": num_views, "views_per_s": round(num_views * steps / elapsed, 3),
           "ms_per_batch": round(ms, 4), "dtype": "bf16 storage, f32 compute",
           "graph": True, "capacity_ok": bool(ok)}
    # precision check of view 0 (same storage formats, kernels and camera)
    vstep = _view_step(mv, 0)
    job = quality_job(vstep, mv.targets[0], seed=1)
    job.update(view=views[0], bf16=True)
    del vstep, mv
    torch.cuda.empty_cache()
    plain = plain_same_cameras(scene, views, dev, steps, warmup)
    out["plain_same_cameras_views_per_s"] = round(plain, 3)
    out["vs_plain_same_cameras"] = round(out["views_per_s"] / plain, 3)
    return out, job


def train_variant(scene, dev, steps, warmup):
    """Iterations/s of the whole training step of example.py's trainer on this
    scene: preprocessing of the raw parameters, the render forward + backward
    (example loss: texture MSE + distortion + normal terms), the preprocessing
    adjoint and an Adam step on every raw tensor (means, log scales, quaternions,
    opacity logits, SH coefficients, texel logits)."""
    import torch

    from gstex_cuda_amd.pipeline import Camera
    from gstex_cuda_amd.train import TrainStep, raw_from_scene

    raw, dims = raw_from_scene(scene, dev)
    cam = Camera(viewmat=torch.from_numpy(scene.viewmat).to(dev),
                 c2w=torch.from_numpy(scene.c2w).to(dev), fx=scene.fx, fy=scene.fy,
                 cx=scene.cx, cy=scene.cy, img_h=scene.img_h, img_w=scene.img_w)
    target = torch.from_numpy(scene.target).to(dev)
    ts = TrainStep(raw, dims, cam, target, lr=1e-3, sh_degree=scene.sh_degree,
                   block_width=scene.block_width, settings=scene.settings)
    for _ in range(warmup):
        ts.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ts.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    out = {"iter_per_s": round(steps / elapsed, 3), "ms_per_step": round(1e3 * elapsed / steps, 4),
           "steps": steps, "dtype": "f32", "optimizer": "adam (fused, one kernel per tensor)",
           "loss": round(ts.loss(), 6)}
    del ts
    torch.cuda.empty_cache()
    return out


def _timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def regularised_variant(dev, steps, warmup):
    """Iterations/s of fwd+bwd on one 1600x1200 view of 1M untextured Gaussians
    (one texel each) with the L2 loss plus the distortion (weight 100) and
    normal (weight 0.05) regularisers."""
    import torch

    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_step

    scene = synthetic.make_scene(3)
    step = scene_to_step(scene, dev)
    target = torch.from_numpy(scene.target).to(dev)
    elapsed = _timed(lambda: step.regularised_step(target, 100.0, 0.05), steps, warmup)
    out = {"workload": f"1 view {scene.img_w}x{scene.img_h}, {scene.num_points} untextured "
                       "Gaussians, distortion 100 + normal 0.05",
           "iter_per_s": round(steps / elapsed, 3), "ms_per_step": round(1e3 * elapsed / steps, 4),
           "intersects": int(step.num_intersects), "dtype": "f32", "steps": steps}
    del step
    torch.cuda.empty_cache()
    return out


def multiview_variant(dev, workload, steps, warmup):
    """Batches/s of a multi-view workload captured as one graph: 16 orbit views
    of mixed resolution over 1M Gaussians in float32 (workload 4), or 8 views at
    1920x1080 over 3M Gaussians with bfloat16 texels and SH coefficients and the
    float64 ray/plane solve for near-edge-on pairs (workload 5)."""
    import torch

    from gstex_cuda_amd import cuda as _C
    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import scene_to_multiview

    from gstex_cuda_amd.autotune import BatchConfig, tune_multiview

    bf16 = workload == 5
    scene = synthetic.make_scene(workload)
    views = synthetic.workload_views(workload, scene)
    prev = _C.set_precise_edge_on(bf16)
    job = None
    try:
        # the batch's launch configuration, chosen by timing the candidates
        # (the 16-view workload also tries 8-pixel tiles: its tile size is tuned)
        widths = (16, 8) if workload == 4 else (16,)
        streams = (4, 8) if workload == 4 else (4,)
        cands = [BatchConfig(q, st, bw) for bw in widths for q in (False, True) for st in streams]
        best, tried = tune_multiview(scene, views, dev, bf16=bf16, candidates=cands)
        scene.block_width = best.block_width
        mv = scene_to_multiview(scene, views, dev, bf16=bf16, queue=best.queue,
                                streams=best.streams)
        mv.capture()
        elapsed = _timed(mv.replay, steps, warmup)
        ok = mv.capacity_ok()
        isects = [int(v.num_intersects) for v in mv.views]
        if bf16:
            # precision check of view 0 (same storage formats, kernels, camera
            # and float64 edge-on solve)
            vstep = _view_step(mv, 0)
            job = quality_job(vstep, mv.targets[0], seed=5)
            job.update(view=views[0], bf16=True, scene=scene)
            del vstep
    finally:
        _C.set_precise_edge_on(prev)
    res = sorted({f"{v.img_w}x{v.img_h}" for v in views})
    out = {"workload": f"{len(views)} views ({', '.join(res)}), {scene.num_points} Gaussians, "
                       f"{scene.texture.shape[0]} texels, SH degree {scene.sh_degree}"
                       + (", bf16 storage, f64 edge-on solve" if bf16 else ", f32"),
           "batches_per_s": round(steps / elapsed, 3),
           "views_per_s": round(len(views) * steps / elapsed, 3),
           "ms_per_batch": round(1e3 * elapsed / steps, 4),
           "intersects_per_batch": int(sum(isects)), "intersects_min_max": [min(isects), max(isects)],
           "graph": True, "capacity_ok": bool(ok), "steps": steps,
           "tuned": {"chosen": {"queue": best.queue, "streams": best.streams,
                                "block_width": best.block_width},
                     "candidates_ms_per_batch": tried}}
    del mv
    torch.cuda.empty_cache()
    # the plain FP32 single-view path on the same cameras, one view at a time
    scene.block_width = 16
    plain = plain_same_cameras(scene, views, dev, steps, warmup)
    out["plain_same_cameras_views_per_s"] = round(plain, 3)
    out["vs_plain_same_cameras"] = round(out["views_per_s"] / plain, 3)
    return out, job


def edit_sample_variant(scene, step, dev, reps=5):
    """Timing of texture_edit (texture_edit.cu:229-346) over the benched view's
    sorted lists, and of texture_sample forward and texel adjoint
    (texture_sample.cu:11-141) on 16M queries over the view's charts; HIP
    events on the launch stream, mean of ``reps`` calls after one warm-up."""
    import torch

    from gstex_cuda_amd import cuda as _C

    def timed_ms(fn):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    t = scene.to_torch(dev)
    H, W, N = scene.img_h, scene.img_w, scene.num_points
    X, C = scene.texture.shape
    torch.manual_seed(0)
    img = torch.rand((H, W, 3), device=dev)
    alpha = torch.ones((H, W), device=dev)
    lo = torch.zeros((H, W), device=dev)
    hi = torch.full((H, W), 1e9, device=dev)
    ids = step.ids[: step.num_intersects]
    bw = scene.block_width

    def edit():
        return _C.texture_edit(step.tiles, (bw, bw, 1), (W, H, 1), (N, 1, 5), X,
                               t["texture_dims"], img, alpha, lo, hi, ids, step.tile_bins,
                               t["opacities"], t["means"], t["scales"], 1.0, t["quats"],
                               t["uv0"], t["umap"], t["vmap"], t["viewmat"], t["c2w"],
                               scene.fx, scene.fy, scene.cx, scene.cy, 0, t["background"])

    edit_ms = timed_ms(edit)
    Q = 1 << 24
    dims = t["texture_dims"][torch.randint(0, N, (Q,), device=dev)].contiguous()
    uvs = torch.rand((Q, 2), device=dev)
    v_out = torch.rand((Q, C), device=dev)
    fwd_ms = timed_ms(lambda: _C.texture_sample_forward((N, 1, C), dims, uvs, t["texture"]))
    bwd_ms = timed_ms(lambda: _C.texture_sample_backward((N, 1, C), dims, uvs, t["texture"],
                                                         v_out))
    # per query: chart row 12 B, uv 8 B, 4 taps x C values read (forward) or
    # added (adjoint), C values written (forward) or read (adjoint); the adjoint
    # also clears the (X, C) texel gradient
    q_bytes = 12 + 8 + 16 * C + 4 * C
    fwd_gbs = Q * q_bytes / (fwd_ms * 1e-3) / 1e9
    bwd_gbs = (Q * q_bytes + X * C * 4) / (bwd_ms * 1e-3) / 1e9
    out = {"texture_edit_ms": round(edit_ms, 4),
           "texture_edit_workload": f"{W}x{H}, {int(step.num_intersects)} intersects, 5 "
                                    "channels, depth window unbounded (records built per call)",
           "texture_sample_queries": Q,
           "texture_sample_fwd_ms": round(fwd_ms, 4),
           "texture_sample_fwd_GBps": round(fwd_gbs, 1),
           "texture_sample_fwd_frac_hbm": round(fwd_gbs / HBM_PEAK_GBS, 4),
           "texture_sample_bwd_ms": round(bwd_ms, 4),
           "texture_sample_bwd_GBps": round(bwd_gbs, 1),
           "texture_sample_bwd_frac_hbm": round(bwd_gbs / HBM_PEAK_GBS, 4)}
    del dims, uvs, v_out
    torch.cuda.empty_cache()
    return out


def _oracle_quality(O, scene, job):
    """The float64 C oracle against a quality job: image PSNR over the sampled
    tiles and the relative L2 error of every gradient of those tiles, on the
    stored values (bfloat16-rounded texels and SH coefficients for a bfloat16
    job) and, for a bfloat16 job, on the original float32 values as well."""
    import torch

    from gstex_cuda_amd import synthetic

    vs = synthetic.view_scene(scene, job["view"]) if job.get("view") is not None else scene
    H, W, bw = vs.img_h, vs.img_w, vs.block_width
    tx = (W + bw - 1) // bw
    pick = job["pick"]
    mask = np.zeros(job["bins"].shape[0], np.uint8)
    mask[pick] = 1
    sel = np.zeros((H, W), bool)
    for t in pick:
        ty, t0 = divmod(int(t), tx)
        sel[ty * bw:(ty + 1) * bw, t0 * bw:(t0 + 1) * bw] = True

    def rounded(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(torch.bfloat16).float().numpy()

    bf = bool(job.get("bf16"))
    sh = vs.sh_coeffs
    inputs = [("stored", rounded(vs.texture) if bf else vs.texture,
               None if sh is None else (rounded(sh) if bf else sh))]
    if bf:
        inputs.append(("float32_inputs", vs.texture, sh))
    z1 = np.zeros((H, W), np.float32)
    z3 = np.zeros((H, W, 3), np.float32)
    out = {"view": 0, "tiles": int(len(pick)), "pixels": int(sel.sum())}
    for name, tex, coeffs in inputs:
        if coeffs is not None:
            cols = O.sh_forward(vs.sh_degree, vs.sh_degree, vs.means - vs.c2w[:3, 3], coeffs,
                                "f64")
        else:
            cols = vs.colors
        sc = O.Scene(means=vs.means, scales=vs.scales, quats=vs.quats, colors=cols,
                     opacities=vs.opacities, uv0=vs.uv0, umap=vs.umap, vmap=vs.vmap, texture=tex,
                     texture_dims=vs.texture_dims, viewmat=vs.viewmat, c2w=vs.c2w,
                     background=vs.background, fx=vs.fx, fy=vs.fy, cx=vs.cx, cy=vs.cy, img_h=H,
                     img_w=W, block_width=bw, settings=vs.settings, ids_sorted=job["ids"],
                     tile_bins=job["bins"], tile_mask=mask)
        f = O.texture_forward(sc, "f64")
        ref = f["out_img"] + f["out_texture"][..., :3]
        img = job["img"].astype(np.float64)
        mse = float(np.mean((img[sel] - ref[sel]) ** 2))
        peak = max(float(np.abs(ref[sel]).max()), 1e-12)
        out[f"image_psnr_db_{name}"] = (round(10.0 * np.log10(peak * peak / mse), 2)
                                        if mse > 0 else float("inf"))
        g = O.texture_backward(sc, f, job["v_img"], z1, z1, z1, job["v_tex"], z3, "f64")
        rel = {}
        for k, ref_g in g.items():
            got = job["grads"][k].astype(np.float64).ravel()
            r = ref_g.astype(np.float64).ravel()
            rel[k] = float(np.linalg.norm(got - r) / max(np.linalg.norm(r), 1e-30))
        out[f"grad_rel_l2_{name}"] = {k: float(f"{v:.3e}") for k, v in rel.items()}
        out[f"grad_rel_l2_max_{name}"] = float(f"{max(rel.values()):.3e}")
    return out


def cpu_threads():
    """Threads for the CPU path: the host's CPUs available to this process,
    capped by OMP_NUM_THREADS when it is set (the share of the host a job on a
    shared GPU box is given)."""
    import os as _os

    n = _os.cpu_count() or 1
    try:
        n = min(n, len(_os.sched_getaffinity(0)))
    except AttributeError:
        pass
    cap = _os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(n, 1)


def cpu_baseline(scene, target_seconds, jobs):
    """The PyTorch CPU path timed on the host cores (the 256x256 / 1k workload
    whole, and the benched frame by a tile sample), then the float64 oracle
    checks of the GPU quality jobs.  Returns (cpu_baseline, {job: check})."""
    import torch

    import oracle as O
    from gstex_cuda_amd import synthetic
    from oracle import torch_path as TP

    threads = cpu_threads()
    torch.set_num_threads(threads)
    host = TP.host_cpu()
    budget = 250_000  # pairs per chunk

    # 256x256, 1k Gaussians, 4x4 RGB texels: 3 warm-up, median of 20
    s1 = synthetic.make_scene(1)
    p1, cam1, tgt1 = TP.leaf_params(s1), TP.camera_of(s1), torch.tensor(s1.target)
    times = []
    for _ in range(23):
        t0 = time.perf_counter()
        TP.train_iteration(p1, cam1, tgt1, s1.sh_degree, s1.block_width, s1.settings,
                           pair_budget=budget)
        times.append(time.perf_counter() - t0)
    c1_ms = 1e3 * float(np.median(times[3:]))

    # the benched frame: full-frame SH, projection and binning, then the render
    # fwd+bwd of a growing random sample of tiles, scaled by intersect share
    p2, cam2, tgt2 = TP.leaf_params(scene), TP.camera_of(scene), torch.tensor(scene.target)
    t0 = time.perf_counter()
    q, ids, bins = TP.preprocess(p2, cam2, scene.sh_degree, scene.block_width)
    t_pre = time.perf_counter() - t0
    col = q.get("colors")
    if "sh_coeffs" in p2:
        q = dict(q)
        q["colors"] = col.detach().requires_grad_(True)
    counts = (bins[:, 1] - bins[:, 0]).numpy().astype(np.int64)
    total = int(counts.sum())
    order = np.random.default_rng(0).permutation(len(counts))
    share, k, chosen, t_render = 0.002, 1, 0, 0.0
    while True:
        k = max(1, int(len(order) * share))
        t0 = time.perf_counter()
        TP.render_backward(q, cam2, ids, bins, scene.block_width, scene.settings, target=tgt2,
                           tiles=torch.from_numpy(order[:k]), pair_budget=budget)
        t_render = time.perf_counter() - t0
        chosen = int(counts[order[:k]].sum())
        if t_render >= 0.3 * target_seconds or share >= 1.0 or chosen == total:
            break
        share = min(1.0, share * max(2.0, 0.5 * target_seconds / max(t_render, 1e-3)))
    t_sh = 0.0
    if "sh_coeffs" in p2 and q["colors"].grad is not None:
        t0 = time.perf_counter()
        col.backward(q["colors"].grad)
        t_sh = time.perf_counter() - t0
    frame_s = t_pre + t_sh + t_render * total / max(chosen, 1)
    cpu = {"value": 1.0 / frame_s, "unit": "iter/s", "cores": threads, "kind": "port",
           "sample": f"PyTorch CPU path (autograd), {threads} threads: full-frame SH + "
                     f"projection + binning {t_pre:.2f} s, render fwd+bwd of {k} of "
                     f"{len(counts)} tiles ({chosen} of {total} intersects) {t_render:.1f} s, "
                     f"scaled by intersect share",
           "os_cpu_count": host["os_cpu_count"], "affinity_cpus": host["affinity_cpus"],
           "cpu_model": host["cpu_model"],
           "c1_256px_1k_fwd_bwd_median_ms": round(c1_ms, 2),
           "c1_256px_1k_iter_per_s": round(1e3 / c1_ms, 3),
           "c1_timing": "3 warm-up, median of 20"}
    del p2, q, ids, bins
    checks = {}
    if jobs:
        O.build()
        for name, job in jobs.items():
            checks[name] = _oracle_quality(O, job.get("scene", scene), job)
    return cpu, checks


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "is not built: every box has a single MI355X.", file=sys.stderr)
        sys.exit(2)
    import torch

    from gstex_cuda_amd import synthetic
    from gstex_cuda_amd.pipeline import StageTimer, scene_to_step

    dev = torch.device("cuda:0")
    scene = synthetic.make_scene(args.workload, num_points=args.num_points)
    step = scene_to_step(scene, dev)
    target = torch.from_numpy(scene.target).to(dev)
    for _ in range(args.warmup):
        step.step(target)
    torch.cuda.synchronize()
    timer = StageTimer()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.step(target, timer)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    if args.copy_calibration_mb > 0:
        nbytes_copy = args.copy_calibration_mb << 20
        src = torch.ones(nbytes_copy // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    I, N = step.num_intersects, step.n
    T = step.tiles[0] * step.tiles[1]
    P = scene.img_h * scene.img_w
    X, C = scene.texture.shape
    kern = {k: timer.mean_ms(k) for k in ("raster_forward", "raster_backward")}
    dom = max(kern, key=kern.get)
    nbytes = raster_bytes(dom, I, T, P, X, C)
    achieved = nbytes / (kern[dom] * 1e-3) / 1e9
    rec_bytes = raster_bytes_record_model(dom, I, T, P, X, C, N)
    D = (scene.sh_degree + 1) ** 2 if scene.sh_degree >= 0 else 0
    sbytes = step_bytes(N, I, P, T, X, C, D)
    step_gbs = sbytes / (elapsed / args.steps) / 1e9
    traffic, traffic_src = (measured_traffic(args.workload, dom) if args.num_points is None
                            else (None, None))
    util, share = (measured_valu(args.workload, dom) if args.num_points is None
                   else (None, None))
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes": int(nbytes),
            "bytes_per_unit": ("I*108 + P*72 + T*8 + 2*X*C*4" if dom == "raster_backward"
                               else "I*108 + P*68 + T*8 + X*C*4"),
            "record_model": {"algorithmic_bytes": int(rec_bytes),
                             "achieved": round(rec_bytes / (kern[dom] * 1e-3) / 1e9, 2),
                             "frac": round(rec_bytes / (kern[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                           5)},
            "step": {"algorithmic_bytes": int(sbytes), "achieved": round(step_gbs, 2),
                     "frac": round(step_gbs / HBM_PEAK_GBS, 5)},
            "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
            "valu_issue_share": share, "valu_lane_utilisation": util}
    variants, jobs = {}, {}
    if args.batched_views > 0:
        b, jobs["bf16_batched_graph"] = batched_variant(
            scene, dev, args.batched_views, max(args.steps // 4, 3), max(args.warmup // 2, 1))
        b["vs_plain"] = round(b["views_per_s"] / value, 3)
        variants["bf16_batched_graph"] = b
    if args.train_steps > 0:
        t = train_variant(scene, dev, args.train_steps, max(args.warmup // 2, 1))
        t["vs_plain"] = round(t["iter_per_s"] / value, 3)
        variants["train_step_adam"] = t
    if args.edit_sample:
        variants["texture_edit_and_sample"] = edit_sample_variant(scene, step, dev)
    # precision check of the benched view (after every use of its lists)
    jobs["headline"] = quality_job(step, target, seed=0)
    del step
    torch.cuda.empty_cache()
    if args.config_steps > 0:
        w = max(args.warmup // 2, 1)
        variants["c3_regularised_1600x1200_1m"] = regularised_variant(dev, args.config_steps, w)
        variants["c4_orbit_16_views_1m_graph"], _ = multiview_variant(dev, 4, args.config_steps,
                                                                      w)
        c5, c5_job = multiview_variant(dev, 5, args.config_steps, w)
        variants["c5_bf16_8x1080p_3m_graph"] = c5
        if c5_job is not None:
            jobs["c5"] = c5_job
    cpu, quality = None, {}
    if args.cpu_baseline_seconds > 0:
        cpu, quality = cpu_baseline(scene, args.cpu_baseline_seconds, jobs)
    if "bf16_batched_graph" in quality:
        variants["bf16_batched_graph"]["vs_f64_oracle"] = quality["bf16_batched_graph"]
    if "c5" in quality:
        variants["c5_bf16_8x1080p_3m_graph"]["vs_f64_oracle"] = quality["c5"]
    line = {"metric": METRIC, "value": round(value, 4), "unit": "iter/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"single view {scene.img_w}x{scene.img_h}, {N} Gaussians, "
                                   f"SH degree {max(scene.sh_degree, 0)}, {X} texels, "
                                   f"FP32 plain path",
                       "num_gaussians": N, "resolution": f"{scene.img_w}x{scene.img_h}",
                       "texels": X, "channels": C, "intersects": I, "block_width": 16},
            "roofline": roof, "cpu_baseline": cpu, "vs_f64_oracle": quality.get("headline"),
            "variants": variants}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
