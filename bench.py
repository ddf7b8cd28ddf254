"""Benchmark: textured-2DGS forward + backward iterations per second on one GPU.

Workload (default): one 1920x1080 view, 500k textured Gaussians, SH degree 3,
2M texels in area-proportional 1x1..16x16 grids, FP32, L2 loss against a
seeded random target.  One step = SH colours, projection, binning (depth sort
+ tile sort + bin edges), splat records, raster forward, loss gradient, raster
backward, gradient finalize, SH adjoint; inputs are resident in HBM before the
timed region.  Prints one JSON line.

Roofline: the slower of the two raster kernels, timed live with HIP events on
the stream it runs on, against its algorithmic bytes per launch (``traffic`` is
the measured HBM bytes per launch of that kernel, read from the committed
rocprofv3 FETCH_SIZE / WRITE_SIZE summary profiles/hbm_traffic_<workload tag>.md):
  forward   I*(4+128) + T*8 + P*(56 + 4C) + X*C*4
  backward  I*(4+128) + T*8 + P*(60 + 4C) + 2*X*C*4 + N*128
(I intersects, T tiles, P pixels, X texels, C channels, N Gaussians).

variants.bf16_batched_graph: views/s of an 8-view batch of the same scene
(seeded cameras around the original, same resolution) with bfloat16 texels and
SH coefficients, the whole batch captured as one graph and replayed; vs_plain
is its views/s over the plain single-view iter/s.

cpu_baseline: the FP32 C oracle (single thread, the reference's semantics) on a
bounded sample of whole tiles of the same frame, forward + backward, scaled to
the full frame by the sampled share of intersects.
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


def raster_bytes(kind, I, T, P, X, C, N):
    if kind == "raster_forward":
        return I * (4 + 128) + T * 8 + P * (56 + 4 * C) + X * C * 4
    return I * (4 + 128) + T * 8 + P * (60 + 4 * C) + 2 * X * C * 4 + N * 128


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
    del mv
    torch.cuda.empty_cache()
    return out


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

    bf16 = workload == 5
    scene = synthetic.make_scene(workload)
    views = synthetic.workload_views(workload, scene)
    prev = _C.set_precise_edge_on(bf16)
    try:
        mv = scene_to_multiview(scene, views, dev, bf16=bf16)
        mv.capture()
        elapsed = _timed(mv.replay, steps, warmup)
        ok = mv.capacity_ok()
        isects = [int(v.num_intersects) for v in mv.views]
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
           "graph": True, "capacity_ok": bool(ok), "steps": steps}
    del mv
    torch.cuda.empty_cache()
    return out


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


def cpu_baseline(scene, step, target_seconds):
    """Time the FP32 oracle on a sample of tiles and scale to one full frame."""
    import oracle as O

    O.build()
    ids = step.ids[: step.num_intersects].cpu().numpy()
    bins = step.tile_bins.cpu().numpy()
    colors = step.colors.cpu().numpy()
    counts = (bins[:, 1] - bins[:, 0]).astype(np.int64)
    total = int(counts.sum())
    rng = np.random.default_rng(0)
    order = rng.permutation(len(counts))
    H, W, C = scene.img_h, scene.img_w, scene.texture.shape[1]
    sc = O.Scene(means=scene.means, scales=scene.scales, quats=scene.quats, colors=colors,
                 opacities=scene.opacities, uv0=scene.uv0, umap=scene.umap, vmap=scene.vmap,
                 texture=scene.texture, texture_dims=scene.texture_dims, viewmat=scene.viewmat,
                 c2w=scene.c2w, background=scene.background, fx=scene.fx, fy=scene.fy,
                 cx=scene.cx, cy=scene.cy, img_h=H, img_w=W, block_width=scene.block_width,
                 settings=scene.settings, ids_sorted=ids, tile_bins=bins)
    v = np.zeros((H, W, 3), np.float32)
    vt = np.zeros((H, W, C), np.float32)
    z = np.zeros((H, W), np.float32)
    # grow the sample until the measured time reaches the target
    share, chosen, elapsed = 0.002, 0, 0.0
    while True:
        k = max(1, int(len(order) * share))
        mask = np.zeros(len(counts), np.uint8)
        mask[order[:k]] = 1
        sc.tile_mask = mask
        t0 = time.perf_counter()
        f = O.texture_forward(sc, "f32")
        O.texture_backward(sc, f, v, z, z, z, vt, v, "f32")
        elapsed = time.perf_counter() - t0
        chosen = int(counts[order[:k]].sum())
        if elapsed >= 0.3 * target_seconds or share >= 1.0 or chosen == total:
            break
        share = min(1.0, share * max(2.0, 0.5 * target_seconds / max(elapsed, 1e-3)))
    frame_s = elapsed * total / max(chosen, 1)
    # the timed run's own image, checked against the float64 oracle on 16 of
    # the sampled tiles (outside the timed region)
    chk = np.zeros(len(counts), np.uint8)
    chk[order[:16]] = 1
    sc.tile_mask = chk
    f64 = O.texture_forward(sc, "f64")
    bw, tx = scene.block_width, (W + scene.block_width - 1) // scene.block_width
    sel = np.zeros((H, W), bool)
    for t in order[:16]:
        ty, tx0 = divmod(int(t), tx)
        sel[ty * bw:(ty + 1) * bw, tx0 * bw:(tx0 + 1) * bw] = True
    img = (step.out_img + step.out_texture[..., :3]).cpu().numpy().astype(np.float64)
    ref = f64["out_img"] + f64["out_texture"][..., :3]
    mse = float(np.mean((img[sel] - ref[sel]) ** 2))
    peak = max(float(np.abs(ref[sel]).max()), 1e-12)
    psnr = float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)
    return {"value": 1.0 / frame_s, "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": f"{k} of {len(counts)} tiles ({chosen} of {total} intersects), "
                      f"FP32 oracle fwd+bwd {elapsed:.1f} s, scaled by intersect share",
            "image_psnr_db_vs_f64_oracle": round(psnr, 2),
            "psnr_sample": f"{int(sel.sum())} pixels of 16 sampled tiles, out_img + out_texture"}


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
    nbytes = raster_bytes(dom, I, T, P, X, C, N)
    achieved = nbytes / (kern[dom] * 1e-3) / 1e9
    traffic, traffic_src = (measured_traffic(args.workload, dom) if args.num_points is None
                            else (None, None))
    util, share = (measured_valu(args.workload, dom) if args.num_points is None
                   else (None, None))
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes": int(nbytes),
            "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
            "valu_issue_share": share, "valu_lane_utilisation": util}
    variants = {}
    if args.batched_views > 0:
        b = batched_variant(scene, dev, args.batched_views, max(args.steps // 4, 3),
                            max(args.warmup // 2, 1))
        b["vs_plain"] = round(b["views_per_s"] / value, 3)
        variants["bf16_batched_graph"] = b
    if args.train_steps > 0:
        t = train_variant(scene, dev, args.train_steps, max(args.warmup // 2, 1))
        t["vs_plain"] = round(t["iter_per_s"] / value, 3)
        variants["train_step_adam"] = t
    if args.edit_sample:
        variants["texture_edit_and_sample"] = edit_sample_variant(scene, step, dev)
    if args.config_steps > 0:
        w = max(args.warmup // 2, 1)
        variants["c3_regularised_1600x1200_1m"] = regularised_variant(dev, args.config_steps, w)
        variants["c4_orbit_16_views_1m_graph"] = multiview_variant(dev, 4, args.config_steps, w)
        variants["c5_bf16_8x1080p_3m_graph"] = multiview_variant(dev, 5, args.config_steps, w)
    cpu = None
    if args.cpu_baseline_seconds > 0:
        cpu = cpu_baseline(scene, step, args.cpu_baseline_seconds)
    line = {"metric": METRIC, "value": round(value, 4), "unit": "iter/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * elapsed / args.steps, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"single view {scene.img_w}x{scene.img_h}, {N} Gaussians, "
                                   f"SH degree {max(scene.sh_degree, 0)}, {X} texels, "
                                   f"FP32 plain path",
                       "num_gaussians": N, "resolution": f"{scene.img_w}x{scene.img_h}",
                       "texels": X, "channels": C, "intersects": I, "block_width": 16},
            "roofline": roof, "cpu_baseline": cpu, "variants": variants}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
