"""Measured choice of raster launch options for a scene shape.

Launch options are timed on the caller's own scene and the fastest
combination is kept:

* ``use_wave_masks`` -- the backward visits only entries the forward composited;
* ``culling`` -- per-wave screen-box culling of list entries;
* ``block_width`` -- tile side (only the caller's by default).

The first two never change results (up to float summation order).  The tile
side does, slightly, as in the reference: a pixel sees every Gaussian whose
+-3 sigma AABB meets its tile, and alpha >= 1/255 reaches past 3 sigma for
opacities above 0.35, so pixels just outside an AABB are lit or not depending
on the tiling.  Pass ``block_widths=(16, 8)`` to let the tuner choose it.

A trial is the HIP-event duration of one whole step (SH, projection, binning,
raster forward, loss gradient, raster backward, finalize) after a warm-up step.  Results are cached per
(N, width, height, channels, settings, texel format) in memory and, when
``cache_path`` is given, in a JSON file.

``tune_multiview`` does the same for a multi-view batch captured as one graph,
over the batch's own knobs:

* ``queue`` -- one raster launch per direction over every view's tiles from a
  longest-first queue, or each view's own launches on the streams;
* ``streams`` -- HIP streams the views' stages run on;
* ``block_width`` -- tile side (only when the caller lists more than one).
"""
from __future__ import annotations

import itertools
import json
import os
from dataclasses import asdict, dataclass
from typing import Dict, Iterable, Optional

import torch

from .pipeline import Camera, RenderStep, StageTimer


@dataclass(frozen=True)
class RasterConfig:
    block_width: int = 16
    use_wave_masks: bool = True
    culling: bool = True


_CACHE: Dict[str, dict] = {}


def _key(params, cam: Camera, settings: int) -> str:
    tex = params["texture"]
    return (f"n{params['means'].shape[0]}_w{cam.img_w}_h{cam.img_h}_c{tex.shape[1]}"
            f"_s{settings}_{str(tex.dtype).split('.')[-1]}")


def _time_config(params, cam, target, cfg: RasterConfig, settings, background, sh_degree,
                 trials: int) -> float:
    step = RenderStep(params, cam, block_width=cfg.block_width, settings=settings,
                      background=background, sh_degree=sh_degree,
                      use_wave_masks=cfg.use_wave_masks, culling=cfg.culling)
    step.step(target)
    torch.cuda.synchronize()
    timer = StageTimer()
    for _ in range(trials):
        timer.start("all")
        step.step(target)
        timer.stop("all")
    torch.cuda.synchronize()
    return timer.mean_ms("all")


def tune(params, cam: Camera, target: torch.Tensor, settings: int = 1 << 8,
         background: Optional[torch.Tensor] = None, sh_degree: int = -1,
         block_widths: Iterable[int] = (16,), trials: int = 3,
         cache_path: Optional[str] = None) -> RasterConfig:
    """Time every candidate configuration on this scene and return the fastest."""
    key = _key(params, cam, settings)
    if cache_path and os.path.exists(cache_path) and not _CACHE:
        with open(cache_path) as fh:
            _CACHE.update(json.load(fh))
    if key in _CACHE:
        return RasterConfig(**_CACHE[key]["best"])
    results = []
    for bw, masks, cull in itertools.product(block_widths, (True, False), (True, False)):
        cfg = RasterConfig(bw, masks, cull)
        ms = _time_config(params, cam, target, cfg, settings, background, sh_degree, trials)
        results.append((ms, cfg))
    results.sort(key=lambda r: r[0])
    best = results[0][1]
    _CACHE[key] = {"best": asdict(best),
                   "timings_ms": {f"bw{c.block_width}_masks{int(c.use_wave_masks)}"
                                  f"_cull{int(c.culling)}": round(ms, 4) for ms, c in results}}
    if cache_path:
        with open(cache_path, "w") as fh:
            json.dump(_CACHE, fh, indent=1)
    return best


def tuned_step(params, cam: Camera, target: torch.Tensor, **kw) -> RenderStep:
    """A RenderStep built with the tuned configuration.

    ``block_widths`` (default: the caller's ``block_width``, else 16) are the tile
    sides the tuner may choose among; the tuned choice replaces any
    ``block_width`` / ``use_wave_masks`` / ``culling`` passed here.  The culling
    choice belongs to the returned step (its raster launches set it and restore
    the previous process-wide value), so other steps are unaffected."""
    bw = kw.pop("block_width", None)
    kw.pop("use_wave_masks", None)
    kw.pop("culling", None)
    widths = tuple(kw.pop("block_widths", (bw if bw is not None else 16,)))
    cfg = tune(params, cam, target, settings=kw.get("settings", 1 << 8),
               background=kw.get("background"), sh_degree=kw.get("sh_degree", -1),
               block_widths=widths)
    return RenderStep(params, cam, block_width=cfg.block_width,
                      use_wave_masks=cfg.use_wave_masks, culling=cfg.culling, **kw)


def last_timings() -> Dict[str, dict]:
    return dict(_CACHE)


@dataclass(frozen=True)
class BatchConfig:
    queue: bool = False
    streams: int = 4
    block_width: int = 16


def tune_multiview(scene, views, device, bf16: bool = False,
                   candidates: Optional[Iterable[BatchConfig]] = None, trials: int = 3):
    """Time each candidate configuration of a multi-view batch of ``scene``
    seen from ``views`` (graph capture, one warm replay, ``trials`` timed
    replays) and return (fastest config, {config name: ms per batch})."""
    from .pipeline import scene_to_multiview

    if candidates is None:
        candidates = [BatchConfig(q, st, scene.block_width) for q in (False, True)
                      for st in (4, 8)]
    bw0 = scene.block_width
    timings = {}
    best, best_ms = None, float("inf")
    try:
        for cfg in candidates:
            scene.block_width = cfg.block_width
            mv = scene_to_multiview(scene, views, device, bf16=bf16, queue=cfg.queue,
                                    streams=cfg.streams)
            mv.capture()
            mv.replay_checked()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(trials):
                mv.replay()
            t1.record()
            t1.synchronize()
            ms = t0.elapsed_time(t1) / trials
            name = f"queue{int(cfg.queue)}_streams{cfg.streams}_bw{cfg.block_width}"
            timings[name] = round(ms, 4)
            if ms < best_ms:
                best, best_ms = cfg, ms
            del mv
            torch.cuda.empty_cache()
    finally:
        scene.block_width = bw0
    return best, timings
