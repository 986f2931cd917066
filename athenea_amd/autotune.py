"""Autotuner for the two launch shapes the reference fixes: the tile side
(16 in the reference, TileRasterizer.cpp:19; frame.slang:34) and the blend's
LDS batch depth (256, splat_blend.slang:150-155).

It times the frame for every (tile, batch) candidate on the caller's own
workload -- the views, the cloud -- and returns the fastest as a
RenderSettings. Results are kept per workload shape (Gaussian count, view
sizes, SH width) for the process's life, so a render loop pays the search
once.

Every candidate draws the same image: a Gaussian that reaches a pixel reaches
its tile by the exact tile test whatever the tile side, and each tile lists
its Gaussians in (depth key, slot) order, so every pixel walks the same
Gaussians in the same order (the GPU parity tests check each shape against
the oracle, and tests/test_gpu_autotune.py the shapes against each other).
"""
from __future__ import annotations

import dataclasses
import time

import torch

from ._lib import AthxError
from .raster import GaussianCloud, RenderSettings, TileRasterizer

TILES = (8, 16, 32)
BATCHES = (32, 64, 128, 256)

_cache: dict[tuple, tuple[RenderSettings, list[dict]]] = {}


def _shape_key(views: list, cloud: GaussianCloud, base: RenderSettings) -> tuple:
    return (cloud.count, int(cloud.sh.shape[1]), str(cloud.sh.dtype),
            tuple((v.width, v.height) for v in views), base.antialias, base.max_sh_degree)


def autotune(raster: TileRasterizer, views: list, cloud: GaussianCloud, base: RenderSettings | None = None,
             tiles=TILES, batches=BATCHES, reps: int = 3, use_cache: bool = True
             ) -> tuple[RenderSettings, list[dict]]:
    """Returns (the fastest settings, every candidate's median ms). A tile
    side whose frame the rasterizer refuses as too large (more pairs than a
    frame holds: small tiles multiply them) is listed with its reason and
    skipped."""
    base = base or RenderSettings()
    key = _shape_key(views, cloud, base)
    if use_cache and key in _cache:
        return _cache[key]
    table = []
    best = None
    for tile in tiles:
        targets = None
        for batch in batches:
            s = dataclasses.replace(base, tile=tile, batch=batch, time_stages=False, count_splats=False)
            try:
                targets, _ = raster.render(views, cloud, s, targets)  # warm-up, and sizes the scratch
            except AthxError as e:
                if e.code != "Unsupported":
                    raise
                table.append({"tile": tile, "batch": batch, "ms": None, "refused": str(e)})
                break
            torch.cuda.synchronize()
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                targets, _ = raster.render(views, cloud, s, targets)
                torch.cuda.synchronize()
                times.append((time.perf_counter() - t0) * 1e3)
            ms = sorted(times)[len(times) // 2]
            table.append({"tile": tile, "batch": batch, "ms": round(ms, 4)})
            if best is None or ms < best[0]:
                best = (ms, s)
    result = (best[1], table)
    if use_cache:
        _cache[key] = result
    return result
