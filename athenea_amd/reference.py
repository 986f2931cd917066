"""The ground truth on the device and the reference's image comparisons
(render::ReferenceRenderer and compareImages / compareHdr,
modules/render/include/athenea/render/ReferenceRenderer.h:20-66), through the
C ABI: athx_reference_render, athx_compare_images, athx_compare_hdr.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from .raster import (PackedGaussians, RenderSettings, RenderTargets, TileRasterizer, _gaussians_c, _ptr,
                     _views_and_targets)


@dataclass
class ImageDifference:
    pixels: int
    p99: int
    max: int
    over2: int


@dataclass
class HdrDifference:
    pixels: int
    rel_mse: float
    p99_relative: float
    max_relative: float


def reference_render(raster: TileRasterizer, view, cloud: PackedGaussians, settings: RenderSettings | None = None,
                     want_records: bool = False) -> tuple[RenderTargets, torch.Tensor | None, int]:
    """One view of a packed cloud, every splat at every pixel in (depth,
    index) order: (targets with colour and depth, the projected records
    [count, 12] when asked for, the pixels reached by more than 1024 splats)."""
    settings = settings or RenderSettings()
    dev = cloud.positions.device
    targets = RenderTargets.allocate(view.width, view.height, dev, False)
    c_views, c_targets = _views_and_targets([view], [targets])
    g = _gaussians_c(cloud)
    s = settings.to_c(cloud.linear)
    records = torch.empty((max(cloud.count, 1), 12), dtype=torch.float32, device=dev) if want_records else None
    over = ctypes.c_uint32(0)
    raster._check(raster._lib.athx_reference_render(raster._ctx, c_views, ctypes.byref(g), ctypes.byref(s),
                                                    c_targets, _ptr(records), ctypes.byref(over)))
    return targets, (records[:cloud.count] if records is not None else None), int(over.value)


def reference_render_instances(raster: TileRasterizer, view, instances: list, settings: RenderSettings | None = None,
                               want_records: bool = False) -> tuple[RenderTargets, torch.Tensor | None, int]:
    """The ground truth of several packed clouds, each with its own
    object-to-world (raster.SplatInstance), under one view made by
    camera.view_for: every drawn splat of all of them at every pixel in
    (depth, index) order, indices the instances' back to back."""
    from . import camera as _camera

    settings = settings or RenderSettings()
    if not instances or any(not isinstance(i.cloud, PackedGaussians) for i in instances):
        raise ValueError("the ground truth reads packed clouds")
    dev = instances[0].cloud.positions.device
    targets = RenderTargets.allocate(view.width, view.height, dev, False)
    c_views, c_targets = _views_and_targets([view], [targets])
    total = sum(i.cloud.count for i in instances)
    keep = []
    c_inst = (_lib.Instance * len(instances))()
    for j, inst in enumerate(instances):
        rows, eye = _camera.instance_transform(view, inst.object_to_world)
        g = _gaussians_c(inst.cloud)
        keep += [rows, eye, g]
        c_inst[j].gaussians = ctypes.pointer(g)
        c_inst[j].object_to_view = rows.ctypes.data
        c_inst[j].eye = eye.ctypes.data
    s = settings.to_c(instances[0].cloud.linear)
    records = torch.empty((max(total, 1), 12), dtype=torch.float32, device=dev) if want_records else None
    over = ctypes.c_uint32(0)
    raster._check(raster._lib.athx_reference_render_instances(raster._ctx, c_views, c_inst, len(instances),
                                                              ctypes.byref(s), c_targets, _ptr(records),
                                                              ctypes.byref(over)))
    del keep
    return targets, (records[:total] if records is not None else None), int(over.value)


def _images(a: torch.Tensor, b: torch.Tensor) -> tuple[int, int]:
    if a.shape != b.shape or a.dim() != 3 or a.shape[2] != 4 or a.dtype != torch.float32 or b.dtype != torch.float32:
        raise ValueError("two float32 [height, width, 4] images of one size")
    if not (a.is_cuda and b.is_cuda and a.is_contiguous() and b.is_contiguous()):
        raise ValueError("contiguous device images")
    return int(a.shape[1]), int(a.shape[0])


def compare_images(raster: TileRasterizer, a: torch.Tensor, b: torch.Tensor) -> ImageDifference:
    """The per-pixel worst channel difference in 8-bit sRGB codes, alpha included."""
    w, h = _images(a, b)
    d = _lib.ImageDifference()
    raster._check(raster._lib.athx_compare_images(raster._ctx, a.data_ptr(), b.data_ptr(), w, h, ctypes.byref(d)))
    return ImageDifference(int(d.pixels), int(d.p99), int(d.max), int(d.over2))


def compare_hdr(raster: TileRasterizer, a: torch.Tensor, b: torch.Tensor) -> HdrDifference:
    """relMSE and the relative difference's percentiles, as radiance."""
    w, h = _images(a, b)
    d = _lib.HdrDifference()
    raster._check(raster._lib.athx_compare_hdr(raster._ctx, a.data_ptr(), b.data_ptr(), w, h, ctypes.byref(d)))
    return HdrDifference(int(d.pixels), float(d.rel_mse), float(d.p99_relative), float(d.max_relative))
