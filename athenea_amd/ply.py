"""A trained splat cloud's PLY or .splat file, read into the packed device
layout the rasterizer draws (io::readSplatPly and io::readDotSplat,
modules/io/src/SplatReaders.cpp:48-240, then scene::CloudLoader::upload,
modules/scene/include/athenea/scene/GpuClouds.h:349-516).

The header is parsed by the library (athx_ply_parse: the reference's header
grammar and checks), the vertex block is copied to the device as the file
holds it, and the library validates, decodes and packs it there
(athx_ingest_ply): nothing of a vertex is read on the host.
"""
from __future__ import annotations

import ctypes
import mmap
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .raster import PackedGaussians, TileRasterizer

# The reference reads a PLY's header from the file's first 64 KiB
# (SplatReaders.cpp:21, kHeaderProbe).
HEADER_PROBE = 1 << 16
# Rest coefficients a colour kept at each harmonics degree.
BASIS_OF_DEGREE = (0, 3, 8, 15)


def parse_header(data: bytes) -> _lib.PlyLayout:
    """The vertex element's layout from the start of a PLY file (athx_ply_parse);
    raises AthxError with the reference's code and reason."""
    lib = _lib.load()
    layout = _lib.PlyLayout()
    message = ctypes.create_string_buffer(512)
    probe = bytes(data[:HEADER_PROBE])
    status = lib.athx_ply_parse(probe, len(probe), ctypes.byref(layout), message, len(message))
    if status != 0:
        raise _lib.AthxError(status, message.value.decode(errors="replace"))
    return layout


@dataclass
class PlyCloud:
    """What read_ply returns: the kept splats in the packed layout (file
    order), each one's vertex index in the file, and the file's vertex count."""

    splats: PackedGaussians
    origin: torch.Tensor  # int32 [kept]
    vertices: int


def kept_per_colour(layout: _lib.PlyLayout, max_degree: int) -> int:
    return min(int(layout.rest_per_colour), BASIS_OF_DEGREE[min(max(int(max_degree), 0), 3)])


def ingest(vertices: torch.Tensor, layout: _lib.PlyLayout, max_degree: int = 3, keep_faint: bool = False,
           raster: TileRasterizer | None = None) -> PlyCloud:
    """The vertex block (a uint8 device tensor of layout.count * layout.stride
    bytes) validated, decoded and packed on its device (athx_ingest_ply)."""
    if vertices.dtype != torch.uint8 or not vertices.is_cuda or not vertices.is_contiguous():
        raise ValueError("vertices must be a contiguous uint8 tensor on the device")
    n, stride = int(layout.count), int(layout.stride)
    if vertices.numel() < n * stride:
        raise ValueError(f"vertices hold {vertices.numel()} bytes; the layout needs {n * stride}")
    device = vertices.device
    keep = kept_per_colour(layout, max_degree)
    words = (keep * 3 + 1) // 2
    positions = torch.empty((n, 4), dtype=torch.float32, device=device)
    shape = torch.empty((n, 4), dtype=torch.int32, device=device)
    sh = torch.empty((n, words), dtype=torch.int32, device=device) if words else None
    origin = torch.empty((n,), dtype=torch.int32, device=device)
    out = _lib.IngestOut(positions.data_ptr(), shape.data_ptr(), sh.data_ptr() if sh is not None else None,
                         origin.data_ptr())
    own = raster is None
    raster = raster or TileRasterizer(device.index or 0)
    try:
        kept = ctypes.c_uint32(0)
        raster._check(raster._lib.athx_ingest_ply(raster._ctx, vertices.data_ptr(), ctypes.byref(layout),
                                                  int(max_degree), 1 if keep_faint else 0, ctypes.byref(out),
                                                  ctypes.byref(kept)))
    finally:
        if own:
            raster.close()
    k = int(kept.value)
    splats = PackedGaussians(positions[:k], shape[:k], sh[:k] if sh is not None else None, rest_per_colour=keep)
    return PlyCloud(splats, origin[:k], n)


def read_ply(path, device="cuda:0", max_degree: int = 3, keep_faint: bool = False,
             raster: TileRasterizer | None = None) -> PlyCloud:
    """A binary little-endian 3DGS PLY into the packed layout on `device`."""
    path = Path(path)
    with open(path, "rb") as fh:
        size = path.stat().st_size
        if size == 0:
            raise _lib.AthxError(3, f"'{path}': not a PLY file")
        with mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_READ) as data:
            try:
                layout = parse_header(data)
            except _lib.AthxError as e:
                raise _lib.AthxError(e.status, f"'{path}': {str(e).split(': ', 1)[-1]}") from None
            n, stride, start = int(layout.count), int(layout.stride), int(layout.data_offset)
            if start + n * stride > size:
                raise _lib.AthxError(5, f"'{path}': the file ends before its vertices do")
            block = np.frombuffer(data[start:start + n * stride], dtype=np.uint8)  # a copy: the map closes
    vertices = torch.from_numpy(block.copy()).to(device)
    return ingest(vertices, layout, max_degree, keep_faint, raster)


def read_dot_splat(path, device="cuda:0", keep_faint: bool = False,
                   raster: TileRasterizer | None = None) -> PlyCloud:
    """A .splat file (32-byte records: position and linear scale floats,
    rgba and rotation bytes) into the packed layout on `device`: its layout
    from the file's size (athx_splat_layout), the records copied to the
    device as they are and decoded there (athx_ingest_ply)."""
    path = Path(path)
    lib = _lib.load()
    layout = _lib.PlyLayout()
    message = ctypes.create_string_buffer(512)
    size = path.stat().st_size
    status = lib.athx_splat_layout(size, ctypes.byref(layout), message, len(message))
    if status != 0:
        raise _lib.AthxError(status, f"'{path}': {message.value.decode(errors='replace')}")
    block = np.fromfile(path, dtype=np.uint8)
    vertices = torch.from_numpy(block).to(device)
    return ingest(vertices, layout, 0, keep_faint, raster)


def read_splats(path, device="cuda:0", max_degree: int = 3, keep_faint: bool = False,
                raster: TileRasterizer | None = None) -> PlyCloud:
    """io::readSplats by extension (SplatReaders.cpp:28-46): .ply or .splat."""
    ext = Path(path).suffix.lower()
    if ext == ".splat":
        return read_dot_splat(path, device, keep_faint, raster)
    if ext == ".ply":
        return read_ply(path, device, max_degree, keep_faint, raster)
    if ext == ".spz":
        # SPZ is a zstd/gzip stream decoded by Niantic's spz library, which the
        # reference links when it finds it (SplatReaders.cpp:275-279).
        raise _lib.AthxError(4, f"'{path}': this build reads no .spz")
    if ext == ".sog" or Path(path).name == "meta.json":
        raise _lib.AthxError(4, f"'{path}': SOG (WebP planes and a palette) is not read by this build")
    raise _lib.AthxError(4, f"'{path}': no splat reader for '{ext}'")
