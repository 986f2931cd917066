"""athenea's packed device layout of a cloud, written from the FP32 arrays
(the encoder side of shaders/athenea/common/packing.slang:1-76 and
scene/splat_encoding.slang, in numpy): what a caller holding an athenea
`scene::GpuSplats` hands the rasterizer as `athx_packed_splats`.

    positions  float32 [n, 4]  x, y, z, opacity (linear)
    shape      uint32  [n, 4]  smallest-three quaternion; f16 log-scales and
                               the f16 base colour, two halves a word
    sh         uint32  [n, words]  f16 rest coefficients, rgb per basis,
                               words = (rest * 3 + 1) // 2

Halves round to nearest even (IEEE binary16, as f32tof16 does). The
quaternion is normalised, its largest component dropped (made non-negative)
and the other three stored as 10-bit unorms of [-sqrt(1/2), sqrt(1/2)],
rounded to nearest.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SH0 = 0.28209479177387814
SQRT_HALF = np.float32(0.70710678)


def halves(values: np.ndarray) -> np.ndarray:
    """IEEE binary16 bit patterns (uint32) of float32 values, round to nearest even."""
    return np.asarray(values, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)


def encode_quaternions(q: np.ndarray) -> np.ndarray:
    """packing.slang:33-57 over rows (x, y, z, w): one word each."""
    q = np.asarray(q, np.float32)
    q = q / np.linalg.norm(q, axis=1, keepdims=True).astype(np.float32)
    largest = np.argmax(np.abs(q), axis=1)
    sign = np.where(q[np.arange(q.shape[0]), largest] < 0, np.float32(-1), np.float32(1))
    q = q * sign[:, None]
    words = largest.astype(np.uint32)
    shift = np.full(q.shape[0], 22, np.uint32)
    for k in range(4):
        keep = largest != k
        unit = np.clip(q[:, k] / SQRT_HALF * np.float32(0.5) + np.float32(0.5), 0.0, 1.0)
        code = np.floor(unit * np.float32(1023.0) + np.float32(0.5)).astype(np.uint32)
        words = np.where(keep, words | (code << shift), words)
        shift = np.where(keep, shift - 10, shift)
    return words


def decode_quaternions(words: np.ndarray) -> np.ndarray:
    """packing.slang:59-76 in float32, the order the kernels compute it."""
    words = np.asarray(words, np.uint32)
    f = np.float32
    s = np.stack([((words >> 22) & 1023), ((words >> 12) & 1023), ((words >> 2) & 1023)], 1).astype(np.float32)
    s = (s / f(1023.0) * f(2.0) - f(1.0)) * SQRT_HALF
    dot = (s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2]
    rest = np.sqrt(np.maximum(f(0.0), f(1.0) - dot)).astype(np.float32)
    largest = words & 3
    q = np.zeros((words.size, 4), np.float32)
    for row in range(words.size):
        nxt = 0
        for k in range(4):
            if k == largest[row]:
                q[row, k] = rest[row]
            else:
                q[row, k] = s[row, nxt]
                nxt += 1
    return q


@dataclass
class PackedCloud:
    """A cloud in the packed layout, host memory."""

    positions: np.ndarray  # float32 [n, 4]
    shape: np.ndarray  # uint32 [n, 4]
    sh: np.ndarray | None  # uint32 [n, words], None at rest_per_colour 0
    rest_per_colour: int

    @property
    def count(self) -> int:
        return int(self.positions.shape[0])

    def to_device(self, device, linear: bool = True):
        import torch

        from .raster import PackedGaussians

        def t(a):
            return None if a is None else torch.from_numpy(np.ascontiguousarray(a).view(
                np.int32 if a.dtype == np.uint32 else a.dtype)).to(device)

        return PackedGaussians(t(self.positions), t(self.shape), t(self.sh), self.rest_per_colour, linear)


def pack_cloud(cloud) -> PackedCloud:
    """Packs an FP32 cloud (means, log_scales, quats (x, y, z, w), opacity
    logits, sh [n, k, 3]) as athenea's encoder does: opacity = sigmoid(logit),
    base colour = 0.5 + SH0 * dc (splat_encoding.slang:103-105), the rest
    coefficients as halves."""
    n = cloud.means.shape[0]
    f = np.float32
    positions = np.empty((n, 4), np.float32)
    positions[:, :3] = cloud.means
    positions[:, 3] = (f(1.0) / (f(1.0) + np.exp(-np.asarray(cloud.opacity_logits, np.float32)))).astype(np.float32)
    sh = np.asarray(cloud.sh, np.float32)
    base = (f(0.5) + f(SH0) * sh[:, 0, :]).astype(np.float32)
    ls = halves(cloud.log_scales)
    col = halves(base)
    shape = np.empty((n, 4), np.uint32)
    shape[:, 0] = encode_quaternions(cloud.quats)
    shape[:, 1] = ls[:, 0] | (ls[:, 1] << 16)
    shape[:, 2] = ls[:, 2] | (col[:, 0] << 16)
    shape[:, 3] = col[:, 1] | (col[:, 2] << 16)
    rest = sh.shape[1] - 1
    if rest == 0:
        return PackedCloud(positions, shape, None, 0)
    h = halves(sh[:, 1:, :].reshape(n, rest * 3))
    words = (rest * 3 + 1) // 2
    if h.shape[1] % 2:
        h = np.concatenate([h, np.zeros((n, 1), np.uint32)], 1)
    packed = (h[:, 0::2] | (h[:, 1::2] << 16)).astype(np.uint32)
    assert packed.shape[1] == words
    return PackedCloud(positions, shape, np.ascontiguousarray(packed), rest)


def pack_normals(normals: np.ndarray) -> np.ndarray:
    """packing.slang:116-127 (packNormal): the octahedral map, 16-bit unorms."""
    n = np.asarray(normals, np.float32)
    l1 = np.abs(n).sum(axis=1, keepdims=True)
    n = np.where(l1 > 1e-20, n / np.maximum(l1, 1e-30), np.array([0, 0, 1], np.float32))
    p = n[:, :2].copy()
    neg = n[:, 2] < 0
    wrap = (1.0 - np.abs(n[:, [1, 0]])) * np.where(n[:, :2] >= 0, 1.0, -1.0)
    p[neg] = wrap[neg]
    q = np.floor(np.clip(p * 0.5 + 0.5, 0, 1) * 65535.0 + 0.5).astype(np.uint32)
    return q[:, 0] | (q[:, 1] << 16)


def pack_pbr(metallic: np.ndarray, roughness: np.ndarray) -> np.ndarray:
    """splat_encoding.slang:189-202 (packPbr) with no transmission: metallic
    in byte 0, roughness in byte 1."""
    m = np.floor(np.clip(metallic, 0, 1) * 255.0 + 0.5).astype(np.uint32)
    r = np.floor(np.clip(roughness, 0, 1) * 255.0 + 0.5).astype(np.uint32)
    return m | (r << 8)


def unpacked_fields(packed: PackedCloud):
    """The FP32-array fields a packed cloud stands for (means, log-scales,
    the decoded quaternion, the opacity's logit, SH with the base colour as
    its constant term), for checkers that take the arrays."""
    from types import SimpleNamespace

    n = packed.count
    f64 = np.float64
    shape = packed.shape.astype(np.uint32)
    half = lambda w: (w & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32)
    log_scales = np.stack([half(shape[:, 1]), half(shape[:, 1] >> 16), half(shape[:, 2])], 1)
    base = np.stack([half(shape[:, 2] >> 16), half(shape[:, 3]), half(shape[:, 3] >> 16)], 1).astype(f64)
    o = packed.positions[:, 3].astype(f64)
    logits = np.log(o / (1.0 - o))
    rest = packed.rest_per_colour
    sh = np.zeros((n, rest + 1, 3), f64)
    sh[:, 0, :] = (base - 0.5) / SH0
    if rest:
        w = packed.sh.astype(np.uint32)
        h = np.stack([w & 0xFFFF, w >> 16], 2).reshape(n, -1)[:, :rest * 3]
        sh[:, 1:, :] = h.astype(np.uint16).view(np.float16).astype(f64).reshape(n, rest, 3)
    return SimpleNamespace(means=packed.positions[:, :3].astype(f64), log_scales=log_scales.astype(f64),
                           quats=decode_quaternions(shape[:, 0]).astype(f64), opacity_logits=logits, sh=sh)
