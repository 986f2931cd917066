"""Seeded synthetic Gaussian clouds and camera walks for the rasterizer's workloads.

One generator, numpy's PCG64 seeded with ``0x5EED0000 + config``, draws every
parameter in a fixed order, so a (config, count) pair always gives the same
cloud; tests/golden holds the first Gaussians of each config to pin it.

  plain_512      100k Gaussians, 512x512. Means U[-1,1]^3; log-scales
                 N(-4.5, 0.5) per axis; rotations normalise(N(0,1)^4);
                 opacity logits N(0,1) (opacity sigmoid(N(0,1))); SH degree
                 3, 16x3 coefficients N(0, 0.2) with +0.5 on the constant
                 term. Pinhole at radius 3 looking at the origin, 60 degree
                 horizontal field of view, near 0.01.
  plain_1080p    1.4M Gaussians at 1920x1080, the same, with log-scales
                 anisotropic: N(-4.5, 0.5) a Gaussian plus N(0, 1) an axis.
  walk_60        2.36M Gaussians, 60 views at 960x540 on a circular walk,
                 radius 4 + 2 sin(6 pi k / 60), fov 65 + 25 sin(2 pi k / 60 + 1).
  mixed_8        20M Gaussians, log-scales N(-6, 1), logits N(-1, 1), FP16
                 SH; 8 views of 3840x2160, 1920x1080, 1280x720, 640x360 (two
                 each) on the same walk.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .camera import Camera, orbit_camera

SH_COEFFS = 16
SH_CHUNK = 1 << 20


@dataclass
class Cloud:
    """Gaussians as numpy arrays in the layout of athx_gaussians."""

    means: np.ndarray  # [n, 3] f32
    log_scales: np.ndarray  # [n, 3] f32
    quats: np.ndarray  # [n, 4] f32 (x, y, z, w)
    opacity_logits: np.ndarray  # [n] f32
    sh: np.ndarray  # [n, 16, 3] f32 (or f16 for a half cloud)

    @property
    def count(self) -> int:
        return int(self.means.shape[0])

    def subset(self, n: int) -> "Cloud":
        return Cloud(self.means[:n], self.log_scales[:n], self.quats[:n], self.opacity_logits[:n], self.sh[:n])


def seed_for(config: int) -> int:
    return 0x5EED0000 + config


def make_cloud(
    n: int,
    seed: int,
    log_scale_mean: float = -4.5,
    log_scale_sd: float = 0.5,
    axis_spread: float = 0.0,
    logit_mean: float = 0.0,
    logit_sd: float = 1.0,
    sh_sd: float = 0.2,
    half_sh: bool = False,
) -> Cloud:
    rng = np.random.Generator(np.random.PCG64(seed))
    means = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    if axis_spread > 0.0:
        base = rng.normal(log_scale_mean, log_scale_sd, size=(n, 1))
        log_scales = (base + rng.normal(0.0, axis_spread, size=(n, 3))).astype(np.float32)
    else:
        log_scales = rng.normal(log_scale_mean, log_scale_sd, size=(n, 3)).astype(np.float32)
    q = rng.normal(0.0, 1.0, size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    quats = q.astype(np.float32)
    logits = rng.normal(logit_mean, logit_sd, size=n).astype(np.float32)
    # The harmonics in chunks of rows: the same draws in the same order as
    # one call, without an [n, 16, 3] float64 array (7.7 GB at 20M).
    sh = np.empty((n, SH_COEFFS, 3), np.float16 if half_sh else np.float32)
    for lo in range(0, n, SH_CHUNK):
        hi = min(lo + SH_CHUNK, n)
        block = rng.normal(0.0, sh_sd, size=(hi - lo, SH_COEFFS, 3))
        block[:, 0, :] += 0.5
        sh[lo:hi] = block
    return Cloud(means, log_scales, quats, logits, sh)


def plain_512_cloud(n: int = 100_000) -> Cloud:
    return make_cloud(n, seed_for(0))


def plain_1080p_cloud(n: int = 1_400_000) -> Cloud:
    return make_cloud(n, seed_for(1), axis_spread=1.0)


def walk_cloud(n: int = 2_360_000) -> Cloud:
    return make_cloud(n, seed_for(3))


def mixed_cloud(n: int = 20_000_000) -> Cloud:
    return make_cloud(n, seed_for(4), log_scale_mean=-6.0, log_scale_sd=1.0, logit_mean=-1.0, half_sh=True)


def front_camera(fov_degrees: float = 60.0, radius: float = 3.0, near: float = 0.01) -> Camera:
    cam = Camera.looking_at((0.0, 0.0, radius), (0.0, 0.0, 0.0)).with_horizontal_fov(fov_degrees)
    cam.lens.near_z = near
    cam.lens.far_z = 1.0e4
    return cam


def walk_cameras(views: int = 60, near: float = 0.01) -> list[Camera]:
    cams = []
    for k in range(views):
        radius = 4.0 + 2.0 * math.sin(6.0 * math.pi * k / 60.0)
        fov = 65.0 + 25.0 * math.sin(2.0 * math.pi * k / 60.0 + 1.0)
        cam = orbit_camera(radius, 2.0 * math.pi * k / 60.0, fov)
        cam.lens.near_z = near
        cam.lens.far_z = 1.0e4
        cams.append(cam)
    return cams


MIXED_SIZES = [(3840, 2160), (3840, 2160), (1920, 1080), (1920, 1080), (1280, 720), (1280, 720), (640, 360), (640, 360)]
