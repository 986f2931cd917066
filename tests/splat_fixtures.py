"""Synthetic splat clouds for the tests, restating the reference's fixtures.

``CloudBuilder`` and ``random_cloud`` follow tests/render/SplatFixtures.h:
20-133: records in the 3DGS encoding (logit opacity, log scale, quaternion
w x y z, DC colour as (colour - 0.5) / SH0, degree-1 rest coefficients
channel-major) and the same 64-bit LCG, drawn in the same order, so a seed
gives the reference's cloud. Here a record becomes the rasterizer's FP32
structure of arrays (quaternion stored x y z w).
"""
from __future__ import annotations

import numpy as np

from athenea_amd.synthetic import Cloud

SH0 = np.float32(0.28209479)
MASK64 = (1 << 64) - 1


class CloudBuilder:
    def __init__(self, sh_coeffs: int = 4):
        self.sh_coeffs = sh_coeffs
        self.rows = []

    def add(self, x, y, z, opacity, sx, sy, sz, wxyz=(1.0, 0.0, 0.0, 0.0), colour=(1.0, 1.0, 1.0), rest=None):
        f = np.float32
        logit = np.log(f(opacity) / (f(1.0) - f(opacity)))
        sh = np.zeros((self.sh_coeffs, 3), np.float32)
        for ch in range(3):
            sh[0, ch] = (f(colour[ch]) - f(0.5)) / SH0
        if rest is not None:
            rest = np.asarray(rest, np.float32)
            per = rest.size // 3
            for ch in range(3):
                for j in range(min(per, self.sh_coeffs - 1)):
                    sh[1 + j, ch] = rest[ch * per + j]
        w, qx, qy, qz = (f(a) for a in wxyz)
        self.rows.append(((f(x), f(y), f(z)), (np.log(f(sx)), np.log(f(sy)), np.log(f(sz))), (qx, qy, qz, w), logit, sh))
        return self

    def cloud(self) -> Cloud:
        n = len(self.rows)
        means = np.array([r[0] for r in self.rows], np.float32).reshape(n, 3)
        log_scales = np.array([r[1] for r in self.rows], np.float32).reshape(n, 3)
        quats = np.array([r[2] for r in self.rows], np.float32).reshape(n, 4)
        logits = np.array([r[3] for r in self.rows], np.float32).reshape(n)
        sh = np.stack([r[4] for r in self.rows]).astype(np.float32).reshape(n, self.sh_coeffs, 3)
        return Cloud(means, log_scales, quats, logits, sh)


class Lcg:
    """SplatFixtures.h:94-101."""

    def __init__(self, seed: int = 0):
        self.state = 0x853C49E6748FEA9B ^ seed

    def next(self) -> np.float32:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & MASK64
        return np.float32((self.state >> 40) & 0xFFFFFF) / np.float32(16777216.0)

    def range(self, lo, hi) -> np.float32:
        lo, hi = np.float32(lo), np.float32(hi)
        return lo + (hi - lo) * self.next()


def random_cloud(count: int, seed: int, min_size=0.004, max_size=0.6, min_opacity=0.02, max_opacity=0.99) -> Cloud:
    """SplatFixtures.h:103-133."""
    b = CloudBuilder()
    rng = Lcg(seed)
    f = np.float32
    for _ in range(count):
        size = np.exp(rng.range(np.log(f(min_size)), np.log(f(max_size))))
        stretch = rng.range(0.05, 1.0)
        rest = [rng.range(-0.3, 0.3) for _ in range(9)]
        x = rng.range(-2.0, 2.0)
        y = rng.range(-1.5, 1.5)
        z = rng.range(-2.0, 2.0)
        opacity = rng.range(min_opacity, max_opacity)
        thin = rng.range(0.1, 1.0)
        turn = [rng.range(-1, 1) for _ in range(4)]
        dc = [rng.range(0.0, 1.2) for _ in range(3)]
        b.add(x, y, z, opacity, size, size * stretch, size * thin, turn, dc, rest)
    return b.cloud()


def concat(*clouds: Cloud) -> Cloud:
    return Cloud(*(np.concatenate([getattr(c, k) for c in clouds]) for k in
                   ("means", "log_scales", "quats", "opacity_logits", "sh")))


def counters_cloud() -> Cloud:
    """The two clouds of test_splat_render.cpp:1831-1862, one after the other."""
    first = CloudBuilder()
    for k in range(10):  # kept
        first.add(-0.9 + 0.2 * k, 0.0, -10.0, 0.8, 0.05, 0.05, 0.05)
    for k in range(4):  # behind the eye
        first.add(0.1 * k, 0.0, 5.0, 0.8, 0.05, 0.05, 0.05)
    for k in range(3):  # too small to leave a mark
        first.add(0.1 * k, 0.3, -10.0, 0.8, 1.0e-5, 1.0e-5, 1.0e-5)
    for k in range(5):  # far to the side
        first.add(200.0 + k, 0.0, -10.0, 0.8, 0.05, 0.05, 0.05)
    second = CloudBuilder()
    for k in range(6):
        second.add(-0.5 + 0.2 * k, -0.4, -8.0, 0.8, 0.05, 0.05, 0.05)
    return concat(first.cloud(), second.cloud())
