"""The binning's shortcut for the tile test, checked against the float test.

A strip entry decides most of its tiles without the float test: a tile is a
sure hit when its box meets the x-extent of the footprint's ellipse at
cutoff - kSlack M, sliced by the tile row's band, and a sure miss when it does
not meet the extent at cutoff + kSlack M (M bounds the test's terms over the
box; kSlack = 256 * 2^-24, covering the test's rounding and the extent's own,
computed in float32 from a determinant taken in FP64). The float test here is min_power_over_box in
float32 in the kernels' operation order (the order the oracle and the GPU
parity tests pin bit for bit). Over random footprints -- thin, turned, huge,
faint, off screen -- and tiles near their edges, a sure hit is never a float
miss and a sure miss never a float hit; with no slack the rule does fail,
so the check can see a wrong margin.
"""
from __future__ import annotations

import numpy as np
import pytest

f32 = np.float32
EPS = 2.0 ** -24


def ulp_jitter(r, rng):
    """r moved by -1, 0 or +1 unit in the last place, element by element: the
    kernels take these reciprocals with the hardware's (within an ulp of the
    correctly rounded one), and the rule must hold for any of them."""
    k = rng.integers(-1, 2, size=r.shape)
    down = np.nextafter(r, f32(-np.inf)).astype(f32)
    up = np.nextafter(r, f32(np.inf)).astype(f32)
    return np.where(k < 0, down, np.where(k > 0, up, r)).astype(f32)


def recip(v, rng):
    r = np.where(v > f32(1e-12), f32(1.0) / np.where(v > f32(1e-12), v, f32(1)), f32(0.0)).astype(f32)
    return np.where(v > f32(1e-12), ulp_jitter(r, rng), f32(0.0)).astype(f32)


def min_power(a, b, c, lox, loy, hix, hiy):
    """frame.slang:354-371 in float32: each edge's minimiser ((-b) x) / c."""
    inside = (lox <= 0) & (hix >= 0) & (loy <= 0) & (hiy >= 0)
    best = np.full(a.shape, f32(3.4e38), f32)
    safe_a = np.where(a > f32(1e-12), a, f32(1.0)).astype(f32)
    safe_c = np.where(c > f32(1e-12), c, f32(1.0)).astype(f32)
    for x in (lox, hix):
        y = np.where(c > f32(1e-12), np.minimum(np.maximum(((-b) * x) / safe_c, loy), hiy), loy).astype(f32)
        best = np.minimum(best, (a * x * x + f32(2.0) * b * x * y + c * y * y).astype(f32))
    for y in (loy, hiy):
        x = np.where(a > f32(1e-12), np.minimum(np.maximum(((-b) * y) / safe_a, lox), hix), lox).astype(f32)
        best = np.minimum(best, (a * x * x + f32(2.0) * b * x * y + c * y * y).astype(f32))
    return np.where(inside, f32(0.0), best)


def band_extent(a, b, c, ia, ic, det, inv_det, C, y0, y1):
    """project.hip band_extent, float32 in the kernel's order: the x-extent
    of {y0 <= y <= y1, q <= C} (det from FP64, so a thin ellipse's does not
    cancel)."""
    ok = C > f32(0)
    sC = np.sqrt(np.maximum(C, f32(0))).astype(f32)
    ym = (sC * np.sqrt(a * inv_det).astype(f32)).astype(f32)
    xg = (sC * np.sqrt(c * inv_det).astype(f32)).astype(f32)
    ya, yb = np.maximum(y0, -ym), np.minimum(y1, ym)
    ok &= ya <= yb
    yr = np.minimum(np.maximum((-b * xg) * ic, ya), yb).astype(f32)
    yl = np.minimum(np.maximum((b * xg) * ic, ya), yb).astype(f32)
    aC = (a * C).astype(f32)
    xb = ((-b * yr + np.sqrt(np.maximum(aC - det * yr * yr, f32(0))).astype(f32)) * ia).astype(f32)
    xa = ((-b * yl - np.sqrt(np.maximum(aC - det * yl * yl, f32(0))).astype(f32)) * ia).astype(f32)
    return ok, xa, xb


def classify(seed, k_slack, n=300000, W=1920, H=1080, T=16):
    rng = np.random.default_rng(seed)
    sx = np.exp(rng.uniform(np.log(0.5), np.log(400), n))
    sy = sx / np.exp(rng.uniform(0, np.log(3000), n))
    th = rng.uniform(0, np.pi, n)
    ct, st = np.cos(th), np.sin(th)
    ca = ct**2 * sx**2 + st**2 * sy**2 + 0.3
    cc = st**2 * sx**2 + ct**2 * sy**2 + 0.3
    cb = ct * st * (sx**2 - sy**2)
    D = ca * cc - cb**2
    a, b, c = f32(cc / D), f32(-cb / D), f32(ca / D)
    alpha = np.exp(rng.uniform(np.log(1 / 255), 0, n))
    cut = np.minimum(2 * np.log(np.maximum(255 * alpha, 1.0001)), 9).astype(f32)
    cx = rng.uniform(-200, W + 200, n).astype(f32)
    cy = rng.uniform(-200, H + 200, n).astype(f32)
    det64 = a.astype(np.float64) * c - b.astype(np.float64) ** 2
    xg = np.sqrt(np.maximum(c * cut / det64, 0))
    yg = np.sqrt(np.maximum(a * cut / det64, 0))
    # Tiles around the ellipse's edges, where the decisions are close.
    ty = np.clip(((cy + rng.uniform(-1.3, 1.3, n) * yg) / T).astype(np.int64), 0, (H + T - 1) // T - 1)
    tx = np.clip(((cx + rng.uniform(-1.3, 1.3, n) * xg) / T).astype(np.int64), 0, (W + T - 1) // T - 1)
    ey, ex = np.minimum((ty + 1) * T, H), np.minimum((tx + 1) * T, W)
    loy = ((ty * T).astype(f32) + f32(0.5)) - cy
    hiy = ((ey - 1).astype(f32) + f32(0.5)) - cy
    lox = ((tx * T).astype(f32) + f32(0.5)) - cx
    hix = ((ex - 1).astype(f32) + f32(0.5)) - cx
    hit = min_power(a, b, c, lox, loy, hix, hiy) <= cut
    # The kernel's rule, in float32 but for the determinant.
    det = (a.astype(np.float64) * c.astype(np.float64) - b.astype(np.float64) ** 2)
    usable = (a > f32(1e-9)) & (c > f32(1e-9)) & (det > 0)
    detf = det.astype(f32)
    inv_det = np.where(usable, ulp_jitter((f32(1.0) / np.where(usable, detf, f32(1))).astype(f32), rng),
                       f32(0)).astype(f32)
    X = np.maximum(np.abs(lox), np.abs(hix))
    Y = np.maximum(np.abs(loy), np.abs(hiy))
    M = (a * X * X + f32(2.0) * np.abs(b) * X * Y + c * Y * Y).astype(f32)
    slack = (f32(k_slack * EPS) * M).astype(f32)
    ia, ic = recip(a, rng), recip(c, rng)
    ok_lo, xa_lo, xb_lo = band_extent(a, b, c, ia, ic, detf, inv_det, (cut - slack).astype(f32), loy, hiy)
    ok_hi, xa_hi, xb_hi = band_extent(a, b, c, ia, ic, detf, inv_det, (cut + slack).astype(f32), loy, hiy)
    sure = usable & ok_lo & (hix >= xa_lo) & (lox <= xb_lo)
    maybe = ~usable | (ok_hi & (hix >= xa_hi) & (lox <= xb_hi))
    return hit, sure, maybe


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sure_decisions_agree_with_the_float_test(seed):
    hit, sure, maybe = classify(seed, 256.0)
    assert not np.any(sure & ~hit), "a sure hit the float test misses"
    assert not np.any(~maybe & hit), "a sure miss the float test touches"
    # Near the edges about 4 in 10 tiles are touched; nearly all are decided.
    assert 0.2 < hit.mean() < 0.7
    assert np.mean(maybe & ~sure) < 5e-3


def test_without_slack_the_rule_fails_somewhere():
    fails = 0
    for seed in range(3, 9):
        hit, sure, maybe = classify(seed, 0.0)
        fails += int(np.sum(sure & ~hit) + np.sum(~maybe & hit))
    assert fails > 0

