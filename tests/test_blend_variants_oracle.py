"""CPU pins of the oracle's walk variants (orc_blend_ex; splat_blend.slang's
composite, matte and traced walks, splat_trace.slang), no GPU.

- The variants cost the colour nothing: without an under layer the image,
  depth, transmittance and walk lengths are orc_blend's, bit for bit.
- The under layer, closed forms for one splat: a layer behind it shows
  through what it leaves, C = w c + T U, depth (w z + T z_u) / 1; a layer in
  front of it hides it, C = U, T = 1, no pair read; a layer of depth 0 is no
  layer.
- The matte: at sampled pixels each id's coverage is the sum of the blend
  weights of its splats (an FP64 replay of the pixel's walk), ranks descend,
  and with no more ids than ranks the coverages add up to the pixel's alpha
  (splat_blend.slang:93-143's layer rule: the layer's own id takes T).
- The trace: the reference's test_splat_trace.cpp:25-118 assertions -- ranks
  name slots of the frame at their record's depth, descend, never twice,
  weights + rest + under = the pixel's alpha within 2e-3, some pixels hold
  more than four -- and the ranks are the four heaviest of the replayed walk.
"""
from __future__ import annotations

import numpy as np

from athenea_amd import camera, synthetic
from oracle import pyoracle as po
from tests.splat_fixtures import CloudBuilder


def frame_of(cloud, size=(120, 90), cam=None):
    cam = cam or synthetic.front_camera()
    view = camera.view_for(cam, *size)
    ov = po.orc_view(view)
    return ov, po.render([ov], cloud)


def halves(words):
    lo = (words & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64)
    hi = (words >> 16).astype(np.uint16).view(np.float16).astype(np.float64)
    return lo, hi


def replay(ov, frame, px, py):
    """The pixel's walk in FP64 from the frame's pairs: (slot, weight) of
    every splat it blends, in order, and its final transmittance."""
    rec = frame["records"].astype(np.float64)
    tiles_x = (ov.width + ov.tile - 1) // ov.tile
    t = (py // ov.tile) * tiles_x + px // ov.tile
    start, end = frame["ranges"][2 * t], frame["ranges"][2 * t + 1]
    T = 1.0
    out = []
    for slot in frame["pair_values"][start:end]:
        r = rec[slot]
        dx, dy = px + 0.5 - r[0], py + 0.5 - r[1]
        power = r[4] * dx * dx + 2 * r[5] * dx * dy + r[6] * dy * dy
        if power > r[3]:
            continue
        alpha = min(0.99, r[7] * np.exp(-0.5 * power))
        if alpha < 1 / 255:
            continue
        out.append((int(slot), T * alpha))
        T *= 1 - alpha
        if T < 1e-4:
            break
    return out, T


def test_variants_cost_the_colour_nothing():
    ov, fr = frame_of(synthetic.plain_512_cloud(8000))
    ex = po.blend_ex([ov], fr, matte=True, traced=True, slot_ids=np.arange(8000, dtype=np.uint32) % 9 + 1)
    for key in ("colour", "depth", "final_t", "last"):
        assert np.array_equal(ex[key][0], fr[key][0]), key


def one_splat(z_world=0.0):
    b = CloudBuilder()
    b.add(0.0, 0.0, z_world, 0.9, 0.2, 0.2, 0.2, colour=(0.8, 0.3, 0.1))
    return b.cloud()


def test_under_layer_behind_shows_through():
    ov, fr = frame_of(one_splat(), size=(64, 64))
    H, W = 64, 64
    U = np.array([0.2, 0.4, 0.6, 1.0], np.float32)
    colour = np.broadcast_to(U, (H, W, 4)).copy()
    depth = np.full((H, W), 5.0, np.float32)  # the splat stands at view z 3
    ex = po.blend_ex([ov], fr, under=[(colour, depth, None)])
    c, d, T = ex["colour"][0][32, 32].astype(np.float64), ex["depth"][0][32, 32], float(ex["final_t"][0][32, 32])
    walk, t64 = replay(ov, fr, 32, 32)
    assert len(walk) == 1 and abs(T - t64) < 1e-6 and T < 0.5
    rec = fr["records"][walk[0][0]].astype(np.float64)
    w = walk[0][1]
    assert np.allclose(c[:3], w * rec[8:11] + T * U[:3], atol=1e-6)
    assert abs(c[3] - ((1 - T) + T * 1.0)) < 1e-6
    assert abs(d - (w * rec[2] + T * 5.0)) < 1e-5
    # Far from the splat the layer is all there is.
    assert np.array_equal(ex["colour"][0][2, 2], U) and ex["depth"][0][2, 2] == np.float32(5.0)


def test_under_layer_in_front_hides_the_splat():
    ov, fr = frame_of(one_splat(), size=(64, 64))
    U = np.array([0.5, 0.5, 0.5, 1.0], np.float32)
    colour = np.broadcast_to(U, (64, 64, 4)).copy()
    depth = np.full((64, 64), 2.0, np.float32)
    ex = po.blend_ex([ov], fr, under=[(colour, depth, None)])
    assert np.array_equal(ex["colour"][0], colour)
    assert np.all(ex["final_t"][0] == 1.0) and np.all(ex["last"][0] == 0)
    assert np.all(ex["depth"][0] == np.float32(2.0))
    none = po.blend_ex([ov], fr, under=[(colour, np.zeros((64, 64), np.float32), None)])
    for key in ("colour", "depth", "final_t", "last"):
        assert np.array_equal(none[key][0], fr[key][0]), key


def test_matte_coverages_are_the_ids_weights():
    cloud = synthetic.plain_512_cloud(6000)
    ids = (np.arange(6000, dtype=np.uint32) % 4) + 11  # four ids: never more than six ranks
    ov, fr = frame_of(cloud)
    H, W = ov.height, ov.width
    U = np.array([0.1, 0.1, 0.1, 1.0], np.float32)
    under_ids = np.full((H, W), 77, np.uint32)
    depth = np.full((H, W), 50.0, np.float32)
    ex = po.blend_ex([ov], fr, under=[(np.broadcast_to(U, (H, W, 4)).copy(), depth, under_ids)], slot_ids=ids,
                     matte=True)
    cr = ex["crypto"][0]
    cov = np.concatenate([cr[..., 1], cr[..., 3]], axis=0)  # [6, H, W] in rank order
    got_ids = np.concatenate([cr[..., 0], cr[..., 2]], axis=0).view(np.uint32)
    alpha = ex["colour"][0][..., 3]
    assert np.abs(cov.sum(axis=0) - alpha).max() < 1e-5  # the layer's own id takes T
    order = np.concatenate([cr[:, :, :, 1:2], cr[:, :, :, 3:4]], axis=3).transpose(0, 3, 1, 2).reshape(6, H, W)
    assert np.all(np.diff(order, axis=0) <= 0)
    rng = np.random.default_rng(0)
    for _ in range(40):
        px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
        walk, T = replay(ov, fr, px, py)
        want = {77: T}
        for slot, w in walk:
            want[int(ids[slot])] = want.get(int(ids[slot]), 0.0) + w
        have = {int(i): float(c) for i, c in zip(got_ids[:, py, px], cov[:, py, px]) if c > 0}
        assert set(have) == {k for k, v in want.items() if v > 0}
        for k, v in have.items():
            assert abs(v - want[k]) < 1e-5


def test_trace_names_the_heaviest_and_loses_no_weight():
    cloud = synthetic.plain_512_cloud(20000)
    ov, fr = frame_of(cloud, size=(140, 100))
    H, W = ov.height, ov.width
    ex = po.blend_ex([ov], fr, traced=True)
    tr = ex["trace"][0]
    first = tr[0]
    empty = first == 0xFFFFFFFE
    slots = tr[:4].astype(np.int64)
    named = tr[:4] != 0xFFFFFFFF
    live = ~empty
    n = fr["records"].shape[0]
    rec = fr["records"]
    # every rank names a slot of the frame, at its record's depth
    for k in range(4):
        m = live & named[k]
        assert np.all(slots[k][m] < n)
        assert np.array_equal(tr[4 + k][m], rec[slots[k][m], 2].view(np.uint32))
    w01 = halves(tr[8])
    w23 = halves(tr[9])
    weights = np.stack([w01[0], w01[1], w23[0], w23[1]])
    rest, under = halves(tr[10])
    weights[~named] = 0.0
    # ranks descend, never twice
    assert np.all(np.diff(weights[:, live], axis=0) <= 0)
    for k in range(4):
        for m in range(k + 1, 4):
            both = live & named[k] & named[m]
            assert not np.any(slots[k][both] == slots[m][both])
    # weights + rest + under = alpha, to what halves hold
    alpha = ex["colour"][0][..., 3].astype(np.float64)
    total = weights.sum(axis=0) + rest + under
    assert np.abs(total - alpha)[live].max() < 2e-3
    assert np.all(alpha[empty] == 0.0)
    assert live.sum() > 1000 and np.count_nonzero(rest[live] > 0) > 0
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(200):
        px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
        walk, _ = replay(ov, fr, px, py)
        if len(walk) < 6:
            continue
        ws = sorted((w for _, w in walk), reverse=True)
        if ws[3] - ws[4] < 1e-6:
            continue  # a tie at the cut: either may hold the rank
        top = {s for s, w in walk if w >= ws[3]}
        assert set(slots[:, py, px].tolist()) == top
        checked += 1
    assert checked > 20
