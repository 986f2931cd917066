"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Each check is one the reference's own tests make (file:line cited), a closed
form, an invariant, or an independent computation (numpy, scipy, brute
force), chosen so that a plausible slip in the restatement -- a dropped term,
a wrong sign or index, a transposed operand -- fails one of them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.special

from athenea_amd import camera, synthetic
from athenea_amd.camera import Camera
from oracle import pyoracle as po
from tests.splat_fixtures import CloudBuilder, counters_cloud, random_cloud


def default_camera_view(width, height, antialias=True, linear=True, tile=16):
    """The reference tests' camera: at the origin looking down -Z with
    focal = haperture, so fx = width (test_splat_render.cpp:130-140)."""
    cam = Camera()
    cam.lens.focal = cam.lens.haperture
    v = camera.view_for(cam, width, height)
    return v, po.orc_view(v, tile=tile, antialias=antialias, linear=linear)


# --- the deterministic exp and log -----------------------------------------------------------------

def _ulps(a: float, b: float) -> float:
    return abs(a - b) / float(np.spacing(np.float32(abs(b))))


def test_expf_within_two_ulp_of_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-87.0, 88.0, 4000), rng.uniform(-5, 5, 4000), [0.0, -0.0, 1.0, -1.0, 1e-8]])
    worst = 0.0
    for x in xs.astype(np.float32):
        want = math.exp(float(x))
        got = po.expf(float(x))
        if want < 1.1754944e-38:  # subnormal results: absolute check
            assert abs(got - want) < 2e-45 * 4
            continue
        worst = max(worst, _ulps(got, want))
    assert worst <= 2.0, worst
    assert po.expf(-100.0) == 0.0 and math.isinf(po.expf(100.0)) and math.isnan(po.expf(float("nan")))


def test_logf_within_two_ulp_of_libm():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 6000)), rng.uniform(0.5, 2.0, 4000), [1.0, 2.0, 0.5, 1.0001]])
    worst = 0.0
    for x in xs.astype(np.float32):
        want = math.log(float(x))
        got = po.logf(float(x))
        if want == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, _ulps(got, want))
    assert worst <= 2.0, worst
    assert math.isinf(po.logf(0.0)) and math.isnan(po.logf(-1.0))


# --- projection: closed forms and the reference's own assertions ------------------------------------

def test_single_splat_lands_where_the_camera_arithmetic_says():
    """test_splat_render.cpp:130-157: (1, 0.5, -10) lands on (192, 136) at 320x240."""
    v, ov = default_camera_view(320, 240, linear=False)
    cloud = CloudBuilder().add(1.0, 0.5, -10.0, 0.9, 0.02, 0.02, 0.02).cloud()
    frame = po.render([ov], cloud)
    alpha = frame["colour"][0][..., 3]
    y, x = np.unravel_index(np.argmax(alpha), alpha.shape)
    assert abs(int(x) - 192) <= 1 and abs(int(y) - 136) <= 1
    rec = frame["records"][0]
    assert rec[0] == pytest.approx(192.0, abs=1e-4) and rec[1] == pytest.approx(136.0, abs=1e-4)
    assert rec[2] == pytest.approx(10.0)


def test_isotropic_conic_closed_form():
    """A sphere of sigma s at depth z on the axis: 2D variance (f s / z)^2 + 0.3,
    so conic a = c = 1 / that, b = 0, and alpha = paid * o + (1 - paid) / 255."""
    v, ov = default_camera_view(256, 256)
    s, z, o = 0.05, 8.0, 0.7
    cloud = CloudBuilder().add(0.0, 0.0, -z, o, s, s, s).cloud()
    rec, touched, keys = po.project(ov, cloud)
    f = 256.0
    var = (f * s / z) ** 2
    a, b, c, alpha = rec[0, 4:8]
    assert a == pytest.approx(1.0 / (var + 0.3), rel=1e-5)
    assert c == pytest.approx(1.0 / (var + 0.3), rel=1e-5)
    assert abs(b) < 1e-9
    paid = var / (var + 0.3)  # sqrt(det_before / det) with det = (var + 0.3)^2
    o32 = 1.0 / (1.0 + math.exp(-math.log(o / (1 - o))))
    assert alpha == pytest.approx(paid * o32 + (1.0 - paid) / 255.0, rel=1e-5)
    assert rec[0, 3] == pytest.approx(min(2.0 * math.log(255.0 * alpha), 9.0), rel=1e-5)
    assert keys[0] == np.float32(z).view(np.uint32) >> 7


def test_rotation_turns_the_footprint():
    """A splat long along x, turned 90 degrees about z (w = cos 45, z = sin 45),
    is long along screen y: its conic swaps a and c."""
    v, ov = default_camera_view(256, 256)
    h = math.sqrt(0.5)
    cloud = CloudBuilder().add(0, 0, -5, 0.9, 0.2, 0.02, 0.02).add(0, 0, -5, 0.9, 0.2, 0.02, 0.02, (h, 0, 0, h)).cloud()
    rec, _, _ = po.project(ov, cloud)
    assert rec[0, 4] == pytest.approx(rec[1, 6], rel=1e-4)
    assert rec[0, 6] == pytest.approx(rec[1, 4], rel=1e-4)
    # Long along x: the conic (the inverse covariance) is small along x.
    assert rec[0, 6] > 10 * rec[0, 4]


def test_antialias_pays_back_dilation_energy():
    """test_splat_render.cpp:208-235: a sub-pixel splat with the compensation
    carries less than 0.6 of the energy it has without."""
    cloud = CloudBuilder().add(0.0, 0.0, -50.0, 0.9, 0.002, 0.002, 0.002).cloud()
    energies = []
    for aa in (True, False):
        _, ov = default_camera_view(64, 64, antialias=aa)
        energies.append(float(po.render([ov], cloud)["colour"][0][..., 3].sum()))
    assert energies[1] > 0.0 and energies[0] < 0.6 * energies[1]


def test_nearer_splat_wins_from_either_side():
    """test_splat_render.cpp:159-206."""
    b = CloudBuilder()
    b.add(0, 0, -1, 0.99, 0.5, 0.5, 0.05, colour=(1, 0, 0))  # red, z = -1
    b.add(0, 0, 1, 0.99, 0.5, 0.5, 0.05, colour=(0, 0, 1))  # blue, z = +1
    b.add(-3, 0, 0, 0.99, 0.3, 0.3, 0.3, colour=(0, 1, 0))  # green, -x
    cloud = b.cloud()

    def draw(eye):
        cam = Camera.looking_at(eye, (0, 0, 0))
        v = camera.view_for(cam, 200, 100)
        return po.render([po.orc_view(v, linear=False)], cloud)["colour"][0]

    front = draw((0, 0, 10))
    assert front[50, 100, 2] > 0.9 and front[50, 100, 0] < 0.1
    assert int(np.argmax(front[50, :, 1])) < 100
    back = draw((0, 0, -10))
    assert back[50, 100, 0] > 0.9 and back[50, 100, 2] < 0.1
    assert int(np.argmax(back[50, :, 1])) > 100


def test_cull_reasons_match_the_reference_counters():
    """test_splat_render.cpp:1831-1917: 28 slots, 16 kept, 4 behind the eye,
    3 too faint, 5 off screen."""
    _, ov = default_camera_view(320, 240)
    rec, touched, keys = po.project(ov, counters_cloud())
    culled = keys[touched == 0]
    reasons = culled & 0xFF
    assert touched.size == 28 and int((touched > 0).sum()) == 16
    assert ((culled & 0xFF000000) == 0xFF000000).all()
    assert int((reasons == 2).sum()) == 4 and int((reasons == 4).sum()) == 3 and int((reasons == 5).sum()) == 5
    assert int((touched[:22] > 0).sum()) == 10 and int((touched[22:] > 0).sum()) == 6


def _real_sh_3dgs(l: int, m: int, d: np.ndarray) -> float:
    """Real spherical harmonics without the Condon-Shortley phase, from
    scipy's complex ones: the convention 3DGS trains (and sh.slang evaluates)."""
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    ylm = scipy.special.sph_harm_y(l, abs(m), theta, phi)
    if m == 0:
        val = ylm.real
    elif m > 0:
        val = math.sqrt(2.0) * (-1) ** m * ylm.real
    else:
        val = math.sqrt(2.0) * (-1) ** m * ylm.imag
    return float((-1) ** abs(m) * val)


@pytest.mark.parametrize("k", range(1, 16))
def test_sh_basis_matches_textbook_real_harmonics(k):
    """Coefficient k alone moves the colour by Y_k(dir) * coefficient, with Y_k
    the textbook real harmonic l = floor(sqrt k), m = k - l^2 - l."""
    l = int(math.isqrt(k))
    m = k - l * l - l
    _, ov = default_camera_view(64, 64)
    rng = np.random.default_rng(k)
    for _ in range(5):
        p = rng.uniform(-0.3, 0.3, 3).astype(np.float32)
        p[2] -= 4.0
        b = CloudBuilder(sh_coeffs=16)
        coef = 0.1
        rest = np.zeros(45, np.float32)
        rest[0 * 15 + (k - 1)] = coef  # channel r, channel-major like a PLY
        b.add(p[0], p[1], p[2], 0.9, 0.05, 0.05, 0.05, colour=(0.5, 0.5, 0.5), rest=rest)
        rec, touched, _ = po.project(ov, b.cloud())
        assert touched[0] > 0
        d = p.astype(np.float64) - np.array([ov.eye[0], ov.eye[1], ov.eye[2]])
        d /= np.linalg.norm(d)
        want = 0.5 + coef * _real_sh_3dgs(l, m, d)
        assert rec[0, 8] == pytest.approx(want, abs=2e-6)
        assert rec[0, 9] == pytest.approx(0.5, abs=1e-6)


def test_srgb_decode_matches_the_curve():
    """color.slang:7-10: a capture's colour c is drawn as ((c + 0.055) / 1.055)^2.4."""
    _, ov = default_camera_view(64, 64, linear=False)
    b = CloudBuilder()
    b.add(0, 0, -4, 0.9, 0.05, 0.05, 0.05, colour=(0.8, 0.03, 0.4))
    rec, _, _ = po.project(ov, b.cloud())
    def curve(c):
        return c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4
    for ch, c in enumerate((0.8, 0.03, 0.4)):
        assert rec[0, 8 + ch] == pytest.approx(curve(c), rel=2e-6, abs=1e-7)


# --- scan, sort, ranges ------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 100003])
def test_scan_matches_cumsum(n):
    """test_algo.cpp:72-109's sizes."""
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n, dtype=np.uint32)
    out, total = po.exclusive_scan(x)
    want = np.concatenate([[0], np.cumsum(x.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(out, want) and total == int(x.astype(np.uint64).sum())


def sort_patterns(n: int, bits: int, seed: int):
    """The five patterns of test_algo.cpp:347-379 (sort_generate.slang)."""
    rng = np.random.default_rng(seed)
    mask = (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
    rand = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    yield "random", rand & np.uint64(mask)
    yield "equal", np.full(n, 12345 & mask, np.uint64)
    yield "sorted", (np.arange(n, dtype=np.uint64) * np.uint64(7)) & np.uint64(mask)
    yield "reverse", (np.arange(n, 0, -1, dtype=np.uint64) * np.uint64(7)) & np.uint64(mask)
    yield "distinct13", (rng.integers(0, 13, n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15 & mask)) & np.uint64(mask)


@pytest.mark.parametrize("n,bits", [(1, 32), (2, 32), (4097, 13), (65537, 32), (300007, 24), (250001, 64),
                                    (100000, 40)])
def test_sort_is_a_stable_sort(n, bits):
    for name, keys in sort_patterns(n, bits, n + bits):
        # Bits above key_bits are carried, not sorted on: set some.
        noisy = keys | (np.uint64(0xAB) << np.uint64(56)) if bits <= 48 else keys
        vals = np.arange(n, dtype=np.uint32)
        k, v = po.sort_pairs(noisy, vals, bits)
        mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64((1 << 64) - 1)
        order = np.argsort(noisy & mask, kind="stable")
        assert np.array_equal(v, order.astype(np.uint32)), name
        assert np.array_equal(k, noisy[order]), name


def test_ranges_by_brute_force():
    rng = np.random.default_rng(5)
    tiles = 37
    keys = np.sort(rng.integers(0, tiles - 3, 5000).astype(np.uint64) << np.uint64(24))
    frame_ranges = np.zeros(2 * tiles, np.uint32)
    po.load().orc_ranges(keys, keys.size, tiles, frame_ranges)
    for t in range(tiles):
        idx = np.nonzero((keys >> np.uint64(24)) == t)[0]
        want = (idx[0], idx[-1] + 1) if idx.size else (0, 0)
        assert tuple(frame_ranges[2 * t:2 * t + 2]) == want


# --- the frame --------------------------------------------------------------------------------------

def test_key_route_equals_the_reference_two_sort_route():
    """One stable sort of (tile << 24 | depth24) over pairs emitted in slot
    order gives exactly the reference's depth sort then stable tile sort
    (frame.slang:6-20; TileRasterizer.cpp:1111-1176)."""
    cloud = random_cloud(3000, 7)
    cam = Camera.looking_at((0.5, 0.3, 6.0), (0, 0, 0))
    v = camera.view_for(cam, 250, 190)
    for tile in (8, 16, 32):
        ov = po.orc_view(v, tile=tile, linear=False)
        ours = po.render([ov], cloud)
        ref = po.render_ref_order(ov, cloud)
        assert ours["pairs"] == ref["pairs"]
        assert np.array_equal(ours["pair_values"], ref["pair_values"])
        assert np.array_equal(ours["ranges"], ref["ranges"])
        assert np.array_equal(ours["colour"][0], ref["colour"])


def srgb_codes(img: np.ndarray) -> np.ndarray:
    c = np.clip(img[..., :3].astype(np.float64), 0.0, 1.0)
    enc = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055)
    return np.round(enc * 255.0)


def test_tiled_blend_matches_the_brute_force_reference():
    """test_splat_render.cpp:49-128 against reference_blend.slang:28-123's
    per-pixel brute force: p99 <= 2 and max <= 16 8-bit sRGB codes, no overflow."""
    cloud = random_cloud(3000, 11)
    cam = Camera.looking_at((0.0, 0.0, 7.0), (0, 0, 0))
    v = camera.view_for(cam, 250, 190)
    ov = po.orc_view(v, linear=False)
    frame = po.render([ov], cloud)
    truth, _, overflow = po.reference_brute_force(ov, frame["records"], frame["touched"])
    assert overflow == 0
    diff = np.abs(srgb_codes(frame["colour"][0]) - srgb_codes(truth))
    assert np.percentile(diff, 99) <= 2 and diff.max() <= 16


def test_walk_invariants_and_the_fp64_replay():
    """Premultiplied alpha over a transparent background is 1 - T; the FP64
    replay of the same decisions stays within 1e-4 of the FP32 image."""
    cloud = synthetic.plain_512_cloud(20000)
    v = camera.view_for(synthetic.front_camera(), 96, 80)
    ov = po.orc_view(v)
    frame = po.render([ov], cloud, with_f64=True)
    col, ft = frame["colour"][0], frame["final_t"][0]
    assert np.allclose(col[..., 3], 1.0 - ft, atol=1e-7)
    assert (ft >= 0).all() and (ft <= 1).all()
    assert np.abs(frame["colour64"][0] - col).max() < 1e-4
    assert frame["last"][0].max() > 0


def test_multi_view_frame_is_each_view_alone():
    cloud = synthetic.plain_512_cloud(5000)
    cams = synthetic.walk_cameras(60)[::20]
    sizes = [(96, 64), (64, 96), (80, 80)]
    views = [po.orc_view(camera.view_for(c, w, h)) for c, (w, h) in zip(cams, sizes)]
    together = po.render(views, cloud)
    for k, ov in enumerate(views):
        alone = po.render([ov], cloud)
        assert np.array_equal(together["colour"][k], alone["colour"][0])


def test_generator_is_pinned_by_the_golden_fixture():
    fixture = np.load(po.HERE.parent / "tests" / "golden" / "synthetic_head.npz")
    for name, make in (("plain_512", synthetic.plain_512_cloud), ("plain_1080p", synthetic.plain_1080p_cloud),
                       ("walk", synthetic.walk_cloud), ("mixed", synthetic.mixed_cloud)):
        c = make(16)
        for field in ("means", "log_scales", "quats", "opacity_logits", "sh"):
            assert np.array_equal(getattr(c, field), fixture[f"{name}_{field}"]), (name, field)


# --- the tile test against brute force ------------------------------------------------------------

def _box_min_f64(a, b, c, lox, loy, hix, hiy):
    """Least of q = a x^2 + 2 b x y + c y^2 over a box, in FP64: zero when the
    box holds the centre; else, q being convex, the least over the four
    edges, each a 1D quadratic minimised where its derivative vanishes."""
    if lox <= 0 <= hix and loy <= 0 <= hiy:
        return 0.0
    q = lambda x, y: a * x * x + 2 * b * x * y + c * y * y  # noqa: E731
    best = np.inf
    for x in (lox, hix):
        y = min(max(-b * x / c, loy), hiy) if c > 0 else loy
        best = min(best, q(x, y))
    for y in (loy, hiy):
        x = min(max(-b * y / a, lox), hix) if a > 0 else lox
        best = min(best, q(x, y))
    return best


def test_box_minimum_matches_dense_sampling():
    """The FP64 box minimum used below, against a dense grid over the box
    (the grid's least is an upper bound within one grid step of it)."""
    rng = np.random.default_rng(9)
    for _ in range(200):
        s1, s2 = np.exp(rng.uniform(-1, 4, 2))
        th = rng.uniform(0, np.pi)
        ct, st = np.cos(th), np.sin(th)
        cov = np.array([[ct * ct * s1 * s1 + st * st * s2 * s2, ct * st * (s1 * s1 - s2 * s2)],
                        [ct * st * (s1 * s1 - s2 * s2), st * st * s1 * s1 + ct * ct * s2 * s2]])
        a, b, c = np.linalg.inv(cov)[0, 0], np.linalg.inv(cov)[0, 1], np.linalg.inv(cov)[1, 1]
        lox, loy = rng.uniform(-60, 40, 2)
        hix, hiy = lox + 15, loy + 15
        xs, ys = np.meshgrid(np.linspace(lox, hix, 301), np.linspace(loy, hiy, 301))
        grid = (a * xs * xs + 2 * b * xs * ys + c * ys * ys).min()
        m = _box_min_f64(a, b, c, lox, loy, hix, hiy)
        assert m <= grid + 1e-12 * max(1.0, grid)
        assert grid - m <= 1e-3 * max(1.0, grid)


@pytest.mark.parametrize("tile", [8, 16])
def test_tile_membership_by_brute_force(tile):
    """The tiles each footprint is paired with (orc_project's counts and
    orc_emit_pairs' pairs), against brute force in FP64: every tile holding a
    pixel centre whose power is within the cut is paired, and every paired
    tile's box of pixel centres comes within the cut (frame.slang:354-379),
    both up to a relative 1e-5 for the FP32 rounding at the edge."""
    cloud = random_cloud(600, 21, min_size=0.01, max_size=0.5)
    cam = Camera.looking_at((0.3, -0.2, 5.0), (0, 0, 0))
    v = camera.view_for(cam, 200, 150)
    ov = po.orc_view(v, tile=tile)
    frame = po.render([ov], cloud)
    rec, touched = frame["records"], frame["touched"]
    tx_n, ty_n = po.tiles_of(ov)
    paired = [set() for _ in range(rec.shape[0])]
    ranges = frame["ranges"].reshape(-1, 2)
    for t in range(tx_n * ty_n):
        for p in range(ranges[t, 0], ranges[t, 1]):
            paired[frame["pair_values"][p]].add(t)
    px = np.arange(ov.width) + 0.5
    py = np.arange(ov.height) + 0.5
    checked = 0
    for s in np.nonzero(touched)[0]:
        cx, cy, _, cut, a, b, c = (float(x) for x in rec[s, :7])
        assert len(paired[s]) == touched[s]
        dx, dy = px - cx, py - cy
        power = a * dx[None, :] ** 2 + 2 * b * dx[None, :] * dy[:, None] + c * dy[:, None] ** 2
        inside = power <= cut * (1 - 1e-5)
        for t in range(tx_n * ty_n):
            tx, ty = t % tx_n, t // tx_n
            x0, y0 = tx * tile, ty * tile
            x1, y1 = min(x0 + tile, ov.width), min(y0 + tile, ov.height)
            if inside[y0:y1, x0:x1].any():
                assert t in paired[s], (s, t)
            if t in paired[s]:
                m = _box_min_f64(a, b, c, x0 + 0.5 - cx, y0 + 0.5 - cy, x1 - 0.5 - cx, y1 - 0.5 - cy)
                assert m <= cut * (1 + 1e-5) + 1e-6, (s, t, m, cut)
                checked += 1
    assert checked > 1000


# --- the image comparisons against numpy ----------------------------------------------------------

def _codes_f64(x):
    c = np.clip(x.astype(np.float64), 0.0, 1.0)
    e = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055) * 255.0
    return e


def test_compare_codes_against_numpy():
    """image_compare.slang's histogram of the per-pixel largest code
    difference (rgb through the sRGB curve, alpha linear), restated in numpy
    FP64; pixels whose code sits within 1e-3 of a rounding boundary are
    redrawn so FP32 and FP64 cannot round them apart."""
    rng = np.random.default_rng(4)
    h, w = 37, 53
    a = rng.uniform(-0.1, 1.1, (h, w, 4)).astype(np.float32)
    b = np.clip(a + rng.normal(0, 0.01, a.shape), -0.2, 1.2).astype(np.float32)
    for img in (a, b):
        for _ in range(20):
            ea = _codes_f64(img)
            ea[..., 3] = np.clip(img[..., 3].astype(np.float64), 0, 1) * 255.0
            near = np.abs(ea - np.floor(ea) - 0.5) < 1e-3
            if not near.any():
                break
            img[near] = rng.uniform(0, 1, int(near.sum())).astype(np.float32)
    ca, cb = _codes_f64(a), _codes_f64(b)
    worst = np.abs(np.rint(ca[..., :3]) - np.rint(cb[..., :3])).max(axis=-1)
    alpha = np.abs(np.rint(np.clip(a[..., 3], 0, 1).astype(np.float64) * 255) -
                   np.rint(np.clip(b[..., 3], 0, 1).astype(np.float64) * 255))
    worst = np.maximum(worst, alpha).astype(np.int64)
    want = np.bincount(worst.ravel(), minlength=256)
    assert np.array_equal(po.compare_codes(a, b), want)
    assert want[1:].sum() > 100  # the noise moves codes


def test_compare_hdr_against_numpy():
    """image_compare_hdr.slang: the per-pixel worst relative difference binned
    at floor((log2 + 16) * 8) + 1 (0 for none), and relMSE's sum
    (channel mean of (a - b)^2 / (b^2 + 1e-2)), against numpy FP64."""
    rng = np.random.default_rng(6)
    h, w = 29, 41
    b = np.exp(rng.normal(0, 2, (h, w, 4))).astype(np.float32)
    a = (b * np.exp(rng.normal(0, 0.3, b.shape))).astype(np.float32)
    a[0, :5] = b[0, :5]  # identical pixels land in bin 0
    for _ in range(20):
        d = np.abs(a.astype(np.float64) - b)
        worst = (d / np.maximum(np.abs(b.astype(np.float64)), 1e-3)).max(axis=-1)
        pos = (np.log2(np.where(worst > 0, worst, 1.0)) + 16) * 8
        near = (worst > 0) & (np.abs(pos - np.round(pos)) < 1e-3)
        if not near.any():
            break
        a[near] = (b[near] * np.float32(1.37)).astype(np.float32)
    d = np.abs(a.astype(np.float64) - b)
    worst = (d / np.maximum(np.abs(b.astype(np.float64)), 1e-3)).max(axis=-1)
    bins = np.where(worst > 0, np.clip(np.floor((np.log2(np.where(worst > 0, worst, 1.0)) + 16) * 8) + 1, 1, 255), 0)
    want = np.bincount(bins.astype(np.int64).ravel(), minlength=256)
    got, total = po.compare_hdr(a, b)
    assert np.array_equal(got, want)
    rel = (d * d / (b.astype(np.float64) ** 2 + 1e-2)).mean(axis=-1).sum()
    assert abs(total - rel) <= 1e-5 * rel
    assert want[0] == 5
