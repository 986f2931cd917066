"""Pins of the oracle's relit colour (splat_relight.slang's dome term with the
plain lobes, light/environment.slang's prepared-sky reader) against what
the mathematics fixes: exhaustive binary16 decoding, the level layout's
closed form, a constant sky (irradiance pi L, radiance L at any roughness),
the GGX fit against a quadrature of the height-correlated Smith GGX
directional albedo, Fresnel's end points, a sky linear in texel position
read back through an independent octahedral map, and the shading normal's
transport to the world under a non-uniform scale."""
from __future__ import annotations

import numpy as np
import pytest

from athenea_amd import camera, sky as skymod
from oracle import pyoracle as po
from splat_fixtures import CloudBuilder


@pytest.fixture(scope="module", autouse=True)
def _lib():
    po.load()


def _attrs(sky, n=1, metallic=0.0, roughness=0.5, albedo=None, normals=None, o2w=skymod.IDENTITY_3X4):
    return skymod.RelightAttributes(np.full(n, metallic, np.float32), np.full(n, roughness, np.float32), sky,
                                    normals, albedo, o2w)


def test_binary16_decoding_is_exact_for_every_half():
    halves = np.arange(1 << 16, dtype=np.uint16)
    want = halves.view(np.float16).astype(np.float32)
    got = np.array([po.half_to_float(int(h)) for h in halves], np.float32)
    finite = np.isfinite(want)
    assert np.array_equal(got[finite].view(np.uint32), want[finite].view(np.uint32))
    assert np.all(np.isnan(got[~finite & np.isnan(want)]))
    assert np.array_equal(got[np.isinf(want)], want[np.isinf(want)])


@pytest.mark.parametrize("base", [256, 512, 1024, 4096])
def test_level_offsets_follow_the_closed_form(base):
    """environment.slang:77-87: levels above the floor are B, B/2, ..., their
    squares summing to (4^(b+1) - 4^(b-m+1)) / 3, the rest the floor's."""
    b = base.bit_length() - 1
    f = 4
    for level in range(9):
        m = min(level, b - f)
        closed = ((1 << (2 * (b + 1))) - (1 << (2 * (b - m + 1)))) // 3 + (level - m) * 16 * 16
        assert po.env_level_offset(level, base) == closed == skymod.level_offset(level, base)
    if base == 256:
        # 256^2 + 128^2 + 64^2 + 32^2 above the floor, four levels of 16^2.
        assert skymod.texels_per_dome(256) == 87040 + 4 * 256


def test_constant_sky_radiance_and_irradiance():
    L = (1.5, 0.25, 3.0)  # binary16-exact
    sky = skymod.constant_sky(L)
    rl = _attrs(sky)
    rng = np.random.default_rng(3)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = float(rng.uniform(0, 1))
        assert np.array_equal(po.env_radiance(rl, d, r), np.float32(L))
        e = po.env_irradiance(rl, d)
        np.testing.assert_allclose(e, np.pi * np.asarray(L), rtol=2e-6)


def test_ggx_fit_against_quadrature():
    """The rational fit (microfacet.slang:66-82) is within a few per cent of
    the split-sum terms of a height-correlated Smith GGX, integrated here by
    importance sampling: a coefficient in the wrong place is off by far more."""
    rng = np.random.default_rng(1)

    def dfg(x, alpha, count=400_000):
        v = np.array([np.sqrt(1 - x * x), 0.0, x])
        u1, u2 = rng.random(count), rng.random(count)
        a2 = alpha * alpha
        ct = np.sqrt((1 - u1) / (1 + (a2 - 1) * u1))
        st = np.sqrt(1 - ct * ct)
        ph = 2 * np.pi * u2
        h = np.stack([st * np.cos(ph), st * np.sin(ph), ct], 1)
        vh = h @ v
        nl = (2 * vh[:, None] * h - v)[:, 2]

        def lam(c):
            c = np.clip(c, 1e-9, 1)
            return (-1 + np.sqrt(1 + a2 * (1 - c * c) / (c * c))) / 2

        g = 1 / (1 + lam(x) + lam(nl))
        w = np.where((nl > 0) & (vh > 0), g * vh / (x * h[:, 2]), 0.0)
        s = (1 - vh) ** 5
        return np.mean(w * (1 - s)), np.mean(w * s)

    for x in (0.2, 0.5, 0.9):
        for alpha in (0.1, 0.4, 0.8):
            a_want, b_want = dfg(x, alpha)
            a, b = po.ggx_albedo(x, alpha)
            assert abs(a - a_want) < 0.035, (x, alpha, a, a_want)
            assert abs(b - b_want) < 0.005, (x, alpha, b, b_want)


def test_fresnel_end_points():
    # A dielectric at 1.5: ((n - 1) / (n + 1))^2 = 0.04 head on, 1 at grazing.
    assert abs(po.fresnel_dielectric(1.0, 1.5) - 0.04) < 1e-7
    assert abs(po.fresnel_dielectric(0.0, 1.5) - 1.0) < 1e-6
    # The artistic index reproduces its reflectivity head on and goes to one at grazing.
    for r in (0.05, 0.3, 0.7, 0.95):
        assert abs(po.fresnel_conductor(r, 1.0) - r) < 2e-6, r
        assert po.fresnel_conductor(r, 1e-4) > 0.99


def _oct_encode(d):
    d = np.asarray(d, np.float64)
    n = d / np.abs(d).sum()
    if n[2] >= 0:
        p = n[:2]
    else:
        p = (1 - np.abs(n[[1, 0]])) * np.where(n[:2] >= 0, 1.0, -1.0)
    return p * 0.5 + 0.5


def test_sky_linear_in_texel_position_reads_back_through_the_map():
    """Texel (x, y) of level k holds (x + y, x, k): bilinear reads of a
    linear function are linear, so the radiance along d at roughness
    (k/7)^2 is (ax + ay, ax, k) at a = clamp(octEncode(d) side - 0.5),
    octEncode restated independently here."""
    base = 256
    words = []
    for k in range(skymod.LEVELS):
        side = skymod.level_side(k, base)
        y, x = np.mgrid[0:side, 0:side]
        rgb = np.stack([x + y, x, np.full_like(x, k)], -1).reshape(-1, 3).astype(np.float64)
        words.append(skymod._pack_half(rgb))
    sky = skymod.PreparedSky(np.concatenate(words).reshape(-1).astype(np.uint32), np.zeros((16, 4), np.float32),
                             np.zeros((2, 4), np.float32), base)
    rl = _attrs(sky)
    rng = np.random.default_rng(5)
    for _ in range(200):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        k = int(rng.integers(0, skymod.LEVELS))
        side = skymod.level_side(k, base)
        at = np.clip(_oct_encode(d) * side - 0.5, 0, side - 1)
        got = po.env_radiance(rl, d, skymod.level_roughness(k))
        np.testing.assert_allclose(got, [at[0] + at[1], at[0], k], atol=2e-3 * side / 16 + 1e-3)


def test_relit_point_under_a_constant_sky_with_a_sun():
    """Constant sky L and a sun of irradiance E along the normal: the body
    reads albedo (1 - m) (L + E / pi), the polish reflects L."""
    L, E = np.array([1.0, 0.5, 2.0]), np.array([3.0, 2.0, 1.0])
    sky = skymod.constant_sky(tuple(L))
    sky.sun[0] = [0.0, 0.0, 1.0, 0.01]
    sky.sun[1, :3] = E
    n = np.array([0.0, 0.0, 1.0])
    wo = np.array([0.6, 0.0, 0.8])
    albedo = np.array([0.2, 0.5, 0.9])
    for metallic in (0.0, 1.0):
        for rough in (0.1, 0.7):
            got = po.relit_point(_attrs(sky), n, wo, albedo, metallic, rough)
            nv = 0.8
            diel = po.ggx_dielectric(rough, nv)
            polish = np.array([diel + (po.ggx_conductor(a, rough, nv) - diel) * metallic for a in albedo])
            want = albedo * (1 - metallic) * (L + E / np.pi) + polish * L
            np.testing.assert_allclose(got, want, rtol=3e-6, atol=1e-7)


def test_relit_projection_normal_and_eye_go_to_the_world():
    """One Gaussian relit through the projection: its stored normal, flipped
    to the disc's side, taken to the world by the inverse transpose of a
    non-uniformly scaled rotation, and the eye through the same rows -- the
    colour must be relit_point's at that normal and direction."""
    b = CloudBuilder(sh_coeffs=1)
    b.add(0.1, -0.05, 0.0, 0.9, 0.05, 0.04, 0.002)
    cloud = b.cloud()
    cam = camera.orbit_camera(3.0, 0.3, 50.0, height=0.6)
    view = camera.view_for(cam, 96, 64)
    ov = po.orc_view(view, tile=16)
    c, s = np.cos(0.7), np.sin(0.7)
    o2w = (2.0 * c, -s, 0.0, 0.5, 2.0 * s, c, 0.0, -0.25, 0.0, 0.0, 0.5, 1.0)  # rotate z, scale (2, 1, 0.5)
    stored = np.array([[0.3, 0.2, -0.93]], np.float32)
    stored /= np.linalg.norm(stored)
    sky = skymod.constant_sky((1.0, 0.75, 0.5))
    rl = _attrs(sky, albedo=np.array([[0.4, 0.6, 0.8]], np.float32), normals=stored, o2w=o2w, roughness=0.3)
    rec, touched, _ = po.project(ov, cloud, relight=rl)
    assert touched[0] > 0
    # The disc's axis is z (the shortest scale); the eye's side of it decides
    # which way the stored normal faces (|dot| >= 0.5 here).
    p = cloud.means[0].astype(np.float64)
    eye = np.array(view.eye, np.float64)
    disc = np.array([0.0, 0.0, 1.0]) * (1 if (eye - p)[2] >= 0 else -1)
    nc = stored[0].astype(np.float64)
    assert abs(nc @ disc) >= 0.5
    nc = nc if nc @ disc >= 0 else -nc
    M = np.array(o2w, np.float64).reshape(3, 4)
    nw = np.linalg.inv(M[:, :3]).T @ nc
    nw /= np.linalg.norm(nw)
    pw = M[:, :3] @ p + M[:, 3]
    ew = M[:, :3] @ eye + M[:, 3]
    wo = (ew - pw) / np.linalg.norm(ew - pw)
    want = po.relit_point(rl, nw, wo, [0.4, 0.6, 0.8], 0.0, 0.3)
    np.testing.assert_allclose(rec[0, 8:11], want, rtol=2e-5)
