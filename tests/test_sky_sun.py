"""The sun found in a sky and the sky without it (the oracle's restatement of
env_sun.slang:154-261 and env_residual.slang:55-78), pinned by what the
construction of the sky fixes: a disc sun of known radius and radiance on a
smooth sky is found at its centre, with a cone of its radius and an irradiance
of its radiance times its solid angle, and leaves no texel brighter than the
sky behind it; an overcast sky has no sun; two lamps eight degrees apart are not
one compact disc; a flat sky's median is the half stop its luminance falls in.
Directions: theta from +y, phi from +x towards +z (texel (i, j) along theta =
pi (j + 0.5) / H, phi = 2 pi (i + 0.5) / W)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle

W, H = 1024, 512


def directions(w=W, h=H):
    theta = np.pi * (np.arange(h) + 0.5) / h
    phi = 2 * np.pi * (np.arange(w) + 0.5) / w
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    return np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (h, w)), st * np.sin(phi)[None, :]], -1)


def unit(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)])


def smooth_sky():
    d = directions()
    base = 0.6 + 0.4 * np.clip(d[..., 1], 0, 1)  # brighter overhead
    return np.stack([base * 0.8, base * 0.9, base * 1.1], -1).astype(np.float32)


def add_disc(img, centre, radius, radiance):
    d = directions()
    inside = (d @ centre) >= np.cos(radius)
    out = img.copy()
    out[inside] = radiance
    return out, inside


def test_disc_sun_found_at_its_centre_with_its_size_and_energy():
    centre = unit(np.radians(50), np.radians(130))
    radius = np.radians(2.5)
    radiance = np.array([40000.0, 38000.0, 35000.0], np.float32)
    img, inside = add_disc(smooth_sky(), centre, radius, radiance)
    sun = pyoracle.sky_sun(img)
    assert sun[0, 3] > 0, "a disc a thousand times the sky is a sun"
    angle = np.degrees(np.arccos(np.clip(sun[0, :3] @ centre, -1, 1)))
    assert angle < 0.5, angle
    # The cone of the disc's solid angle, and its irradiance: radiance times
    # the solid angle 2 pi (1 - cos r), both within the texels' coarseness at
    # the 256-wide level (a texel there is 1.4 degrees across).
    assert abs(sun[0, 3] - radius) < 0.35 * radius, (sun[0, 3], radius)
    omega = 2 * np.pi * (1 - np.cos(radius))
    np.testing.assert_allclose(sun[1, :3], radiance * omega, rtol=0.15)
    # The threshold sits between the sky and the sun.
    assert 1.5 < sun[1, 3] < 0.5 * float(radiance @ np.array([0.2126, 0.7152, 0.0722]))
    res = pyoracle.sky_residual(img, sun)
    lum = res @ np.array([0.2126, 0.7152, 0.0722], np.float32)
    assert lum.max() < 2.0, "no sun texel is left"
    # Only texels of the disc are cut; the sky away from it is untouched.
    cut = np.any(res != img, axis=-1)
    assert not np.any(cut & ~inside)
    assert np.array_equal(res[~cut], img[~cut])


def test_overcast_has_no_sun():
    sun = pyoracle.sky_sun(smooth_sky())
    assert sun[0, 3] == 0 and np.all(sun[1, :3] == 0)
    res = pyoracle.sky_residual(smooth_sky(), sun)
    assert np.array_equal(res, smooth_sky())


def test_two_lamps_are_not_one_disc():
    img = smooth_sky()
    for phi in (120.0, 128.0):
        img, _ = add_disc(img, unit(np.radians(60), np.radians(phi)), np.radians(1.0), np.float32(50000.0))
    sun = pyoracle.sky_sun(img)
    assert sun[0, 3] == 0, "a region spread past twice a disc's mean square angle stays in the sky"
    # Its direction is then the brightest texel's, which is one of the lamps.
    near = [np.degrees(np.arccos(np.clip(sun[0, :3] @ unit(np.radians(60), np.radians(p)), -1, 1)))
            for p in (120.0, 128.0)]
    assert min(near) < 1.0, near


def test_flat_sky_median_and_no_sun():
    # Luminance 1 everywhere: the median is the half stop [2^0, 2^0.5), read at
    # its quarter-stop centre 2^0.25, and the peak (1) is no sun; the threshold
    # is 16 times the median.
    img = np.ones((H, W, 3), np.float32)
    sun = pyoracle.sky_sun(img)
    assert sun[0, 3] == 0
    assert sun[1, 3] == pytest.approx(16.0 * 2.0 ** 0.25, rel=1e-6)
    # One texel at 100: still the median's threshold (16 * 1.19 > 0.02 * 100),
    # and 100 > 32 * 1.19: found, a single texel is compact, so it is a sun.
    img[200, 300] = 100.0
    sun = pyoracle.sky_sun(img)
    assert sun[1, 3] == pytest.approx(16.0 * 2.0 ** 0.25, rel=1e-6)
    # The two halvings to the 256-wide level dilute the texel to 7.2, under
    # 32 * 1.19: no sun.
    assert sun[0, 3] == 0
