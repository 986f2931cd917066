"""The sun found in a sky on the device (athx_sky_sun) and the sky without it
(athx_sky_residual), through the C ABI, against the oracle's restatement of
env_sun.slang and env_residual.slang: the threshold, whether a sun was found
and the residual exactly (integer histogram, float tables and float tests on
both sides); the direction, cone and irradiance to the rounding of their FP64
sums, which run in different orders. And a sky prepared from an equirect with
its sun taken out (prepare_sky_device(sun="extract"), Environment::build's
order): its sun words the oracle's, its levels and harmonics the host
quadrature's of the oracle's residual."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import TileRasterizer, sky as skymod
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def directions(w, h):
    theta = np.pi * (np.arange(h) + 0.5) / h
    phi = 2 * np.pi * (np.arange(w) + 0.5) / w
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    return np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (h, w)), st * np.sin(phi)[None, :]], -1)


def sky_with(w, h, suns, seed):
    """A lognormal sky with discs (theta, phi, radius in degrees, radiance)."""
    img = skymod.lognormal_equirect(w, h, sigma=0.3, seed=seed).astype(np.float32)
    d = directions(w, h)
    for th, ph, r, rad in suns:
        c = np.array([np.sin(np.radians(th)) * np.cos(np.radians(ph)), np.cos(np.radians(th)),
                      np.sin(np.radians(th)) * np.sin(np.radians(ph))])
        img[(d @ c) >= np.cos(np.radians(r))] = rad
    return img


CASES = [
    ("disc", 1024, 512, [(45, 200, 2.0, 30000.0)], 1),
    ("low sun", 2048, 1024, [(84, 20, 0.6, 90000.0)], 2),
    ("overcast", 512, 256, [], 3),
    ("two lamps", 1024, 512, [(60, 120, 1.0, 50000.0), (60, 128, 1.0, 50000.0)], 4),
    ("odd size", 1000, 499, [(30, 300, 3.0, 20000.0)], 5),
]


@pytest.mark.parametrize("name,w,h,suns,seed", CASES, ids=[c[0] for c in CASES])
def test_device_sun_and_residual_match_the_oracle(raster, name, w, h, suns, seed):
    say("sky sun", name, w, h)
    img = sky_with(w, h, suns, seed)
    want = pyoracle.sky_sun(img)
    got = skymod.sun_of(img, raster)
    assert got[1, 3] == want[1, 3], (got[1, 3], want[1, 3])  # the threshold
    assert (got[0, 3] > 0) == (want[0, 3] > 0)
    np.testing.assert_allclose(got[0, :3], want[0, :3], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(got[0, 3], want[0, 3], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(got[1, :3], want[1, :3], rtol=1e-5, atol=1e-6)
    res = skymod.residual_of(img, want, raster).cpu().numpy()
    assert np.array_equal(res, pyoracle.sky_residual(img, want))


def test_prepared_sky_takes_its_sun_out(raster):
    w, h = 128, 64
    img = sky_with(w, h, [(40, 70, 4.0, 20000.0)], 7)
    dev = skymod.prepare_sky_device(img, sun="extract", raster=raster)
    torch.cuda.synchronize()
    sun = pyoracle.sky_sun(img)
    assert sun[0, 3] > 0
    np.testing.assert_allclose(dev.sun.cpu().numpy(), sun, rtol=1e-5, atol=1e-6)
    host = skymod.prepare_sky(pyoracle.sky_residual(img, sun))
    words = dev.texels.cpu().numpy().view(np.uint32).reshape(-1, 2)
    got = np.stack([words[:, 0] & 0xFFFF, words[:, 0] >> 16, words[:, 1] & 0xFFFF], 1).astype(np.int64)
    hw = host.texels.reshape(-1, 2).astype(np.uint32)
    want = np.stack([hw[:, 0] & 0xFFFF, hw[:, 0] >> 16, hw[:, 1] & 0xFFFF], 1).astype(np.int64)
    diff = np.abs(got - want)
    assert diff.max() <= 1 and (diff == 0).mean() > 0.995
    np.testing.assert_allclose(dev.sh.cpu().numpy()[:, :3], host.sh[:, :3], rtol=1e-6, atol=1e-6)
    # The sun is out of the harmonics: the residual's constant term is the
    # smooth sky's, far under the sky with its sun.
    with_sun = skymod.prepare_sky(img)
    assert host.sh[0, 1] < 0.5 * with_sun.sh[0, 1]
