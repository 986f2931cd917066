"""The sky prepared on the device (athx_prepare_sky) against the host
quadrature it restates (athenea_amd.sky.prepare_sky, FP64 numpy): the same
texel count as the layout's (environment.slang:66-92), each f16 texel equal
or one unit in the last place apart (the two FP64 sums run in different
orders before their one rounding), the radiance coefficients to float
rounding."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import TileRasterizer, sky as skymod

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def halves(words: np.ndarray) -> np.ndarray:
    w = words.reshape(-1, 2).astype(np.uint32)
    h = np.stack([w[:, 0] & 0xFFFF, w[:, 0] >> 16, w[:, 1] & 0xFFFF], 1).astype(np.uint16)
    return h


@pytest.mark.parametrize("size,base", [((64, 32), 256), ((128, 64), 256), ((48, 24), 512)])
def test_device_sky_matches_the_host_quadrature(raster, size, base):
    say("sky", size, base)
    img = skymod.lognormal_equirect(*size)
    host = skymod.prepare_sky(img, base_side=base)
    dev = skymod.prepare_sky_device(img, base_side=base, raster=raster)
    torch.cuda.synchronize()
    got = halves(dev.texels.cpu().numpy().view(np.uint32))
    want = halves(host.texels)
    assert got.shape == want.shape
    diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
    assert (diff == 0).mean() > 0.995, (diff == 0).mean()
    assert diff.max() <= 1
    np.testing.assert_allclose(dev.sh.cpu().numpy()[:, :3], host.sh[:, :3], rtol=1e-6, atol=1e-6)


def test_a_constant_sky_is_constant(raster):
    img = np.full((32, 64, 3), 0.75, np.float32)
    dev = skymod.prepare_sky_device(img, raster=raster)
    got = halves(dev.texels.cpu().numpy().view(np.uint32)).view(np.float16).astype(np.float32)
    assert np.abs(got - 0.75).max() <= 2 ** -11 * 0.75 * 2
    sh = dev.sh.cpu().numpy()
    # the constant term: 0.75 * (the texels' solid angles, 4 pi to the
    # midpoint rule's error) * Y0; the others vanish but for that error
    assert sh[0, 0] == pytest.approx(0.75 * 4 * np.pi * 0.28209479177387814, rel=1e-3)
    assert np.abs(sh[1:, :3]).max() < 2e-3 * sh[0, 0]
