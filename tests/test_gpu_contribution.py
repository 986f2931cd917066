"""athx_contribution (RenderSettings::contributionMode 1 to 4) on the GPU
against orc_contribution over the same frame: every slot's words equal
(integer atomics, so the sums do not depend on their order)."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, synthetic
from athenea_amd.raster import GaussianCloud
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


@pytest.mark.parametrize("tile,batch", [(16, 256), (8, 64), (32, 128)])
def test_contributions_match_the_oracle(raster, tile, batch):
    print("contribution", tile, batch, flush=True)
    sys.stdout.flush()
    cloud_np = synthetic.plain_512_cloud(20000)
    view = camera.view_for(synthetic.front_camera(), 160, 120)
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0")
    settings = RenderSettings(tile=tile, batch=batch)
    targets, _ = raster.render([view], cloud, settings, keep_backward=True)
    rng = np.random.default_rng(tile)
    dl = rng.normal(0, 2.0, (120, 160, 4)).astype(np.float32)
    w1 = raster.contribution([view], cloud, settings, targets, mode=1)
    w4 = raster.contribution([view], cloud, settings, targets, mode=4, dl_dcolour=[torch.from_numpy(dl).to("cuda:0")])
    mask = rng.uniform(-0.2, 1.2, (120, 160)).astype(np.float32)
    w3 = raster.contribution([view], cloud, settings, targets, mode=3, dl_dcolour=[torch.from_numpy(mask).to("cuda:0")])
    drawn = targets[0].colour.contiguous()
    w2 = raster.contribution([view], cloud, settings, targets, mode=2, dl_dcolour=[drawn])
    torch.cuda.synchronize()
    ov = po.orc_view(view, tile=tile)
    frame = po.render([ov], cloud_np)
    want1 = po.contribution(ov, frame, 20000, mode=1)
    want4 = po.contribution(ov, frame, 20000, mode=4, dl=dl)
    assert np.array_equal(w1.cpu().numpy().view(np.uint32), want1)
    assert np.array_equal(w4.cpu().numpy().view(np.uint32), want4)
    assert want1.sum() > 0 and (want4[:, :3].view(np.int32) != 0).any()
    want3 = po.contribution(ov, frame, 20000, mode=3, dl=mask)
    want2 = po.contribution(ov, frame, 20000, mode=2, dl=drawn.cpu().numpy())
    assert np.array_equal(w3.cpu().numpy().view(np.uint32), want3)
    assert np.array_equal(w2.cpu().numpy().view(np.uint32), want2)
    assert (want3[:, 0] < want3[:, 1]).any() and want2[:, 0].any()
