"""The autotuner's candidates draw the same image, and it returns one of them."""
from __future__ import annotations

import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, synthetic
from athenea_amd.autotune import autotune
from athenea_amd.raster import GaussianCloud

pytestmark = pytest.mark.gpu


def test_every_tile_and_batch_draws_the_same_image():
    print("autotune shapes", flush=True)
    raster = TileRasterizer(0)
    cloud = GaussianCloud.from_numpy(synthetic.plain_512_cloud(40000), "cuda:0")
    views = [camera.view_for(synthetic.front_camera(), 200, 150)]
    ref, _ = raster.render(views, cloud, RenderSettings(tile=16, batch=256))
    want = ref[0].colour.clone()
    for tile in (8, 16, 32):
        for batch in (32, 96, 256):
            got, _ = raster.render(views, cloud, RenderSettings(tile=tile, batch=batch))
            assert torch.equal(got[0].colour, want), (tile, batch)
    best, table = autotune(raster, views, cloud, reps=2)
    assert len(table) == 12
    assert (best.tile, best.batch) in {(r["tile"], r["batch"]) for r in table}
    again, _ = autotune(raster, views, cloud)
    assert again is best
    raster.close()
