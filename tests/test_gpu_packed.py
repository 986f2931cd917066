"""athenea's packed device layout (athx_packed_splats) on the GPU against the
oracle's decode of the same words, through the C ABI: records, keys, ranges
and images bit-exact with the FP32 oracle, images within 1e-4 of the FP64
replay, at every SH width the layout holds; relit from the packed normal and
pbr words too."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, packing, sky as skymod, synthetic
from athenea_amd._lib import AthxError
from oracle import pyoracle as po
from tests.test_gpu_parity import check_frame

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def render_packed(raster, packed, views, relight=None, settings=RenderSettings()):
    cloud = packed.to_device("cuda:0")
    if relight is not None:
        cloud.relight = relight.to_device("cuda:0")
        # the packed words, read instead of the float arrays
        cloud.relight.pbr = torch.from_numpy(relight.pbr.view(np.int32)).to("cuda:0")
        cloud.relight.normals_packed = torch.from_numpy(relight.normals_packed.view(np.int32)).to("cuda:0")
        cloud.relight.metallic = None
        cloud.relight.roughness = None
        cloud.relight.normals = None
    targets, stats = raster.render(views, cloud, settings, keep_backward=True)
    bufs = raster.frame_buffers()
    torch.cuda.synchronize()
    ovs = [po.orc_view(v, tile=settings.tile) for v in views]
    frame = po.render(ovs, packed, with_f64=True, relight=relight, f64_cloud=packing.unpacked_fields(packed))
    return targets, stats, bufs, frame


@pytest.mark.parametrize("coeffs", [1, 4, 9, 16])
def test_packed_frame_is_bit_exact(raster, coeffs):
    say("packed coeffs", coeffs)
    cloud = synthetic.plain_512_cloud(30000)
    cloud.sh = np.ascontiguousarray(cloud.sh[:, :coeffs, :])
    packed = packing.pack_cloud(cloud)
    view = camera.view_for(synthetic.front_camera(), 256, 192)
    check_frame(*render_packed(raster, packed, [view]))


def test_packed_relit_multi_view(raster):
    say("packed relit multi-view")
    cloud = synthetic.plain_512_cloud(20000)
    packed = packing.pack_cloud(cloud)
    sky = skymod.prepare_sky(skymod.lognormal_equirect(), sun=((0.2, 1.0, 0.4), 0.02, (3.0, 3.0, 2.5)))
    attrs = skymod.relight_attributes(20000, sky, with_albedo=False)
    attrs.pbr = packing.pack_pbr(attrs.metallic, attrs.roughness)
    attrs.normals_packed = packing.pack_normals(attrs.normals)
    views = [camera.view_for(camera.orbit_camera(3.0, a, 60.0), 160, 120) for a in (0.3, 2.0)]
    out = render_packed(raster, packed, views, relight=attrs)
    check_frame(*out)


def test_packed_backward_is_refused(raster):
    say("packed backward refused")
    packed = packing.pack_cloud(synthetic.plain_512_cloud(2000))
    cloud = packed.to_device("cuda:0")
    view = camera.view_for(synthetic.front_camera(), 64, 64)
    targets, _ = raster.render([view], cloud, RenderSettings(), keep_backward=True)
    grads = {"means": torch.zeros((2000, 3), device="cuda:0")}
    with pytest.raises(AthxError) as e:
        raster.render_backward([view], cloud, RenderSettings(), targets, [torch.ones_like(targets[0].colour)], grads)
    assert e.value.status == 4  # Unsupported
