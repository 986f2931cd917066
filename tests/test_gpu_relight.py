"""The relit colour on the GPU against the oracle, through the C ABI: the
records (relit colour included), keys, ranges and images bit-exact with the
FP32 oracle, the images within 1e-4 of the FP64 replay, over the shapes of
the relight inputs (stored normals or the disc's axis, an albedo or the SH
colour, a sun or none, a turned and scaled object-to-world, binary16 SH)."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, sky as skymod, synthetic
from athenea_amd._lib import AthxError
from athenea_amd.raster import GaussianCloud
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


@pytest.fixture(scope="module")
def prepared():
    sky = skymod.prepare_sky(skymod.lognormal_equirect(), sun=((0.3, 1.0, 0.2), 0.02, (4.0, 3.5, 3.0)))
    return sky


TURNED = (0.0, -1.5, 0.0, 0.25, 1.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.75, -0.5)  # a quarter turn about z, scaled


def render_relit(raster, cloud_np, attrs, views, settings=RenderSettings(), half=False):
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0", relight=attrs)
    if half:
        cloud.sh = cloud.sh.half()
    targets, stats = raster.render(views, cloud, settings, keep_backward=True)
    bufs = raster.frame_buffers()
    torch.cuda.synchronize()
    ovs = [po.orc_view(v, tile=settings.tile) for v in views]
    ref_cloud = cloud_np
    if half:
        ref_cloud = type(cloud_np)(cloud_np.means, cloud_np.log_scales, cloud_np.quats, cloud_np.opacity_logits,
                                   cloud_np.sh.astype(np.float16).astype(np.float32))
    frame = po.render(ovs, ref_cloud, with_f64=True, relight=attrs)
    return targets, stats, bufs, frame


@pytest.mark.parametrize("albedo,normals,o2w,sun", [
    (True, True, skymod.IDENTITY_3X4, True),
    (False, False, skymod.IDENTITY_3X4, True),
    (True, False, TURNED, False),
    (False, True, TURNED, True),
])
def test_relit_frame_is_bit_exact(raster, prepared, albedo, normals, o2w, sun):
    say("relit", albedo, normals, o2w[:2], sun)
    cloud = synthetic.plain_512_cloud(30000)
    sky = prepared
    if not sun:
        sky = skymod.PreparedSky(prepared.texels, prepared.sh, np.zeros((2, 4), np.float32), prepared.base_side)
    attrs = skymod.relight_attributes(cloud.means.shape[0], sky, with_albedo=albedo, with_normals=normals)
    attrs.object_to_world = o2w
    view = camera.view_for(synthetic.front_camera(), 256, 192)
    out = render_relit(raster, cloud, attrs, [view])
    flags = out[3]["flags"]
    assert np.all((flags[out[3]["touched"] > 0] & 2048) != 0), "every visible record carries the relit flag"
    check_frame(*out)


def test_relit_multi_view_with_half_harmonics(raster, prepared):
    say("relit multi-view fp16 sh")
    cloud = synthetic.plain_512_cloud(20000)
    attrs = skymod.relight_attributes(cloud.means.shape[0], prepared, with_albedo=False)
    views = [camera.view_for(camera.orbit_camera(3.0, a, 60.0), 160, 120) for a in (0.0, 1.0, 2.5)]
    check_frame(*render_relit(raster, cloud, attrs, views, half=True))


def test_relit_backward_is_refused(raster, prepared):
    say("relit backward refused")
    cloud_np = synthetic.plain_512_cloud(2000)
    attrs = skymod.relight_attributes(2000, prepared)
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0", relight=attrs)
    view = camera.view_for(synthetic.front_camera(), 64, 64)
    targets, _ = raster.render([view], cloud, RenderSettings(), keep_backward=True)
    grads = {"means": torch.zeros_like(cloud.means)}
    with pytest.raises(AthxError) as e:
        raster.render_backward([view], cloud, RenderSettings(), targets, [torch.ones_like(targets[0].colour)], grads)
    assert e.value.status == 4  # Unsupported
