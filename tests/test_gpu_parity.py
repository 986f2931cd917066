"""Parity of the HIP path with the oracle, through the C ABI, on the GPU.

The bar: bit-exact projection records, tile counts, depth keys, scan offsets,
sorted 64-bit keys, pair values and tile ranges; images within 1e-4 max abs
(linear RGB and alpha) of the FP64 oracle. Sizes keep the oracle to seconds.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, synthetic
from athenea_amd._lib import AthxError
from athenea_amd.camera import Camera
from athenea_amd.raster import GaussianCloud, SplatInstance
from oracle import pyoracle as po
from tests.splat_fixtures import counters_cloud, random_cloud
from tests.test_oracle import sort_patterns

pytestmark = pytest.mark.gpu

IMAGE_TOL = 1e-4  # max abs, linear RGB and alpha, against the FP64 oracle


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def u32(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def render_both(raster, cloud_np, views, settings: RenderSettings, linear=True, f64=True):
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0", linear=linear)
    targets, stats = raster.render(views, cloud, settings, keep_backward=True)
    bufs = raster.frame_buffers()
    torch.cuda.synchronize()
    ovs = [po.orc_view(v, tile=settings.tile, antialias=settings.antialias, linear=linear,
                       sh_limit=settings.max_sh_degree, depth_mode=0 if settings.depth == "mean" else 1,
                       depth_threshold=settings.depth_threshold) for v in views]
    frame = po.render(ovs, cloud_np, with_f64=f64)
    return targets, stats, bufs, frame


def check_frame(targets, stats, bufs, frame, f64=True):
    touched = u32(bufs["touched"])
    assert np.array_equal(touched, frame["touched"]), "tile counts"
    assert np.array_equal(u32(bufs["depth_keys"]), frame["depth_keys"]), "depth keys and cull reasons"
    vis = touched > 0
    rec = bufs["records"].cpu().numpy()
    assert np.array_equal(rec[vis].view(np.uint32), frame["records"][vis].view(np.uint32)), "records"
    # A slot's first pair (tiles counted in depth order); defined for visible slots.
    assert np.array_equal(u32(bufs["offsets"])[vis], frame["offsets"][vis]), "offsets"
    assert stats.pairs == frame["pairs"]
    assert np.array_equal(bufs["pair_keys"].cpu().numpy().view(np.uint64), frame["pair_keys"]), "sorted keys"
    assert np.array_equal(u32(bufs["pair_values"]), frame["pair_values"]), "pair values"
    assert np.array_equal(u32(bufs["ranges"]).reshape(-1), frame["ranges"]), "tile ranges"
    for k, t in enumerate(targets):
        col = t.colour.cpu().numpy()
        assert np.array_equal(col, frame["colour"][k]), f"view {k}: image differs from the FP32 oracle"
        assert np.array_equal(t.final_t.cpu().numpy(), frame["final_t"][k])
        assert np.array_equal(t.last.cpu().numpy().view(np.uint32), frame["last"][k])
        if f64:
            err = float(np.abs(col.astype(np.float64) - frame["colour64"][k]).max())
            assert err <= IMAGE_TOL, f"view {k}: {err} from the FP64 oracle"


def test_plain_512_frame_is_bit_exact(raster):
    """100k Gaussians at 512x512 (the CPU-runnable workload), linear."""
    say("plain_512 100k")
    cloud = synthetic.plain_512_cloud()
    view = camera.view_for(synthetic.front_camera(), 512, 512)
    check_frame(*render_both(raster, cloud, [view], RenderSettings()))


@pytest.mark.parametrize("tile,batch", [(8, 64), (16, 32), (16, 128), (32, 256)])
def test_tile_and_batch_shapes(raster, tile, batch):
    say("tile", tile, "batch", batch)
    cloud = synthetic.plain_512_cloud(30000)
    view = camera.view_for(synthetic.front_camera(), 300, 220)
    check_frame(*render_both(raster, cloud, [view], RenderSettings(tile=tile, batch=batch)))


def test_reference_random_cloud_srgb_partial_tiles(raster):
    """test_splat_render.cpp:49-128's scene: 3000 splats at 250x190 (partial
    tiles), sRGB colours, a camera off axis."""
    say("random cloud sRGB")
    cloud = random_cloud(3000, 7)
    view = camera.view_for(Camera.looking_at((0.5, 0.3, 6.0), (0, 0, 0)), 250, 190)
    check_frame(*render_both(raster, cloud, [view], RenderSettings(), linear=False))


def test_orthographic_no_antialias_background(raster):
    say("orthographic")
    cloud = random_cloud(2000, 3)
    cam = Camera.looking_at((0.0, 0.0, 6.0), (0, 0, 0))
    cam.lens.orthographic = True
    cam.lens.focal = cam.lens.haperture / 5.0
    view = camera.view_for(cam, 160, 120, background=(0.1, 0.2, 0.3, 0.5))
    check_frame(*render_both(raster, cloud, [view], RenderSettings(antialias=False), linear=False))


def test_threshold_depth_and_sh_limit(raster):
    say("depth threshold, sh limit 1")
    cloud = synthetic.plain_512_cloud(20000)
    view = camera.view_for(synthetic.front_camera(), 128, 128)
    check_frame(*render_both(raster, cloud, [view], RenderSettings(depth="threshold", depth_threshold=0.3,
                                                                   max_sh_degree=1)))


def test_multi_view_batch(raster):
    """Views of different sizes and visible counts in one launch."""
    say("multi-view")
    cloud = synthetic.walk_cloud(40000)
    cams = synthetic.walk_cameras(60)
    views = [camera.view_for(cams[k], w, h) for k, (w, h) in zip((0, 13, 29, 47), [(240, 135), (96, 96), (200, 50),
                                                                                      (64, 160)])]
    check_frame(*render_both(raster, cloud, views, RenderSettings()))


def test_half_precision_sh(raster):
    """FP16-stored SH: the same frame as FP32 SH holding the same values."""
    say("fp16 sh")
    cloud = synthetic.plain_512_cloud(20000)
    cloud.sh = cloud.sh.astype(np.float16)
    view = camera.view_for(synthetic.front_camera(), 128, 96)
    g16 = GaussianCloud.from_numpy(cloud, "cuda:0")
    t16, _ = raster.render([view], g16, RenderSettings())
    cloud.sh = cloud.sh.astype(np.float32)
    t32, _ = raster.render([view], GaussianCloud.from_numpy(cloud, "cuda:0"), RenderSettings())
    assert torch.equal(t16[0].colour, t32[0].colour)


def test_counters_match_the_reference_test():
    """test_splat_render.cpp:1831-1917: two clouds as two instances, tags 41
    and 42, nothing before a counted frame, a quiet frame leaving the last."""
    say("counters")
    raster = TileRasterizer(0)
    try:
        cam = Camera()
        cam.lens.focal = cam.lens.haperture
        view = camera.view_for(cam, 320, 240)
        both = counters_cloud()
        first = GaussianCloud.from_numpy(both.subset(22), "cuda:0", linear=False)
        rest = synthetic.Cloud(*(np.ascontiguousarray(getattr(both, k)[22:]) for k in
                                 ("means", "log_scales", "quats", "opacity_logits", "sh")))
        second = GaussianCloud.from_numpy(rest, "cuda:0", linear=False)
        instances = [SplatInstance(first), SplatInstance(second)]
        assert raster.latest_counters() is None
        for tag in (41, 42):
            _, stats = raster.render_instances([view], instances, RenderSettings(count_splats=True, counters_tag=tag))
            c = raster.latest_counters()
            assert c is not None and c.tag == tag
            assert c.slots == 28 and c.visible == 16 and c.visible == stats.visible
            assert c.culled["depth"] == 4 and c.culled["faint"] == 3 and c.culled["offscreen"] == 5
            assert c.culled["edit"] == 0 and c.culled["unprojected"] == 0
            assert sum(c.culled.values()) + c.visible == c.slots
            assert c.pairs == stats.pairs and c.max_tiles >= 1
            assert len(c.clouds) == 2 and c.clouds[0].visible == 10 and c.clouds[1].visible == 6
            assert c.clouds[0].pairs + c.clouds[1].pairs == c.pairs
        raster.render_instances([view], instances, RenderSettings(count_splats=False))
        still = raster.latest_counters()
        assert still is not None and still.tag == 42
    finally:
        raster.close()


def test_counters_agree_with_the_exact_tile_counts(raster):
    """The counters' pairs and most tiles, taken from the masks the emission
    decided, against the frame's exact per-slot tile counts."""
    say("counters vs exact counts")
    cloud = GaussianCloud.from_numpy(synthetic.plain_1080p_cloud(200000), "cuda:0")
    views = [camera.view_for(synthetic.front_camera(), 640, 360), camera.view_for(synthetic.walk_cameras(60)[5], 320, 200)]
    _, stats = raster.render(views, cloud, RenderSettings(count_splats=True, counters_tag=7))
    c = raster.latest_counters()
    touched = u32(raster.frame_buffers()["touched"]).astype(np.int64)
    assert c.tag == 7 and c.slots == 2 * 200000
    assert c.visible == stats.visible == int((touched > 0).sum())
    assert c.pairs == stats.pairs == int(touched.sum())
    assert c.max_tiles == int(touched.max()) > 16
    assert len(c.clouds) == 1 and c.clouds[0].visible == c.visible and c.clouds[0].pairs == c.pairs


def test_rendering_twice_gives_the_same_bits(raster):
    say("determinism")
    cloud = GaussianCloud.from_numpy(synthetic.plain_512_cloud(50000), "cuda:0")
    view = camera.view_for(synthetic.front_camera(), 256, 256)
    a, _ = raster.render([view], cloud, RenderSettings())
    first = a[0].colour.clone()
    b, _ = raster.render([view], cloud, RenderSettings())
    assert torch.equal(first, b[0].colour)


@pytest.mark.parametrize("n", [1, 7, 4095, 4096, 4097, 100003, 1 << 22, (1 << 24) + 4099])
def test_scan(raster, n):
    """test_algo.cpp:72-109's sizes, and one past 2^24 elements, where the scan
    runs as partition sums, their scan and the partitions scanned from them."""
    say("scan", n)
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n, dtype=np.uint32)
    out, total = raster.exclusive_scan(torch.from_numpy(x.view(np.int32)).cuda())
    torch.cuda.synchronize()
    want, want_total = po.exclusive_scan(x)
    assert np.array_equal(u32(out), want) and int(total.item()) == want_total


@pytest.mark.parametrize("n,bits", [(1, 32), (2, 32), (4097, 13), (65537, 32), (300007, 24), (250001, 64),
                                    (100000, 40), (2_000_003, 37)])
def test_sort(raster, n, bits):
    """test_algo.cpp:347-379's sizes and patterns, against the oracle's sort
    (itself pinned to numpy's stable argsort)."""
    for name, keys in sort_patterns(n, bits, n + bits):
        say("sort", n, bits, name)
        noisy = keys | (np.uint64(0xAB) << np.uint64(56)) if bits <= 48 else keys
        vals = np.arange(n, dtype=np.uint32)
        kt = torch.from_numpy(noisy.view(np.int64).copy()).cuda()
        vt = torch.from_numpy(vals.view(np.int32).copy()).cuda()
        raster.sort_pairs(kt, vt, bits)
        torch.cuda.synchronize()
        wk, wv = po.sort_pairs(noisy, vals, bits)
        assert np.array_equal(kt.cpu().numpy().view(np.uint64), wk), name
        assert np.array_equal(u32(vt), wv), name


def test_bad_arguments_raise_invalid_argument(raster):
    cloud = GaussianCloud.from_numpy(synthetic.plain_512_cloud(100), "cuda:0")
    view = camera.view_for(synthetic.front_camera(), 64, 64)
    for bad in (RenderSettings(tile=12), RenderSettings(batch=48), RenderSettings(batch=512)):
        with pytest.raises(AthxError) as e:
            raster.render([view], cloud, bad)
        assert e.value.code == "InvalidArgument"
    zero = camera.view_for(synthetic.front_camera(), 64, 64)
    zero.width = 0
    with pytest.raises((AthxError, ValueError)):
        raster.render([zero], cloud, RenderSettings())


def test_empty_cloud_draws_the_background(raster):
    cloud = GaussianCloud.from_numpy(synthetic.plain_512_cloud(0), "cuda:0")
    view = camera.view_for(synthetic.front_camera(), 40, 30, background=(0.25, 0.5, 0.75, 1.0))
    t, stats = raster.render([view], cloud, RenderSettings())
    assert stats.pairs == 0
    col = t[0].colour.cpu().numpy()
    assert np.allclose(col, [0.25, 0.5, 0.75, 1.0])


def test_many_views_take_the_long_scans(raster):
    """60 views of a 300k-Gaussian cloud: 18M (view, Gaussian) slots, past the
    2^24 where the compaction and the strip-entry scan run as partition sums,
    their scan, and the partitions scanned from them. Bit-exact with the
    oracle, as every frame."""
    say("60 views x 300k")
    cloud = synthetic.walk_cloud(300000)
    views = [camera.view_for(c, 64, 48) for c in synthetic.walk_cameras(60)]
    check_frame(*render_both(raster, cloud, views, RenderSettings(), f64=False), f64=False)
