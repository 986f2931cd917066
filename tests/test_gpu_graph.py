"""The wait-free frame (athx_render_async) and its hipGraph capture
(athx_graph_*), against the eager frame, through the C ABI.

Replays are checked several times over, with PyTorch work and host
synchronisations between them, and after the Gaussians change in place.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from athenea_amd import RenderSettings, TileRasterizer, camera, synthetic
from athenea_amd.autograd import rasterize
from athenea_amd.raster import FrameGraph, GaussianCloud

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def walk_views(count=4, sizes=((160, 90), (96, 96), (128, 64), (80, 120))):
    cams = synthetic.walk_cameras(60)
    return [camera.view_for(cams[k * 13], w, h) for k, (w, h) in zip(range(count), sizes)]


@pytest.fixture(scope="module")
def raster():
    r = TileRasterizer(0)
    yield r
    r.close()


def test_async_frame_equals_eager(raster):
    say("async == eager")
    cloud = GaussianCloud.from_numpy(synthetic.walk_cloud(30000), "cuda:0")
    views = walk_views()
    eager, stats = raster.render(views, cloud, RenderSettings())
    want = [t.colour.clone() for t in eager]
    want_bufs = raster.frame_buffers()
    for cap in (stats.pairs, stats.pairs * 2 + 7):
        targets = raster.render_async(views, cloud, RenderSettings(), pair_capacity=cap)
        pairs, got_cap = raster.frame_pairs()
        assert pairs == stats.pairs and got_cap == cap
        for k, t in enumerate(targets):
            assert torch.equal(t.colour, want[k])
        bufs = raster.frame_buffers()
        for name in ("pair_keys", "pair_values", "ranges", "touched", "depth_keys"):
            assert torch.equal(bufs[name], want_bufs[name]), name


def test_async_frame_reports_an_overflow(raster):
    say("async overflow")
    cloud = GaussianCloud.from_numpy(synthetic.walk_cloud(30000), "cuda:0")
    views = walk_views()
    _, stats = raster.render(views, cloud, RenderSettings())
    cap = max(stats.pairs // 3, 2)
    raster.render_async(views, cloud, RenderSettings(), pair_capacity=cap)
    pairs, got_cap = raster.frame_pairs()
    assert pairs == stats.pairs and got_cap == cap and pairs > got_cap
    # A frame after an overflow is whole again.
    targets = raster.render_async(views, cloud, RenderSettings(), pair_capacity=stats.pairs)
    eager, _ = raster.render(views, cloud, RenderSettings())
    assert torch.equal(targets[0].colour, eager[0].colour)


def test_graph_replays_equal_the_eager_frame(raster):
    say("graph replays")
    cloud_np = synthetic.walk_cloud(30000)
    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0")
    views = walk_views()
    graph = FrameGraph(views, cloud, RenderSettings())
    eager, stats = raster.render(views, cloud, RenderSettings())
    for rep in range(4):
        graph.launch()
        noise = torch.randn(1000, 1000, device="cuda") @ torch.randn(1000, 1000, device="cuda")
        torch.cuda.synchronize()
        pairs, cap = graph.pairs()
        assert pairs == stats.pairs and cap >= pairs
        for k in range(len(views)):
            assert torch.equal(graph.targets[k].colour, eager[k].colour), (rep, k)
        del noise
    # The Gaussians change in place: the next replay draws the new values.
    cloud.means.add_(0.05)
    graph.launch()
    torch.cuda.synchronize()
    moved, _ = raster.render(views, cloud, RenderSettings())
    for k in range(len(views)):
        assert torch.equal(graph.targets[k].colour, moved[k].colour)
    graph.close()


def test_graph_adjoint_equals_the_eager_adjoint(raster):
    say("graph adjoint")
    cloud_np = synthetic.plain_1080p_cloud(20000)
    view = camera.view_for(synthetic.front_camera(), 128, 96)
    settings = RenderSettings()
    params = [torch.from_numpy(a.copy()).cuda().requires_grad_(True)
              for a in (cloud_np.means, cloud_np.log_scales, cloud_np.quats, cloud_np.opacity_logits, cloud_np.sh)]
    colours, _ = rasterize(*params, [view], raster, settings)
    target = torch.rand_like(colours[0])
    ((colours[0] - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    eager_grads = [p.grad.clone() for p in params]

    cloud = GaussianCloud.from_numpy(cloud_np, "cuda:0")
    dl = torch.zeros((96, 128, 4), dtype=torch.float32, device="cuda")
    grads = {name: torch.zeros_like(t) for name, t in zip(
        ("means", "log_scales", "quats", "opacity_logits", "sh"),
        (cloud.means, cloud.log_scales, cloud.quats, cloud.opacity_logits, cloud.sh))}
    graph = FrameGraph([view], cloud, settings, dl_dcolour=[dl], grads=grads)
    for rep in range(3):
        for g in grads.values():
            g.zero_()
        graph.launch()
        dl.copy_(2.0 * (graph.targets[0].colour - target))
        graph.launch_backward()
        torch.cuda.synchronize()
        assert torch.equal(graph.targets[0].colour, colours[0].detach())
        for name, e in zip(("means", "log_scales", "quats", "opacity_logits", "sh"), eager_grads):
            got = grads[name]
            scale = float(e.abs().max()) or 1.0
            assert float((got - e).abs().max()) <= 1e-5 * scale, (rep, name)
    graph.close()
