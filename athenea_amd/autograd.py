"""``torch.autograd.Function`` over the rasterizer: forward is athx_render,
backward is athx_render_backward (the hand-written adjoint).

The reference has no such path; its only gradient is a Gaussian's colour,
accumulated by the blend for NeuralFit (shaders/athenea/splat/splat_blend.slang:
381-398; modules/render/src/NeuralFit.cpp:226-228). Here every parameter gets
one: means, log-scales, quaternions, opacity logits and SH coefficients.
"""
from __future__ import annotations

import torch

from .raster import GaussianCloud, RenderSettings, TileRasterizer


class _Rasterize(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, raster, views, settings, linear):
        cloud = GaussianCloud(means.detach().contiguous(), log_scales.detach().contiguous(), quats.detach().contiguous(),
                              opacity_logits.detach().contiguous(), sh.detach().contiguous(), linear)
        targets, _ = raster.render(views, cloud, settings, keep_backward=True)
        ctx.raster = raster
        ctx.views = views
        ctx.settings = settings
        ctx.cloud = cloud
        ctx.targets = targets
        ctx.sh_dtype = sh.dtype
        for t in targets:
            ctx.mark_non_differentiable(t.depth)
        return tuple(t.colour for t in targets) + tuple(t.depth for t in targets)

    @staticmethod
    def backward(ctx, *grad_outputs):
        n = len(ctx.views)
        cloud = ctx.cloud
        dl = []
        for k in range(n):
            g = grad_outputs[k]
            dl.append(torch.zeros_like(ctx.targets[k].colour) if g is None else g)
        # Written whole by the adjoint (athx_grads.overwrite): no clearing pass.
        grads = {
            "means": torch.empty_like(cloud.means),
            "log_scales": torch.empty_like(cloud.log_scales),
            "quats": torch.empty_like(cloud.quats),
            "opacity_logits": torch.empty_like(cloud.opacity_logits),
            "sh": torch.empty(cloud.sh.shape, dtype=torch.float32, device=cloud.sh.device),
        }
        targets = ctx.targets
        if ctx.raster.frame_serial() != targets[0].frame_serial:
            # A later render on this rasterizer replaced the frame's pairs and
            # records (two forwards before one backward). The frame is drawn
            # again from the saved parameters -- the same bits, since the
            # render is deterministic -- and its adjoint taken from that.
            targets, _ = ctx.raster.render(ctx.views, cloud, ctx.settings, keep_backward=True)
        ctx.raster.render_backward(ctx.views, cloud, ctx.settings, targets, dl, grads, overwrite=True)
        g_sh = grads["sh"].to(ctx.sh_dtype)
        return (grads["means"], grads["log_scales"], grads["quats"], grads["opacity_logits"], g_sh,
                None, None, None, None)


def rasterize(
    means: torch.Tensor,
    log_scales: torch.Tensor,
    quats: torch.Tensor,
    opacity_logits: torch.Tensor,
    sh: torch.Tensor,
    views: list,
    raster: TileRasterizer,
    settings: RenderSettings | None = None,
    linear: bool = True,
) -> tuple[list[torch.Tensor], list[torch.Tensor]]:
    """Renders every view; returns (colours [h, w, 4], depths [h, w]), both
    per view. Colours are differentiable with respect to every parameter.
    Several calls may share one ``raster`` before one backward pass: a call
    whose frame a later render replaced draws it again in its backward."""
    settings = settings or RenderSettings()
    outs = _Rasterize.apply(means, log_scales, quats, opacity_logits, sh, raster, views, settings, linear)
    n = len(views)
    return list(outs[:n]), list(outs[n:])
