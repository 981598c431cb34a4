"""torch.autograd over the differentiable frame: forward through
TileRasterizer.render_diff, backward through TileRasterizer.backward (the
HIP blend and preprocess backward kernels); and over a batch of views of one
cloud through render_diff_views / backward_views. Gradients flow to means,
log-scales, quaternions, opacity logits and spherical-harmonic coefficients;
the depth image is not differentiated.
"""
from __future__ import annotations

import torch

from . import raster
from ._abi import LrtError


class RasterizeSplats(torch.autograd.Function):
    """colour [H, W, 4] = raster(means, log_scales, quats, opacity_logits, sh)."""

    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, rast: raster.TileRasterizer, camera,
                settings: raster.RenderSettings, filter3d=None, object_to_world=raster.IDENTITY):
        params = [t.detach().contiguous() for t in (means, log_scales, quats, opacity_logits, sh)]
        splats = raster.DiffSplats(*params, filter3d=filter3d)
        targets = raster.RenderTargets(settings.width, settings.height, device=means.device)
        rast.render_diff(camera, splats, settings, targets, object_to_world)
        ctx.rast, ctx.splats, ctx.serial = rast, splats, rast._diff_serial
        return targets.colour

    @staticmethod
    def backward(ctx, grad_colour):
        rast = ctx.rast
        if getattr(rast, "_diff_serial", None) != ctx.serial:
            raise LrtError(3, "the rasteriser drew another differentiable frame after this one")
        s = ctx.splats
        grads = {"means": torch.empty_like(s.means), "log_scales": torch.empty_like(s.log_scales),
                 "quats": torch.empty_like(s.quats), "opacity_logits": torch.empty_like(s.opacity_logits),
                 "sh": torch.empty_like(s.sh)}
        rast.backward(grad_colour.contiguous().float(), grads)
        return (grads["means"], grads["log_scales"], grads["quats"], grads["opacity_logits"], grads["sh"],
                None, None, None, None, None)


def rasterize(means, log_scales, quats, opacity_logits, sh, rast, camera, settings, filter3d=None,
              object_to_world=raster.IDENTITY):
    """The differentiable frame's colour image, bottom row first."""
    return RasterizeSplats.apply(means, log_scales, quats, opacity_logits, sh, rast, camera, settings, filter3d,
                                 object_to_world)


class RasterizeSplatViews(torch.autograd.Function):
    """(colour_v [H_v, W_v, 4] for each camera v) = raster(means, log_scales, quats, opacity_logits, sh)."""

    @staticmethod
    def forward(ctx, means, log_scales, quats, opacity_logits, sh, rast: raster.TileRasterizer, cameras, sizes,
                settings: raster.RenderSettings, filter3d=None, object_to_world=raster.IDENTITY):
        params = [t.detach().contiguous() for t in (means, log_scales, quats, opacity_logits, sh)]
        splats = raster.DiffSplats(*params, filter3d=filter3d)
        views = [raster.View(cam, raster.RenderTargets(w, h, device=means.device)) for cam, (w, h) in
                 zip(cameras, sizes)]
        rast.render_diff_views(views, splats, settings, object_to_world)
        ctx.rast, ctx.splats, ctx.serial = rast, splats, rast._diff_serial
        return tuple(v.targets.colour for v in views)

    @staticmethod
    def backward(ctx, *grad_colours):
        rast = ctx.rast
        if getattr(rast, "_diff_serial", None) != ctx.serial:
            raise LrtError(3, "the rasteriser drew another differentiable frame after this batch")
        s = ctx.splats
        grads = {"means": torch.empty_like(s.means), "log_scales": torch.empty_like(s.log_scales),
                 "quats": torch.empty_like(s.quats), "opacity_logits": torch.empty_like(s.opacity_logits),
                 "sh": torch.empty_like(s.sh)}
        rast.backward_views([g.contiguous().float() for g in grad_colours], grads)
        return (grads["means"], grads["log_scales"], grads["quats"], grads["opacity_logits"], grads["sh"],
                None, None, None, None, None, None)


def rasterize_views(means, log_scales, quats, opacity_logits, sh, rast, cameras, sizes, settings, filter3d=None,
                    object_to_world=raster.IDENTITY):
    """The colour image of every camera (sizes: (width, height) each) of one batch, bottom row first."""
    return RasterizeSplatViews.apply(means, log_scales, quats, opacity_logits, sh, rast, cameras, sizes, settings,
                                     filter3d, object_to_world)
