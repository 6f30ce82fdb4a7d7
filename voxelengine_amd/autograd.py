"""torch.autograd.Function over the C ABI's shading pair (vx_shade_forward /
vx_shade_backward): radiance of the ambient pass as a function of the fp32
palette and the directional light, visibility held fixed.

Used by the gradient tests; the forward and the adjoint run in libvx_amd.so.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api


class ShadeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, palette: torch.Tensor, light: torch.Tensor, vctx: api.Context, view: dict,
                gbuf: api.GBuffer, vis: api.Visibility):
        pal = palette.detach().contiguous()
        lt = light.detach().to("cpu", torch.float32).numpy().astype(np.float32)
        n = view["width"] * view["height"]
        rgb = torch.empty((n, 3), dtype=torch.float32, device=pal.device)
        api.shade_forward(vctx, view, gbuf, vis, pal, lt, rgb)
        ctx.save_for_backward(pal)
        ctx.light = lt
        ctx.light_device = light.device
        ctx.vctx, ctx.view, ctx.gbuf, ctx.vis = vctx, view, gbuf, vis
        return rgb

    @staticmethod
    def backward(ctx, grad_rgb: torch.Tensor):
        (pal,) = ctx.saved_tensors
        dpal = torch.empty((256, 6), dtype=torch.float32, device=pal.device)
        dlight = torch.empty(6, dtype=torch.float32, device=pal.device)
        api.shade_backward(ctx.vctx, ctx.view, ctx.gbuf, ctx.vis, pal, ctx.light, grad_rgb.contiguous(), dpal,
                           dlight)
        return dpal, dlight.to(ctx.light_device), None, None, None, None


def shade(palette, light, vctx, view, gbuf, vis):
    """rgb (npx, 3) = shading(palette (256, 6), light (6,)); differentiable in both."""
    return ShadeFunction.apply(palette, light, vctx, view, gbuf, vis)
