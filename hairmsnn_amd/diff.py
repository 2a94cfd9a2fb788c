"""torch.autograd harness over the path tracer and its path-replay adjoint.

`render_hair(scene, sigma_a, params, spp)` renders the spp-pass mean image
with per-strand absorption `sigma_a` (num_strands x 3) and global
`params` = (beta_m, beta_n, alpha [rad]); its backward calls
hmsnn_adjoint_render with dL/d(image). Tensors live on the host: the C ABI
takes host buffers and the renderer keeps its own device state, so torch is
only the autograd front end here.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Scene


class RenderHair(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigma_a: torch.Tensor, params: torch.Tensor, scene: Scene, spp: int) -> torch.Tensor:
        if sigma_a.shape != (scene.num_strands, 3):
            raise ValueError(f"sigma_a must be ({scene.num_strands}, 3), got {tuple(sigma_a.shape)}")
        if params.shape != (3,):
            raise ValueError("params must hold (beta_m, beta_n, alpha)")
        bm, bn, al = (float(x) for x in params.detach().cpu().double())
        scene.set_hair_params(bm, bn, al)
        scene.set_strand_sigma_a(sigma_a.detach().cpu().float().numpy())
        img = scene.render_path_tracing(int(spp))[..., :3]
        ctx.scene, ctx.spp = scene, int(spp)
        ctx.dtypes = (sigma_a.dtype, params.dtype)
        return torch.from_numpy(np.ascontiguousarray(img)).to(sigma_a.dtype)

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        gs, gg = ctx.scene.adjoint(grad_img.detach().cpu().float().numpy(), ctx.spp)
        d_sigma = torch.from_numpy(gs).to(ctx.dtypes[0]) if ctx.needs_input_grad[0] else None
        d_params = torch.from_numpy(gg).to(ctx.dtypes[1]) if ctx.needs_input_grad[1] else None
        return d_sigma, d_params, None, None


def render_hair(scene: Scene, sigma_a: torch.Tensor, params: torch.Tensor, spp: int) -> torch.Tensor:
    """H x W x 3 mean image (render-buffer order, row 0 = bottom), differentiable
    in sigma_a and params."""
    return RenderHair.apply(sigma_a, params, scene, spp)
