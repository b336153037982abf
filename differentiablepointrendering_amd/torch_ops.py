"""torch.autograd bridge: the rendered image as a differentiable function of
the surfel opacities, with the adjoint pass as its backward.

    img = render_raw(renderer, opacity, "camera1")   # (H,W,4) HDR, cuda
    loss = 0.5 * ((img[..., :3] - target) ** 2).mean()
    loss.backward()                                 # opacity.grad from the adjoint

The forward sets the opacities from the tensor (no BVH rebuild) and renders
into device memory; the backward sends dL/dI to the adjoint pass as a device
tensor and returns its opacity gradient (the reference's adjoint computes
only that one, SyclBridge.cpp:169-294).  Opacities are in BVH order, the
order of Renderer.get_point_parameters().  Nothing crosses PCIe except the
opacity mirror the renderer keeps on the host for rebuilds.
"""
from __future__ import annotations

import torch


class _RenderRaw(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opacity: torch.Tensor, renderer, camera: str):
        renderer.set_opacity_device(opacity)
        img = renderer.render_forward_device(camera)[camera + "_raw"]
        ctx.renderer = renderer
        ctx.camera = camera
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        dl = grad_img[..., :3].contiguous()
        grads = ctx.renderer.render_backward_device({ctx.camera: dl})
        return grads["opacity"], None, None


def render_raw(renderer, opacity: torch.Tensor, camera: str) -> torch.Tensor:
    """(H,W,4) HDR image of `camera` for the given (N,) opacities; the
    gradient with respect to `opacity` is the renderer's adjoint."""
    if opacity.dim() != 1 or opacity.numel() != renderer.point_count:
        raise RuntimeError(f"opacity must be ({renderer.point_count},), got {tuple(opacity.shape)}")
    return _RenderRaw.apply(opacity, renderer, camera)
