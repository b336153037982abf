"""torch.autograd bridge: the rendered image as a differentiable function of
the surfel parameters, with the adjoint pass as its backward.

    img = render_raw(renderer, opacity, "camera1")   # (H,W,4) HDR, cuda
    loss = 0.5 * ((img[..., :3] - target) ** 2).mean()
    loss.backward()                                 # opacity.grad from the adjoint

render_raw differentiates the opacities only: it sets them from the tensor on
the device (no BVH rebuild) and renders into device memory; the backward
sends dL/dI to the adjoint pass as a device tensor and returns its opacity
gradient (the reference's adjoint computes only that one,
SyclBridge.cpp:169-294).

render_surfels differentiates every surfel parameter -- position,
tangent_u, tangent_v (or a rotation through quaternion_tangents, in torch),
scale, albedo, opacity, beta -- with the full-gradient adjoint (the renderer
must be made with full_gradients = 1).  The forward applies the parameters
the way the reference's apply_point_optimization does (render_hooks.py:
306-316: in place, the BVH kept), device to device, and renders either a camera of the scene
(render_forward: the photon map is traced again) or any view over the
current photon map (render_view: the map stays fixed, so a forward and the
next are the same function of the parameters, which finite differences
need).  The backward regenerates the adjoint paths from their seeds (raygen
seed 42 per pixel, then the ray's rank), so nothing of the forward's paths is
stored; only the image and dL/dI cross between the passes.  As in the
reference's training loop (training.py:715-740), rebuild_bvh() reorders the
surfels: fetch the parameters again after it.

Parameters are in BVH order, the order of Renderer.get_point_parameters().
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import torch

PARAMS = ("position", "tangent_u", "tangent_v", "scale", "albedo", "opacity", "beta")


class _RenderRaw(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opacity: torch.Tensor, renderer, camera: str):
        renderer.set_opacity_device(opacity)
        img = renderer.render_forward_device(camera)[camera + "_raw"]
        ctx.renderer = renderer
        ctx.camera = camera
        ctx.device = opacity.device
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        dl = grad_img[..., :3].contiguous()
        grads = ctx.renderer.render_backward_device({ctx.camera: dl})
        return grads["opacity"].to(ctx.device), None, None


def render_raw(renderer, opacity: torch.Tensor, camera: str) -> torch.Tensor:
    """(H,W,4) HDR image of `camera` for the given (N,) opacities; the
    gradient with respect to `opacity` is the renderer's adjoint."""
    if opacity.dim() != 1 or opacity.numel() != renderer.point_count:
        raise RuntimeError(f"opacity must be ({renderer.point_count},), got {tuple(opacity.shape)}")
    return _RenderRaw.apply(opacity, renderer, camera)


class _RenderSurfels(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, view, *params):
        # The parameters go to the renderer device to device (no host copy).
        renderer.apply_point_optimization_device(dict(zip(PARAMS, params)))
        if view[0] == "camera":
            img = renderer.render_forward_device(view[1])[view[1] + "_raw"]
        else:
            img = renderer.render_view_device(view[1], view[2], view[3])["raw"]
        ctx.renderer = renderer
        ctx.view = view
        ctx.devices = tuple(p.device for p in params)
        return img

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        dl = grad_img[..., :3].contiguous()
        v = ctx.view
        if v[0] == "camera":
            grads = ctx.renderer.render_backward_device({v[1]: dl})
        else:
            grads = ctx.renderer.render_view_backward_device(v[1], v[2], dl)
        # Each gradient on its parameter's device (autograd requires it).
        return (None, None) + tuple(grads[k].to(d) for k, d in zip(PARAMS, ctx.devices))


def render_surfels(renderer, params: Dict[str, torch.Tensor], camera: Optional[str] = None,
                   view: Optional[Tuple] = None) -> torch.Tensor:
    """(H,W,4) HDR image as a differentiable function of every surfel
    parameter in `params` (keys of PARAMS, BVH order).  The image is on the
    renderer's GPU; parameters may live on any device and receive their
    gradients on it (GPU parameters never pass through host memory).

    camera="name": render_forward of a scene camera (photon map traced
    again).  view=(Camera, spp, seed): render_view over the current photon
    map (build it first with renderer.build_photon_map())."""
    if (camera is None) == (view is None):
        raise RuntimeError("render_surfels: give exactly one of camera= and view=")
    n = renderer.point_count
    shapes = {"position": (n, 3), "tangent_u": (n, 3), "tangent_v": (n, 3), "scale": (n, 2), "albedo": (n, 3),
              "opacity": (n,), "beta": (n,)}
    for k in PARAMS:
        if k not in params:
            raise RuntimeError(f"render_surfels: missing parameter '{k}'")
        if tuple(params[k].shape) != shapes[k]:
            raise RuntimeError(f"render_surfels: '{k}' must be {shapes[k]}, got {tuple(params[k].shape)}")
    geometry = any(params[k].requires_grad for k in PARAMS if k != "opacity")
    if geometry and not renderer.full_gradients:
        raise RuntimeError("render_surfels: gradients beyond opacity need a renderer made with full_gradients=1")
    key = ("camera", camera) if camera is not None else ("view",) + tuple(view)
    return _RenderSurfels.apply(renderer, key, *[params[k] for k in PARAMS])


def quaternion_tangents(q: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Tangents (first two columns of the rotation matrix) of quaternions
    q = (w, x, y, z), normalised inside, differentiable; a surfel's rotation
    parameter for render_surfels."""
    q = q / q.norm(dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    tu = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], dim=-1)
    tv = torch.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], dim=-1)
    return tu, tv
