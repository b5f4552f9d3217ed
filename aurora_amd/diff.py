"""Differentiable rendering harness: torch.autograd.Function over the C-ABI.

The reference renders only (src/render/monteCarloRaytracer.cpp:81-120); this
is the differentiable form of the AUR_MODE_PBR path loop.  Parameters:

* albedo   (n_shaders, 3) float32 -- uploaded as diffuse_color with diffuse 1
* radiance (n_lights, 3)  float32 -- uploaded as color with intensity 1

forward: aur_frame_render (device image copied out with aur_frame_copy)
backward: aur_frame_adjoint with dLoss/dI as a float4 device buffer.

With a fixed seed the renders use common random numbers, so the image is a
smooth function of the parameters when Russian roulette is off (rr_start >=
max_depth) and the gradient is exactly the derivative of the estimate.

decorrelate=True replays the backward pass with an independent seed: for a
loss such as ||I - T||^2 the product of dLoss/dI (from the forward samples)
and dI/dtheta (from independent samples) is then an unbiased estimate of the
gradient of ||E[I] - T||^2, whereas reusing the forward samples adds the
gradient of the estimator's variance.
"""
from __future__ import annotations

import numpy as np

from . import native, scenes


class DiffRenderer:
    """Owns a context with the scene's geometry built once; parameters change
    per call."""

    def __init__(self, scene: scenes.Scene, spp: int, max_depth: int, seed: int = 1,
                 rr_start: int | None = None, device: int = 0, ctx: native.Context | None = None,
                 decorrelate: bool = False):
        self.scene = scene
        self.decorrelate = decorrelate
        self.ctx = ctx or native.Context(device)
        self.ctx.upload_scene(scene)
        self.ctx.rebuild_nmh()
        self.spp, self.max_depth, self.seed = spp, max_depth, seed
        self.rr_start = max_depth if rr_start is None else rr_start
        self.ctx.frame_create(scene.width, scene.height, spp)
        self.shaders = scene.shaders.copy()
        self.shaders["diffuse"] = 1.0
        self.lights = scene.lights.copy()
        self.lights["intensity"] = 1.0

    @property
    def npix(self) -> int:
        return self.scene.width * self.scene.height

    def opts(self, spp: int | None = None, seed: int | None = None) -> native.RenderOpts:
        return native.Context.opts(native.MODE_PBR, spp or self.spp, self.max_depth, self.rr_start,
                                   self.seed if seed is None else seed)

    def set_params(self, albedo, radiance) -> None:
        a = albedo.detach().float().cpu().numpy() if hasattr(albedo, "detach") else np.asarray(albedo)
        r = radiance.detach().float().cpu().numpy() if hasattr(radiance, "detach") else np.asarray(radiance)
        if a.shape != (self.shaders.shape[0], 3) or r.shape != (self.lights.shape[0], 3):
            raise ValueError("albedo must be (n_shaders, 3) and radiance (n_lights, 3)")
        self.shaders["diffuse_color"] = a
        self.lights["color"] = r
        self.ctx.upload_shaders(self.shaders)
        self.ctx.upload_lights(self.lights)

    def render(self, albedo, radiance, spp: int | None = None, seed: int | None = None):
        """Image as a (npix, 4) float32 CUDA tensor (RGBA, row 0 = bottom)."""
        import torch

        self.set_params(albedo, radiance)
        if spp and spp != self.spp:
            self.ctx.frame_create(self.scene.width, self.scene.height, spp)
        self.ctx.frame_render(self.scene.camera, self.opts(spp, seed))
        out = torch.empty((self.npix, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.ctx.frame_copy_ptr(out.data_ptr())
        if spp and spp != self.spp:
            self.ctx.frame_create(self.scene.width, self.scene.height, self.spp)
        return out

    def adjoint(self, albedo, radiance, d_image, seed: int | None = None):
        """(grad_albedo, grad_radiance) CUDA tensors for dLoss/dI = d_image
        ((npix, 4) or (npix, 3) CUDA tensor)."""
        import torch

        self.set_params(albedo, radiance)
        d = d_image.detach().float()
        if d.shape[-1] == 3:
            d = torch.nn.functional.pad(d, (0, 1))
        d = d.reshape(self.npix, 4).contiguous()
        ga = torch.empty((self.shaders.shape[0], 3), dtype=torch.float32, device="cuda")
        gr = torch.empty((max(self.lights.shape[0], 1), 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.ctx.frame_adjoint_ptr(self.scene.camera, self.opts(seed=seed), d.data_ptr(), ga.data_ptr(),
                                   gr.data_ptr())
        return ga, gr[: self.lights.shape[0]]

    def close(self) -> None:
        self.ctx.close()


class RenderFunction:
    """image = RenderFunction.apply(renderer, albedo, radiance) -> (npix, 3)."""

    _fn = None

    @classmethod
    def apply(cls, renderer: DiffRenderer, albedo, radiance):
        if cls._fn is None:
            cls._fn = _make_function()
        return cls._fn.apply(renderer, albedo, radiance)


def _make_function():
    import torch

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, renderer, albedo, radiance):
            img = renderer.render(albedo, radiance)
            ctx.renderer = renderer
            # the forward's seed, or an independent one for a decorrelated replay
            ctx.seed = (renderer.seed ^ 0x5BD1E995) & 0xFFFFFFFF if renderer.decorrelate else renderer.seed
            ctx.save_for_backward(albedo.detach().clone(), radiance.detach().clone())
            return img[:, :3].contiguous()

        @staticmethod
        def backward(ctx, grad_out):
            albedo, radiance = ctx.saved_tensors
            ga, gr = ctx.renderer.adjoint(albedo, radiance, grad_out.contiguous(), seed=ctx.seed)
            return None, ga, gr

    return _Render


def render(renderer: DiffRenderer, albedo, radiance):
    """Differentiable render: (npix, 3) image tensor."""
    return RenderFunction.apply(renderer, albedo, radiance)
