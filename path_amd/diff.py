"""Differentiable rendering: a torch.autograd.Function around the mega-kernel
render (forward) and the path-replay adjoint (backward).

Parameters are every material's albedo color (`diffuse` for Diffuse
materials, `specular` for Glossy / Mirror / DielectricRefractive) and emission,
as (n_materials, 3) tensors. The forward pass writes them into the scene
(geometry and BVH stay resident) and renders; the backward pass replays every
path from its Sobol coordinates on the GPU and returns dLoss/dalbedo and
dLoss/demission. The reference has no gradients; its forward estimator is
Renderer::render (src/render/renderer.cu:10-45).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .renderer import Renderer, Scene, Settings


def _albedo_field(m: L.PtMaterial) -> str:
    return "diffuse" if m.type == L.BSDF["Diffuse"] else "specular"


class DiffRenderer:
    """One scene + camera + sample budget, rendered as a function of material
    parameters."""

    def __init__(self, renderer: Renderer, scene: Scene, film_to_world, x_dim: int,
                 y_dim: int, samples_per: int, settings: Settings | None = None):
        self.r = renderer
        self.scene = scene
        self.ftw = np.ascontiguousarray(np.asarray(film_to_world).ravel(), np.float32)
        self.x_dim, self.y_dim, self.spp = x_dim, y_dim, samples_per
        s = settings.copy() if settings is not None else Settings()
        s.kernel_approach = "MegaKernel"
        self.settings = s

    def parameters(self):
        """Current (albedo, emission) as float32 tensors of shape (n, 3)."""
        mats = self.scene.materials()
        alb = [list(getattr(m, _albedo_field(m))) for m in mats]
        em = [list(m.emission) for m in mats]
        return (torch.tensor(alb, dtype=torch.float32), torch.tensor(em, dtype=torch.float32))

    def set_parameters(self, albedo, emission) -> None:
        a = np.asarray(albedo, np.float32).reshape(-1, 3)
        e = np.asarray(emission, np.float32).reshape(-1, 3)
        for i, m in enumerate(self.scene.materials()):
            getattr(m, _albedo_field(m))[:] = [float(x) for x in a[i]]
            m.emission[:] = [float(x) for x in e[i]]
            self.scene.set_material(i, m)

    def render(self) -> np.ndarray:
        img, _ = self.r.render(self.scene, self.ftw, self.x_dim, self.y_dim, self.spp,
                               self.settings, output="FloatRGB")
        return img

    def adjoint(self, d_image: np.ndarray):
        ga, ge, _ = self.r.adjoint(self.scene, self.ftw, self.x_dim, self.y_dim, self.spp,
                                   d_image, self.settings)
        return ga, ge

    def __call__(self, albedo: torch.Tensor, emission: torch.Tensor) -> torch.Tensor:
        return RenderFunction.apply(albedo, emission, self)


class RenderFunction(torch.autograd.Function):
    """image = render(albedo, emission); backward = path-replay adjoint."""

    @staticmethod
    def forward(ctx, albedo, emission, dr: DiffRenderer):
        alb = albedo.detach().to("cpu", torch.float32).numpy().copy()
        em = emission.detach().to("cpu", torch.float32).numpy().copy()
        dr.set_parameters(alb, em)
        img = torch.from_numpy(dr.render())
        ctx.dr = dr
        # The adjoint depends on the parameters the image was rendered with;
        # another forward through the same DiffRenderer before this backward
        # changes the scene, so keep this call's values.
        ctx.params = (alb, em)
        ctx.devices = (albedo.device, emission.device)
        return img.to(albedo.device)

    @staticmethod
    def backward(ctx, grad_image):
        dr = ctx.dr
        cur_alb, cur_em = (t.numpy() for t in dr.parameters())
        alb, em = ctx.params
        swap = not (np.array_equal(cur_alb, alb) and np.array_equal(cur_em, em))
        if swap:
            dr.set_parameters(alb, em)
        try:
            ga, ge = dr.adjoint(grad_image.detach().to("cpu", torch.float32).contiguous().numpy())
        finally:
            if swap:
                dr.set_parameters(cur_alb, cur_em)
        return (torch.from_numpy(ga).to(ctx.devices[0]), torch.from_numpy(ge).to(ctx.devices[1]),
                None)
