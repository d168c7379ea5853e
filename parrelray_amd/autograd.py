"""torch.autograd wrapper of the Kerr render and its adjoint kernel.

    render = KerrOrbitRender(ctx, template_frame, orbit)
    img = render(params)          # params: float64 tensor (spin, r_cam, incl, sphere_phi)
    loss = ((img - target) ** 2).mean()
    loss.backward()               # params.grad from prr_adjoint (checkpointed reverse sweep)

The forward pass is prr_render of the orbiting frame; the backward pass
calls prr_adjoint with grad_output as dL/dI.  The image is the linear-RGB
SSAA result (no FXAA: the adjoint differentiates the SSAA image).
"""
import copy

import torch

from . import _abi as A
from . import api


class _KerrOrbitFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, module):
        p = params.detach().cpu().double().tolist()
        orbit = copy.copy(module.orbit)
        orbit.spin, orbit.r_cam, orbit.incl, orbit.sphere_phi = p
        frame = api.frame_from_orbit(module.tmpl, orbit)
        img = torch.empty((frame.height, frame.width, 3), dtype=torch.float32, device=module.device)
        module.ctx.render_device([frame], rgbf=[img])
        ctx.module = module
        ctx.orbit = orbit
        ctx.save_for_backward(params)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        (params,) = ctx.saved_tensors
        module = ctx.module
        g = module.ctx.adjoint(module.tmpl, ctx.orbit, grad_img.contiguous().float())
        return torch.tensor(g, dtype=params.dtype, device=params.device), None


class KerrOrbitRender(torch.nn.Module):
    """Differentiable render of an orbiting camera around a Kerr hole."""

    def __init__(self, ctx, tmpl, orbit, device="cuda"):
        super().__init__()
        if tmpl.fxaa:
            raise ValueError("the differentiable render works on the SSAA image: fxaa must be 0")
        self.ctx, self.tmpl, self.orbit, self.device = ctx, tmpl, orbit, device

    def forward(self, params):
        return _KerrOrbitFn.apply(params, self)


def default_orbit(spin, r_cam, incl, sphere_phi, sphere_r=8.0, sphere_radius=1.0, sphere_tex=1, cam_phi=0.0):
    o = A.Orbit()
    o.spin, o.r_cam, o.incl, o.sphere_phi = spin, r_cam, incl, sphere_phi
    o.cam_phi, o.sphere_r, o.sphere_radius, o.sphere_tex, o.has_sphere = cam_phi, sphere_r, sphere_radius, sphere_tex, 1
    return o
