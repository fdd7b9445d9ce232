"""torch.autograd.Function over the ray caster and its hand-written adjoint.

``render_tf(engine, cam, tf_rgba, opacity, sigma=None)`` returns the RGBA
image (height, width, 4) rendered by the engine's ray caster from its current
LIC volume and field.  Gradients flow to the 256x4 RGBA transfer function and
to the per-voxel opacity scale; both are computed by ``vv_render_adjoint``
(checkpointed, deterministic) on the GPU.

The backward pass sets the transfer function, opacity table and opacity scale
of its own forward pass again before the adjoint, so several renders may share
one engine before a single backward.  The field, the LIC volume and the LIC
parameters must stay as they were at the forward pass.

The forward pass renders with ``vv_render_record``: when nothing changed the
engine's scene in between, the adjoint reuses its checkpoints instead of
running the forward walk again (a changed transfer function or opacity scale
is detected by the library and only costs that walk).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import Camera, Context


class _RenderTF(torch.autograd.Function):
    @staticmethod
    def forward(fctx, tf_rgba, sigma, engine: Context, cam: Camera, opacity: np.ndarray):
        if not tf_rgba.is_cuda:
            raise ValueError("tf_rgba must be a CUDA tensor")
        rgba = tf_rgba.detach().to(torch.float32).contiguous().cpu().numpy().ravel()
        opacity = np.ascontiguousarray(opacity, np.float32).copy()
        sig = None if sigma is None else sigma.detach().to(torch.float32).contiguous().clone()
        _RenderTF._apply_state(engine, rgba, opacity, sig)
        img = torch.empty(cam.height * cam.width * 4, dtype=torch.float32, device=tf_rgba.device)
        engine.render_record(cam, out=img)
        engine.synchronize()
        fctx.engine = engine
        fctx.cam = cam
        fctx.state = (rgba, opacity, sig)
        fctx.sigma_serial = engine.sigma_serial
        fctx.has_sigma = sigma is not None
        fctx.sigma_shape = None if sigma is None else sigma.shape
        return img.view(cam.height, cam.width, 4)

    @staticmethod
    def _apply_state(engine: Context, rgba: np.ndarray, opacity: np.ndarray, sigma, serial=None) -> None:
        # set_tf with unchanged tables leaves the scene as it is; the opacity
        # scale is set again only if anything set it since (serial)
        engine.set_tf(rgba, opacity)
        if serial is None or engine.sigma_serial != serial:
            engine.set_opacity_scale(sigma)

    @staticmethod
    def backward(fctx, grad_img):
        engine, cam = fctx.engine, fctx.cam
        g = grad_img.detach().to(torch.float32).contiguous().view(-1)
        dtf = torch.empty(1024, dtype=torch.float32, device=g.device)
        dsig = None
        if fctx.has_sigma and fctx.needs_input_grad[1]:
            dsig = torch.empty(int(np.prod(fctx.sigma_shape)), dtype=torch.float32, device=g.device)
        torch.cuda.current_stream().synchronize()
        # the state of this forward pass (another forward may have run since)
        _RenderTF._apply_state(engine, *fctx.state, serial=fctx.sigma_serial)
        engine.render_adjoint(cam, g, dtf, dsig)
        engine.synchronize()
        dsig_out = None if dsig is None else dsig.view(fctx.sigma_shape)
        return dtf.view(256, 4), dsig_out, None, None, None


def render_tf(engine: Context, cam: Camera, tf_rgba: torch.Tensor, opacity: np.ndarray,
              sigma: torch.Tensor | None = None) -> torch.Tensor:
    """Differentiable ray cast: image (H, W, 4) as a function of the
    (256, 4) RGBA transfer function and the optional opacity scale."""
    return _RenderTF.apply(tf_rgba, sigma, engine, cam, opacity)
