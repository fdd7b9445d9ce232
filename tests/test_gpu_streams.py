"""Stream ordering between torch and the engine's own (non-blocking)
stream: a field produced by torch on a side stream is read only once
complete, and a LIC volume downloaded into a torch tensor is complete
before torch work queued after the call reads it."""
import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from vectorvisualization_amd import fields

pytestmark = pytest.mark.gpu


def _lic(ctx, v, n):
    ctx.upload_field(v, (n, n, n))
    ctx.compute_lic()
    out = torch.empty(n ** 3, dtype=torch.float32, device="cuda")
    ctx.download_lic(out)
    return out


def test_torch_side_stream_ordering(ctx):
    n = 128
    ctx.generate_noise(3, vv.NOISE_UNIFORM, 0.0, (n, n, n))
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = 8
    ctx.set_lic_params(p)
    ref_field = fields.rankine_vortices_torch(n, count=24, device="cuda")
    torch.cuda.synchronize()
    ref = _lic(ctx, ref_field, n).clone()
    torch.cuda.synchronize()

    s = torch.cuda.Stream()
    for _ in range(3):
        with torch.cuda.stream(s):
            v = fields.rankine_vortices_torch(n, count=24, device="cuda")  # queued on s
            out = _lic(ctx, v, n)
            got = out * 1.0  # on s, after the engine's work
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_render_into_tensor_then_torch_reads(ctx):
    n = 32
    ctx.upload_field(fields.abc_flow(n), (n, n, n))
    ctx.generate_noise(1, vv.NOISE_BINARY, 0.05, (n, n, n))
    ctx.compute_lic()
    cam = vv.default_camera(128, 128, ctx.field_geometry()[2])
    host = ctx.render(cam)
    img = torch.full((128 * 128 * 4,), -1.0, device="cuda")
    ctx.render(cam, out=img)
    np.testing.assert_array_equal(img.cpu().numpy(), host)  # .cpu() runs on torch's stream
