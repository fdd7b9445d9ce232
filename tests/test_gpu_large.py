"""A field whose padded FP32 texels span more than 2 GiB (1024x1024x130,
2.22 GB): the LIC kernel addresses it with 32-bit byte offsets, so voxels
past the 2 GiB mark must come out exactly as the oracle's.  The ray cast of
the same volume is checked as well.  Boxes at both ends of the volume are
compared (the oracle computes only those voxels)."""
import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from oracle import oracle as orc

from conftest import rel_err

pytestmark = pytest.mark.gpu

DIMS = (1024, 1024, 130)
F = B = 4


def _field_torch(dims):
    nx, ny, nz = dims
    z = (torch.arange(nz, device="cuda", dtype=torch.float32).view(nz, 1, 1) + 0.5) / nz
    y = (torch.arange(ny, device="cuda", dtype=torch.float32).view(1, ny, 1) + 0.5) / ny
    x = (torch.arange(nx, device="cuda", dtype=torch.float32).view(1, 1, nx) + 0.5) / nx
    v = torch.empty((nz, ny, nx, 3), dtype=torch.float32, device="cuda")
    tau = 2.0 * np.pi
    v[..., 0] = torch.sin(tau * 3.0 * z) + 0.7 * torch.cos(tau * 5.0 * y)
    v[..., 1] = torch.sin(tau * 4.0 * x) + 0.5 * torch.cos(tau * 2.0 * z)
    v[..., 2] = torch.sin(tau * 2.0 * y) + 0.9 * torch.cos(tau * 3.0 * x)
    return v


def test_lic_and_render_past_2gib(ctx):
    assert (DIMS[0] + 2) * (DIMS[1] + 2) * (DIMS[2] + 2) * 16 > 2 ** 31
    v = _field_torch(DIMS)
    ctx.upload_field(v, DIMS)
    ctx.generate_noise(5, vv.NOISE_UNIFORM, 0.0, DIMS)
    p = vv.default_lic_params()
    p.steps_forward, p.steps_backward = F, B
    p.step_size_lic = 0.01
    ctx.set_lic_params(p)
    assert ctx.get_tuning("lic.cache") == 1
    ctx.compute_lic()
    n = DIMS[0] * DIMS[1] * DIMS[2]
    lic = torch.empty(n, dtype=torch.float32, device="cuda")
    ctx.download_lic(lic)
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(96, 96, cen)
    rgba, op = vv.fields.tf_linear_ramp()
    ctx.set_tf(rgba, op)
    img, nsamp = ctx.render(cam, count_samples=True)
    torch.cuda.synchronize()

    host = v.cpu().numpy()
    del v
    f4 = orc.encode_field(host, DIMS)
    del host
    noise = orc.gen_noise(5, vv.NOISE_UNIFORM, 0.0, DIMS)
    op_ = orc.params(steps_forward=F, steps_backward=B, step_size_lic=0.01)
    lut = np.full(256, 255, np.uint8)
    lic3 = lic.view(DIMS[2], DIMS[1], DIMS[0])
    for box in ((0, 16, 0, 16, 0, 8), (1008, 1024, 1000, 1024, 122, 130)):
        print(f"box {box}", flush=True)
        ref = orc.lic_box(f4, DIMS, noise, 1, DIMS, lut, op_, box, scale=sc)
        got = lic3[box[4]:box[5], box[2]:box[3], box[0]:box[1]].cpu().numpy().ravel()
        assert np.any(ref != 0)
        assert rel_err(got, ref) <= 1e-5

    licv = lic.cpu().numpy()
    view = orc.view_setup(96, 96, center=cen, extent=ext, scale=sc)
    op_.light[0], op_.light[1], op_.light[2] = list(cam.light_pos)
    ref, ns, total = orc.raycast(f4, DIMS, licv, 1, DIMS, rgba, op, op_, view)
    assert np.max(ref) > 0.01
    assert rel_err(img, ref) <= 1e-5
    assert nsamp == total
