"""vv_render_adjoint_noalloc through the C ABI against vv_render_adjoint.

The allocation-free entry point runs the same kernels on a static device
array instead of allocated scratch, so its gradients must equal the
allocating entry point's bit for bit.  Two calls in a row, the second with a
smaller view (the array keeps the first call's contents), with and without
the opacity-scale gradient, in both scratch layouts (checkpoints with
segment re-runs for a view whose per-segment columns do not fit the array,
columns otherwise); a view too large for the array is refused with
VV_EINVAL before any kernel runs."""
import ctypes as C

import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from vectorvisualization_amd import _lib, fields

from workloads import lic_setup

pytestmark = pytest.mark.gpu


def _scene(ctx):
    lic_setup(ctx, n=20, kind="rankine", L=6, density=0.3)
    p = ctx.get_lic_params()
    p.step_size_vol = 1.0 / 256.0
    ctx.set_lic_params(p)
    ctx.compute_lic()
    rgba, op = fields.tf_random(3)
    ctx.set_tf(rgba, op)
    d, _ = ctx.lic_info()
    sigma = np.random.default_rng(4).uniform(0.3, 1.2, int(np.prod(d))).astype(np.float32)
    ctx.set_opacity_scale(sigma)
    return int(np.prod(d))


def _cam(ctx, w, h):
    cam = vv.default_camera(w, h, ctx.field_geometry()[2])
    for q, v in enumerate((0.2, -0.1, 0.05, 0.97)):
        cam.quat[q] = v
    return cam


def test_noalloc_adjoint_equals_allocating_adjoint(ctx):
    nv = _scene(ctx)
    # 320x288: the per-segment columns exceed the array, so this call keeps
    # checkpoints and re-runs segments (the allocating call keeps columns);
    # the smaller views fit the column layout
    for w, h, with_sigma in ((320, 288, True), (72, 60, True), (40, 28, True), (52, 44, False)):
        print(f"view {w}x{h}, opacity-scale gradient {with_sigma}", flush=True)
        cam = _cam(ctx, w, h)
        dimg = torch.randn(w * h * 4, dtype=torch.float32, device="cuda",
                           generator=torch.Generator(device="cuda").manual_seed(w))
        ref_tf = torch.empty(1024, dtype=torch.float32, device="cuda")
        ref_sg = torch.empty(nv, dtype=torch.float32, device="cuda") if with_sigma else None
        ctx.render_adjoint(cam, dimg, ref_tf, ref_sg)
        ctx.synchronize()
        out_tf = torch.full((1024,), float("nan"), dtype=torch.float32, device="cuda")
        out_sg = torch.full((nv,), float("nan"), dtype=torch.float32, device="cuda") if with_sigma else None
        torch.cuda.synchronize()
        st = vv.LIB.vv_render_adjoint_noalloc(ctx.handle, C.byref(cam), dimg.data_ptr(), out_tf.data_ptr(),
                                              out_sg.data_ptr() if with_sigma else None)
        assert st == _lib.VV_OK, vv.LIB.vv_last_error(ctx.handle)
        ctx.synchronize()
        assert torch.count_nonzero(ref_tf) > 0
        assert torch.equal(out_tf, ref_tf)
        if with_sigma:
            assert torch.count_nonzero(ref_sg) > 0
            assert torch.equal(out_sg, ref_sg)


def test_noalloc_adjoint_refuses_views_beyond_its_array(ctx):
    nv = _scene(ctx)
    cam = _cam(ctx, 1024, 1024)
    dimg = torch.zeros(1024 * 1024 * 4, dtype=torch.float32, device="cuda")
    dtf = torch.zeros(1024, dtype=torch.float32, device="cuda")
    dsg = torch.zeros(nv, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    st = vv.LIB.vv_render_adjoint_noalloc(ctx.handle, C.byref(cam), dimg.data_ptr(), dtf.data_ptr(),
                                          dsg.data_ptr())
    assert st == _lib.VV_EINVAL
    assert "scratch" in vv.LIB.vv_last_error(ctx.handle).decode()
    # the context stays usable
    small = _cam(ctx, 24, 20)
    g = torch.ones(24 * 20 * 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    assert vv.LIB.vv_render_adjoint_noalloc(ctx.handle, C.byref(small), g.data_ptr(), dtf.data_ptr(),
                                            None) == _lib.VV_OK
    ctx.synchronize()
    assert torch.count_nonzero(dtf) > 0
