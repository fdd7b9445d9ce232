"""SPEED_OF_FLOW (VV/shader/inc_lic.glsl:107-110,120-122; key '6' in
VV/3DLIC.cpp:429-432): each Heun displacement scaled by the field magnitude
|v|/max at the step's start point.  The LIC volume in every storage, the
fused caster, view-aligned slicing and a batch against the oracle (whose
SPEED_OF_FLOW is pinned by a closed form in tests/test_oracle.py), within
1e-5 relative with exact sample counts; the batch equal to single calls bit
for bit."""
import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from oracle import oracle as orc

from conftest import rel_err
from workloads import lic_setup, oracle_lic
from test_gpu_fused import _fused_case, _slices_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,grad", [(vv.STORE_FP32, False), (vv.STORE_FP16, False), (vv.STORE_UNORM8, False),
                                          (vv.STORE_FP32, True)])
def test_speed_of_flow_lic_volume(ctx, storage, grad):
    c = lic_setup(ctx, n=20, kind="rankine", L=8, h=0.02, noise_kind=vv.NOISE_UNIFORM, storage=storage,
                  gradient_scale=3.0, grad=grad, speed=True)
    ctx.compute_lic()
    lic = ctx.download_lic()
    ref = oracle_lic(c)
    c["op"].speed_of_flow = 0
    plain = oracle_lic(c)
    e = rel_err(lic, ref)
    print(f"speed of flow LIC: rel err {e:.2e}; change against plain LIC {rel_err(plain, ref):.2e}", flush=True)
    assert rel_err(plain, ref) > 1e-3  # the mode changes the result
    assert e <= 1e-5


def test_speed_of_flow_fused_and_slices(ctx):
    c = lic_setup(ctx, n=16, kind="rankine", L=6, h=0.02, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0,
                  speed=True)
    img, ref, nsamp, total = _fused_case(ctx, c, 40, 32)
    assert nsamp == total and np.max(ref) > 0.01
    assert rel_err(img, ref) <= 1e-5
    c = lic_setup(ctx, n=16, kind="rankine", L=6, h=0.02, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0,
                  speed=True)
    img, ref, nsamp, total = _slices_case(ctx, c, 40, 32)
    assert nsamp == total and total > 0
    assert rel_err(img, ref) <= 1e-5


def test_speed_of_flow_batch_matches_single(ctx):
    import test_gpu_batch as tb

    tb._setup(ctx)
    p = ctx.get_lic_params()
    p.speed_of_flow = 1
    ctx.set_lic_params(p)
    lics, centers, imgs, total = tb._single(ctx, vv.STORE_FP32)
    b, outs = tb._batch(ctx, vv.STORE_FP32, centers)
    b.run()
    torch.cuda.synchronize()
    tb._check(b, outs, lics, imgs, total)
    # a changed LIC step reaches the kernels captured with by-value arguments
    p.step_size_lic = 0.007
    ctx.set_lic_params(p)
    lics2, _, imgs2, total2 = tb._single(ctx, vv.STORE_FP32)
    assert not np.array_equal(lics2[0], lics[0])
    b.run()
    tb._check(b, outs, lics2, imgs2, total2)
    b.close()
