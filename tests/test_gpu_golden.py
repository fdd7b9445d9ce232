"""The HIP path against the committed golden fixtures (tests/golden/
oracle_golden.npz) without running the oracle: byte outputs bit-exact,
floating-point outputs within 1e-5 relative of the FP32 fixtures, the
adjoint within 2e-3 of the FP64 one."""
from pathlib import Path

import numpy as np
import pytest

import vectorvisualization_amd as vv
from vectorvisualization_amd import fields

from conftest import rel_err

pytestmark = pytest.mark.gpu

G = np.load(Path(__file__).resolve().parent / "golden" / "oracle_golden.npz")
INV255 = np.float32(1.0 / 255.0)


def test_golden_encodings(ctx):
    fd = (10, 9, 8)
    v0, v1 = G["enc_v0"], G["enc_v1"]
    ctx.upload_field(v0, fd, v1=v1, lerp_t=0.3)
    np.testing.assert_array_equal(ctx.download_field(fd), G["enc_f32"])
    ctx.upload_field(v0, fd, v1=v1, lerp_t=0.3, encoding=vv.ENCODE_RAW_MAGNITUDE)
    np.testing.assert_array_equal(ctx.download_field(fd), G["enc_raw_f32"])
    ctx.upload_field(v0, fd, v1=v1, lerp_t=0.3, storage=vv.STORE_UNORM8)
    np.testing.assert_array_equal(ctx.download_field(fd), G["enc_u8"].astype(np.float32) * INV255)


def test_golden_noise_and_gradients(ctx):
    nd = (16, 12, 10)
    ctx.generate_noise(7, vv.NOISE_UNIFORM, 0.0, nd)
    np.testing.assert_array_equal(ctx.download_noise(), G["noise_uniform"])
    ctx.generate_noise(7, vv.NOISE_BINARY, 0.05, nd, with_gradients=True)
    got = ctx.download_noise().reshape(-1, 4)
    np.testing.assert_array_equal(got[:, 3], G["noise_binary"])
    np.testing.assert_array_equal(got[:, :3].ravel(), G["noise_grad"])


@pytest.mark.parametrize("kind,name,L", [(vv.FILTER_TRIANGLE, "tri", 30), (vv.FILTER_COS2, "cos2", 60),
                                         (vv.FILTER_GAUSSIAN, "gauss", 40)])
def test_golden_filters(kind, name, L):
    lut, inv = vv.make_filter(kind, L)
    np.testing.assert_array_equal(lut, G[f"filter_{name}_lut"])
    assert np.float32(inv) == G[f"filter_{name}_inv"][0]


def _setup(ctx, grad=False, storage=vv.STORE_FP32):
    n = 12
    dims = (n, n, n)
    ctx.upload_field(G["lic_field"], dims, storage=storage)
    if grad:
        ctx.generate_noise(3, vv.NOISE_BINARY, 0.3, dims, with_gradients=True)
    else:
        ctx.generate_noise(3, vv.NOISE_UNIFORM, 0.0, dims)
    lut, inv = vv.make_filter(vv.FILTER_GAUSSIAN, 8)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = 8
    p.step_size_lic = 0.02
    p.gradient_scale = 3.0
    p.step_size_vol = 1.0 / 128.0
    p.illum_mode = vv.ILLUM_GRADIENT if grad else vv.ILLUM_LIC
    p.tf_index = vv.TF_INDEX_B if grad else vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    ctx.compute_lic()
    ctx.set_tf(*fields.tf_magnitude())
    cen = ctx.field_geometry()[2]
    cam = vv.default_camera(32, 32, cen)
    for i, q in enumerate((0.2, 0.3, 0.1, 0.927)):
        cam.quat[i] = q
    return cam


def test_golden_lic_and_images(ctx):
    cam = _setup(ctx)
    lic = ctx.download_lic()
    assert rel_err(lic, G["lic_f32"]) <= 1e-5
    assert rel_err(lic, G["lic_f64"]) <= 1e-4
    img, ns = ctx.render(cam, count_samples=True)
    assert ns == int(G["img_lic_ns"].sum())
    assert rel_err(img, G["img_lic"]) <= 1e-5
    cam = _setup(ctx, storage=vv.STORE_UNORM8)
    assert rel_err(ctx.download_lic(), G["lic_u8"]) <= 1e-5
    cam = _setup(ctx, grad=True)
    assert rel_err(ctx.download_lic(), G["lic4_f32"]) <= 1e-5
    img, ns = ctx.render(cam, count_samples=True)
    assert ns == int(G["img_grad_ns"].sum())
    assert rel_err(img, G["img_grad"]) <= 1e-5


def test_golden_adjoint(ctx):
    cam = _setup(ctx)
    dtf, dsig = ctx.render_adjoint(cam, G["adj_dimage"].astype(np.float32))
    e_tf = np.max(np.abs(dtf - G["adj_dtf"])) / np.max(np.abs(G["adj_dtf"]))
    e_sg = np.max(np.abs(dsig - G["adj_dsigma"])) / np.max(np.abs(G["adj_dsigma"]))
    assert e_tf <= 2e-3 and e_sg <= 2e-3
