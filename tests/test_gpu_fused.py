"""Fused on-the-fly LIC ray caster (lic3d_fragment.glsl) on the GPU against
the oracle's fused caster (which the CPU tests tie to the LIC-volume caster
and to its own termination rule).  Tolerance: 1e-5 relative in FP32, sample
counts exact."""
import numpy as np
import pytest

import vectorvisualization_amd as vv
from oracle import oracle as orc

from conftest import rel_err
from workloads import lic_setup

pytestmark = pytest.mark.gpu


def _fused_case(ctx, c, w, h, quat=(0.15, 0.25, 0.05, 0.95), illum=None, luts=None, tf_index=vv.TF_INDEX_MAG,
                sigma=None, step=1.0 / 64.0):
    p = c["params"]
    p.tf_index = tf_index
    p.step_size_vol = step
    if illum is not None:
        p.illum_mode = illum
    ctx.set_lic_params(p)
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(w, h, cen)
    for i in range(4):
        cam.quat[i] = quat[i]
    rgba, op = vv.fields.tf_magnitude(0.05, 0.4)
    ctx.set_tf(rgba, op)
    if sigma is not None:
        ctx.set_opacity_scale(sigma)
    img, nsamp = ctx.render_fused(cam, count_samples=True)
    view = orc.view_setup(w, h, quat=[float(np.float32(q)) for q in quat], center=cen, extent=ext, scale=sc)
    op_ = c["op"]
    op_.tf_index = tf_index
    op_.step_size_vol = step
    op_.alpha_correction = float(np.float32(step) * np.float32(128.0))
    op_.light[0], op_.light[1], op_.light[2] = list(cam.light_pos)
    if illum is not None:
        op_.illum_mode = illum
    if luts is not None:
        orc.attach_luts(op_, luts)
    ref, ns, total = orc.raycast_fused(c["f4"], c["dims"], c["noise"], c["nch"], c["dims"], c["lut"], rgba, op,
                                       op_, view, scalar=c["scalar"],
                                       sdims=c["dims"] if c["scalar"] is not None else None,
                                       sigma=sigma, sigma_dims=c["dims"] if sigma is not None else None)
    print(f"fused: rel_err {rel_err(img, ref):.2e} samples {nsamp}", flush=True)
    return img, ref, nsamp, total


@pytest.mark.parametrize("storage", [vv.STORE_FP32, vv.STORE_FP16, vv.STORE_UNORM8])
def test_fused_cached_path(ctx, storage):
    c = lic_setup(ctx, n=16, kind="rankine", L=6, noise_kind=vv.NOISE_UNIFORM, storage=storage,
                  gradient_scale=3.0)
    img, ref, nsamp, total = _fused_case(ctx, c, 40, 32)
    assert nsamp == total and np.max(ref) > 0.01
    assert rel_err(img, ref) <= 1e-5


def test_fused_generic_paths(ctx):
    # gate (scalar volume) and freq != 1 take the generic per-sample LIC
    n = 14
    gate = np.random.default_rng(2).integers(0, 256, n ** 3, dtype=np.uint8)
    c = lic_setup(ctx, n=n, kind="abc", L=5, freq=1.7, gate=gate, gradient_scale=3.0)
    img, ref, nsamp, total = _fused_case(ctx, c, 36, 30)
    assert nsamp == total
    assert rel_err(img, ref) <= 1e-5


@pytest.mark.parametrize("illum", [vv.ILLUM_GRADIENT, vv.ILLUM_MALLO, vv.ILLUM_ZOECKLER])
def test_fused_illumination(ctx, illum):
    grad = illum == vv.ILLUM_GRADIENT
    c = lic_setup(ctx, n=14, kind="rankine", L=5, grad=grad, density=0.3, gradient_scale=3.0)
    luts = orc.illum_luts()
    img, ref, nsamp, total = _fused_case(ctx, c, 32, 28, illum=illum, luts=luts)
    assert nsamp == total and np.max(ref) > 0.01
    assert rel_err(img, ref) <= 1e-5


def test_fused_with_opacity_scale(ctx):
    n = 12
    c = lic_setup(ctx, n=n, kind="rankine", L=4, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0)
    ctx.compute_lic()  # the opacity scale lives on the LIC grid
    sigma = np.random.default_rng(5).uniform(0.3, 2.5, n ** 3).astype(np.float32)
    img, ref, nsamp, total = _fused_case(ctx, c, 30, 30, sigma=sigma)
    assert nsamp == total
    assert rel_err(img, ref) <= 1e-5


def test_fused_needs_noise(ctx):
    ctx.upload_field(np.ones((8, 8, 8, 3), np.float32), (8, 8, 8))
    with pytest.raises(vv.VVError):
        ctx.render_fused(vv.default_camera(8, 8))


# ---- view-aligned slicing ---------------------------------------------------------
def _slices_case(ctx, c, w, h, quat=(0.1, 0.3, -0.05, 0.95), illum=None, luts=None, step=1.0 / 64.0):
    p = c["params"]
    p.tf_index = vv.TF_INDEX_MAG
    p.step_size_vol = step
    if illum is not None:
        p.illum_mode = illum
    ctx.set_lic_params(p)
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(w, h, cen)
    for i in range(4):
        cam.quat[i] = quat[i]
    rgba, op = vv.fields.tf_magnitude(0.02, 0.5)
    ctx.set_tf(rgba, op)
    img, nsamp = ctx.render_slices(cam, count_samples=True)
    view = orc.view_setup(w, h, quat=[float(np.float32(q)) for q in quat], center=cen, extent=ext, scale=sc)
    op_ = c["op"]
    op_.tf_index = vv.TF_INDEX_MAG
    op_.step_size_vol = step
    op_.alpha_correction = float(np.float32(step) * np.float32(128.0))
    op_.light[0], op_.light[1], op_.light[2] = list(cam.light_pos)
    if illum is not None:
        op_.illum_mode = illum
    if luts is not None:
        orc.attach_luts(op_, luts)
    ref, ns, total = orc.render_slices(c["f4"], c["dims"], c["noise"], c["nch"], c["dims"], c["lut"], rgba, op,
                                       op_, view, scalar=c["scalar"],
                                       sdims=c["dims"] if c["scalar"] is not None else None)
    print(f"slices: rel_err {rel_err(img, ref):.2e} samples {nsamp}", flush=True)
    return img, ref, nsamp, total


@pytest.mark.parametrize("storage", [vv.STORE_FP32, vv.STORE_FP16, vv.STORE_UNORM8])
def test_slices_cached_path(ctx, storage):
    c = lic_setup(ctx, n=16, kind="rankine", L=6, noise_kind=vv.NOISE_UNIFORM, storage=storage,
                  gradient_scale=3.0)
    img, ref, nsamp, total = _slices_case(ctx, c, 40, 32)
    assert nsamp == total and total > 0 and np.max(ref) > 0.01
    assert rel_err(img, ref) <= 1e-5


def test_slices_generic_and_line_illumination(ctx):
    c = lic_setup(ctx, n=14, kind="abc", L=5, freq=1.5, gradient_scale=3.0)
    img, ref, nsamp, total = _slices_case(ctx, c, 36, 30, illum=vv.ILLUM_MALLO, luts=orc.illum_luts())
    assert nsamp == total and total > 0
    assert rel_err(img, ref) <= 1e-5


@pytest.mark.parametrize("slices", [False, True])
def test_ray_groups_give_identical_images(ctx, slices):
    # every lanes-per-ray setting must reproduce the one-lane caster bit for
    # bit, including rays that stop early (an opaque TF)
    c = lic_setup(ctx, n=16, kind="rankine", L=5, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0)
    p = c["params"]
    p.tf_index = vv.TF_INDEX_MAG
    p.step_size_vol = 1.0 / 96.0
    ctx.set_lic_params(p)
    cen = ctx.field_geometry()[2]
    cam = vv.default_camera(37, 29, cen)
    cam.quat[0], cam.quat[1], cam.quat[3] = 0.2, 0.1, 0.97
    for alpha_hi in (0.3, 1.0):
        ctx.set_tf(*vv.fields.tf_magnitude(0.1, alpha_hi))
        imgs = {}
        for G in (1, 2, 4, 8, 16, 32, 64):
            ctx.set_tuning("render.ray_group", G)
            render = ctx.render_slices if slices else ctx.render_fused
            imgs[G] = render(cam, count_samples=True)
        ctx.set_tuning("render.ray_group", 0)
        ref, nref = imgs[1]
        assert nref > 0
        for G, (img, n) in imgs.items():
            np.testing.assert_array_equal(img, ref, err_msg=f"G={G}")
            assert n == nref, (G, n, nref)


@pytest.mark.parametrize("slices", [False, True])
def test_ray_groups_identical_for_off_centre_views(ctx, slices):
    # panned views (the volume near the top or bottom edge, partly outside
    # the image, or behind the camera): the ray-group kernels' pixel tiles
    # must give the one-lane caster's image bit for bit
    c = lic_setup(ctx, n=16, kind="rankine", L=5, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0)
    p = c["params"]
    p.tf_index = vv.TF_INDEX_MAG
    p.step_size_vol = 1.0 / 96.0
    ctx.set_lic_params(p)
    cen = ctx.field_geometry()[2]
    ctx.set_tf(*vv.fields.tf_magnitude(0.1, 0.6))
    render = ctx.render_slices if slices else ctx.render_fused
    drawn = 0
    for pos, dist in (((0.0, 1.1, 0.0), 4.0), ((0.0, -1.1, 0.0), 4.0), ((0.3, 1.6, 0.0), 4.0),
                      ((0.0, -1.6, 0.0), 4.0), ((0.0, 0.0, 0.9), 0.3)):
        cam = vv.default_camera(41, 67, cen)
        cam.pos[0], cam.pos[1], cam.pos[2] = pos
        cam.dist = dist
        imgs = {}
        for G in (1, 2, 8):
            ctx.set_tuning("render.ray_group", G)
            imgs[G] = render(cam, count_samples=True)
        ctx.set_tuning("render.ray_group", 0)
        ref, nref = imgs[1]
        drawn += nref > 0
        for G, (img, n) in imgs.items():
            np.testing.assert_array_equal(img, ref, err_msg=f"pos={pos} G={G}")
            assert n == nref, (pos, G, n, nref)
    assert drawn == 4


@pytest.mark.parametrize("group", [1, 0])
def test_slices_gradient_illumination(ctx, group):
    c = lic_setup(ctx, n=14, kind="rankine", L=5, grad=True, density=0.3, gradient_scale=3.0)
    ctx.set_tuning("render.ray_group", group)
    img, ref, nsamp, total = _slices_case(ctx, c, 34, 30, illum=vv.ILLUM_GRADIENT)
    ctx.set_tuning("render.ray_group", 0)
    assert nsamp == total and total > 0
    assert np.all(np.isfinite(img))
    assert rel_err(img, ref) <= 1e-5


def _blend_case(ctx, c, w, h, quat=(0.1, 0.3, -0.05, 0.95), illum=None, luts=None, step=1.0 / 64.0):
    """vv_render_slices_blend against the oracle's blended slices (bytes)."""
    p = c["params"]
    p.step_size_vol = step
    if illum is not None:
        p.illum_mode = illum
    ctx.set_lic_params(p)
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(w, h, cen)
    for i in range(4):
        cam.quat[i] = quat[i]
    rgba, op = vv.fields.tf_magnitude(0.02, 0.5)
    ctx.set_tf(rgba, op)
    img, nsamp = ctx.render_slices_blend(cam, count_samples=True)
    view = orc.view_setup(w, h, quat=[float(np.float32(q)) for q in quat], center=cen, extent=ext, scale=sc)
    op_ = c["op"]
    op_.step_size_vol = step
    op_.alpha_correction = float(np.float32(step) * np.float32(128.0))
    op_.light[0], op_.light[1], op_.light[2] = list(cam.light_pos)
    if illum is not None:
        op_.illum_mode = illum
    if luts is not None:
        orc.attach_luts(op_, luts)
    ref, ns, total = orc.render_slices_blend(c["f4"], c["dims"], c["noise"], c["nch"], c["dims"], c["lut"], rgba,
                                             op, op_, view, scalar=c["scalar"],
                                             sdims=c["dims"] if c["scalar"] is not None else None)
    d = np.abs(img.astype(np.int32) - ref.astype(np.int32))
    print(f"blend: bytes differing {np.count_nonzero(d)} of {d.size}, max {d.max()}; samples {nsamp} "
          f"(oracle {total})", flush=True)
    return img, ref, d, nsamp, total


@pytest.mark.parametrize("storage,group", [(vv.STORE_FP32, 0), (vv.STORE_FP32, 1), (vv.STORE_FP16, 0),
                                           (vv.STORE_UNORM8, 0)])
def test_slices_blend_into_8bit_framebuffer(ctx, storage, group):
    """The reference's default slicing output: bytes equal to the oracle's.
    GPU and oracle LIC values agree to about 1e-6 relative, so a byte may
    differ by one step where a blend result lies within that of a rounding
    boundary; the bound allows one step on 0.1 % of the bytes."""
    c = lic_setup(ctx, n=16, kind="rankine", L=6, noise_kind=vv.NOISE_UNIFORM, storage=storage,
                  gradient_scale=3.0)
    ctx.set_tuning("render.ray_group", group)
    img, ref, d, nsamp, total = _blend_case(ctx, c, 40, 32)
    ctx.set_tuning("render.ray_group", 0)
    assert total > 0 and np.any(ref[3::4] == 255) and np.any(ref[0::4] < 250)
    assert d.max() <= 1 and np.count_nonzero(d) <= 0.001 * d.size
    assert nsamp == total or abs(nsamp - total) <= 0.001 * total


def test_slices_blend_line_illumination(ctx):
    c = lic_setup(ctx, n=14, kind="abc", L=5, freq=1.5, gradient_scale=3.0)
    img, ref, d, nsamp, total = _blend_case(ctx, c, 36, 30, illum=vv.ILLUM_MALLO, luts=orc.illum_luts())
    assert total > 0
    assert d.max() <= 1 and np.count_nonzero(d) <= 0.001 * d.size


@pytest.mark.parametrize("slices", [False, True])
def test_ray_groups_identical_for_tiny_views(ctx, slices):
    # views smaller than one pixel tile of the ray-group kernels: most lanes
    # of a block have no pixel
    c = lic_setup(ctx, n=12, kind="rankine", L=4, noise_kind=vv.NOISE_UNIFORM, gradient_scale=3.0)
    p = c["params"]
    p.tf_index = vv.TF_INDEX_MAG
    p.step_size_vol = 1.0 / 64.0
    ctx.set_lic_params(p)
    cen = ctx.field_geometry()[2]
    ctx.set_tf(*vv.fields.tf_magnitude(0.1, 0.6))
    render = ctx.render_slices if slices else ctx.render_fused
    for w, h in ((1, 1), (1, 7), (3, 1), (65, 3)):
        cam = vv.default_camera(w, h, cen)
        imgs = {}
        for G in (1, 2, 8, 64):
            ctx.set_tuning("render.ray_group", G)
            imgs[G] = render(cam, count_samples=True)
        ctx.set_tuning("render.ray_group", 0)
        ref, nref = imgs[1]
        assert nref > 0, (w, h)
        for G, (img, n) in imgs.items():
            np.testing.assert_array_equal(img, ref, err_msg=f"{w}x{h} G={G}")
            assert n == nref, (w, h, G, n, nref)
