"""GPU adjoint of the ray caster: against the FP64 oracle adjoint (which the
CPU tests pin by central differences) within 2e-3 of the largest gradient
entry, determinism, and the torch.autograd
shim against central finite differences (h = 1e-3) on 32 TF entries plus an
Adam fit of the transfer function."""
import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from vectorvisualization_amd import fields
from oracle import oracle as orc

from workloads import lic_setup

pytestmark = pytest.mark.gpu


_LUTS = None


def _scene(ctx, n=16, grad=False, step=1.0 / 128.0, W=48, H=40, quat=(0.2, -0.1, 0.05, 0.97),
           sigma_seed=None, tf_seed=3, op_scale=1.0, illum=None, storage=vv.STORE_FP32):
    global _LUTS
    c = lic_setup(ctx, n=n, kind="rankine", grad=grad, L=6, density=0.3, storage=storage)
    p = c["params"]
    p.step_size_vol = step
    if illum is not None:
        p.illum_mode = illum
        c["op"].illum_mode = illum
        if _LUTS is None:
            _LUTS = orc.illum_luts()
        orc.attach_luts(c["op"], _LUTS)
    ctx.set_lic_params(p)
    ctx.compute_lic()
    ext, sc, cen = ctx.field_geometry()
    cam = vv.default_camera(W, H, cen)
    for q in range(4):
        cam.quat[q] = quat[q]
    rgba, op = fields.tf_random(tf_seed)
    op = (op * op_scale).astype(np.float32)
    ctx.set_tf(rgba, op)
    sigma = None
    if sigma_seed is not None:
        sigma = np.random.default_rng(sigma_seed).uniform(0.3, 1.2, n ** 3).astype(np.float32)
        ctx.set_opacity_scale(sigma)
    oparams = c["op"]
    oparams.step_size_vol = step
    oparams.alpha_correction = float(np.float32(step) * np.float32(128.0))
    oparams.light[0], oparams.light[1], oparams.light[2] = list(cam.light_pos)
    view = orc.view_setup(W, H, quat=[float(np.float32(x)) for x in quat], center=cen, extent=ext,
                          scale=sc)
    return c, cam, rgba, op, sigma, oparams, view


@pytest.mark.parametrize("grad,step,with_sigma,illum,storage", [
    (False, 1.0 / 128.0, True, None, vv.STORE_FP32),
    (True, 1.0 / 256.0, True, None, vv.STORE_FP32),
    (False, 1.0 / 256.0, False, None, vv.STORE_FP32),
    (False, 1.0 / 128.0, True, vv.ILLUM_MALLO, vv.STORE_FP32),
    (True, 1.0 / 256.0, True, vv.ILLUM_ZOECKLER, vv.STORE_FP32),
    (False, 1.0 / 128.0, True, None, vv.STORE_UNORM8)])
def test_adjoint_matches_oracle(ctx, grad, step, with_sigma, illum, storage):
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, grad=grad, step=step,
                                                    sigma_seed=7 if with_sigma else None, illum=illum,
                                                    storage=storage)
    W, H = cam.width, cam.height
    dimg = np.random.default_rng(11).normal(size=W * H * 4).astype(np.float32)
    dtf, dsig = ctx.render_adjoint(cam, dimg)
    ldims, lch = ctx.lic_info()
    licv = ctx.download_lic()
    rdtf, rdsig = orc.raycast_adjoint(c["f4"].astype(np.float64), c["dims"], licv, lch, ldims, rgba,
                                      op, oparams, view, dimg, sigma=sigma)
    e_tf = np.max(np.abs(dtf - rdtf)) / np.max(np.abs(rdtf))
    e_sg = np.max(np.abs(dsig - rdsig)) / np.max(np.abs(rdsig))
    print(f"adjoint rel err: dTF {e_tf:.2e}  dsigma {e_sg:.2e}", flush=True)
    assert np.max(np.abs(rdtf)) > 0 and np.max(np.abs(rdsig)) > 0
    # FP32 reverse sweep against the FP64 oracle: the rounding of the G
    # recurrence grows with the number of samples per ray (~1e-4 at 1/128,
    # ~1e-3 at 1/256); bound 2e-3 of the largest entry
    assert e_tf <= 2e-3 and e_sg <= 2e-3


def test_adjoint_deterministic_and_window_invariant(ctx):
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, n=24, W=96, H=80, sigma_seed=2)
    dimg = np.random.default_rng(5).normal(size=96 * 80 * 4).astype(np.float32)
    a_tf, a_sg = ctx.render_adjoint(cam, dimg)
    b_tf, b_sg = ctx.render_adjoint(cam, dimg)
    ctx.set_tuning("adjoint.window", 0)
    c_tf, c_sg = ctx.render_adjoint(cam, dimg)
    ctx.set_tuning("adjoint.window", 1)
    np.testing.assert_array_equal(a_tf, b_tf)
    np.testing.assert_array_equal(a_sg, b_sg)
    # integer accumulation: LDS window and direct atomics give the same bits
    np.testing.assert_array_equal(a_tf, c_tf)
    np.testing.assert_array_equal(a_sg, c_sg)


def _autograd_scene(ctx):
    from vectorvisualization_amd.autograd import render_tf

    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, n=20, W=64, H=64, op_scale=0.05)
    target_tf = torch.tensor(fields.tf_random(3)[0].reshape(256, 4), device="cuda")
    with torch.no_grad():
        target = render_tf(ctx, cam, target_tf, op).clone()
    return render_tf, cam, op, target


def test_autograd_matches_central_differences(ctx):
    render_tf, cam, op, target = _autograd_scene(ctx)
    tf0 = torch.tensor(fields.tf_random(5)[0].reshape(256, 4), device="cuda", requires_grad=True)

    def loss_of(tf):
        img = render_tf(ctx, cam, tf, op)
        return ((img.double() - target.double()) ** 2).sum()

    L = loss_of(tf0)
    L.backward()
    g = tf0.grad.detach().cpu().numpy().ravel()
    rng = np.random.default_rng(0)
    cand = np.flatnonzero(np.abs(g) > 1e-3 * np.abs(g).max())
    idx = rng.choice(cand, 32, replace=False)
    h = 1e-3
    worst = 0.0
    with torch.no_grad():
        for k in idx:
            e = torch.zeros(1024, device="cuda")
            e[k] = h
            fp = loss_of(tf0.detach() + e.view(256, 4)).item()
            fm = loss_of(tf0.detach() - e.view(256, 4)).item()
            fd = (fp - fm) / (2 * h)
            err = abs(fd - g[k]) / max(abs(fd), 1e-3 * np.abs(g).max())
            worst = max(worst, err)
    print(f"autograd vs central FD (32 TF entries): worst rel err {worst:.2e}", flush=True)
    assert worst < 2e-2


def test_autograd_backward_uses_its_own_forward_state(ctx):
    """Two forward passes on one engine (different TFs, one with an opacity
    scale), then one backward: each gradient equals that of a forward and
    backward pair on its own (the adjoint sums are deterministic, so equal)."""
    render_tf, cam, op, target = _autograd_scene(ctx)
    d, _ = ctx.lic_info()
    sig0 = torch.tensor(np.random.default_rng(4).uniform(0.5, 1.2, int(np.prod(d))).astype(np.float32),
                        device="cuda")

    def fresh():
        tf_a = torch.tensor(fields.tf_random(5)[0].reshape(256, 4), device="cuda", requires_grad=True)
        tf_b = torch.tensor(fields.tf_random(7)[0].reshape(256, 4), device="cuda", requires_grad=True)
        return tf_a, tf_b, sig0.clone().requires_grad_(True)

    def loss(tf, sig=None):
        return ((render_tf(ctx, cam, tf, op, sig) - target) ** 2).sum()

    tf_a, tf_b, sig = fresh()
    loss(tf_a, sig).backward()
    loss(tf_b).backward()
    ref = (tf_a.grad.clone(), sig.grad.clone(), tf_b.grad.clone())
    tf_a, tf_b, sig = fresh()
    (loss(tf_a, sig) + loss(tf_b)).backward()
    assert torch.equal(tf_a.grad, ref[0])
    assert torch.equal(sig.grad, ref[1])
    assert torch.equal(tf_b.grad, ref[2])


def test_adam_fits_transfer_function(ctx):
    render_tf, cam, op, target = _autograd_scene(ctx)
    tf = torch.tensor(fields.tf_random(5)[0].reshape(256, 4), device="cuda", requires_grad=True)
    opt = torch.optim.Adam([tf], lr=0.02)
    losses = []
    for _ in range(200):
        opt.zero_grad()
        loss = ((render_tf(ctx, cam, tf, op) - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tf.clamp_(0.0, 1.0)
        losses.append(loss.item())
    print(f"Adam 200 steps: loss {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < 0.05 * losses[0]


@pytest.mark.parametrize("factor", [1.0 / (48 * 40 * 4), 1e6])
def test_adjoint_of_mean_and_scaled_losses_matches_oracle(ctx, factor):
    """dL/dimage of a .mean() loss (entries ~1e-4 here, ~1e-7 at 1024^2) and
    of a loss scaled by 1e6: the fixed-point scale follows the gradient's
    magnitude, so both keep the accuracy of the unscaled case."""
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, sigma_seed=7)
    W, H = cam.width, cam.height
    dimg = (np.random.default_rng(11).normal(size=W * H * 4) * factor).astype(np.float32)
    dtf, dsig = ctx.render_adjoint(cam, dimg)
    ldims, lch = ctx.lic_info()
    rdtf, rdsig = orc.raycast_adjoint(c["f4"].astype(np.float64), c["dims"], ctx.download_lic(), lch, ldims,
                                      rgba, op, oparams, view, dimg, sigma=sigma)
    e_tf = np.max(np.abs(dtf - rdtf)) / np.max(np.abs(rdtf))
    e_sg = np.max(np.abs(dsig - rdsig)) / np.max(np.abs(rdsig))
    print(f"factor {factor:.3g}: dTF {e_tf:.2e}  dsigma {e_sg:.2e}", flush=True)
    assert np.max(np.abs(rdtf)) > 0 and np.max(np.abs(rdsig)) > 0
    assert e_tf <= 2e-3 and e_sg <= 2e-3


def test_adjoint_power_of_two_scaling_is_exact(ctx):
    """Multiplying dL/dimage by 2^e multiplies every gradient by exactly 2^e:
    the scale exponent moves by -e and the integer sums do not change."""
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, n=20, W=64, H=56, sigma_seed=3)
    dimg = np.random.default_rng(2).normal(size=64 * 56 * 4).astype(np.float32)
    base_tf, base_sg = ctx.render_adjoint(cam, dimg)
    assert np.max(np.abs(base_tf)) > 0
    for e in (20, -40, 60):
        s = np.float32(2.0 ** e)
        tf_e, sg_e = ctx.render_adjoint(cam, dimg * s)
        np.testing.assert_array_equal(tf_e, base_tf * s)
        np.testing.assert_array_equal(sg_e, base_sg * s)


def test_adjoint_of_zero_and_nonfinite_image_gradients(ctx):
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, sigma_seed=7)
    n = cam.width * cam.height * 4
    dtf, dsig = ctx.render_adjoint(cam, np.zeros(n, np.float32))
    assert not np.any(dtf) and not np.any(dsig)
    g = np.random.default_rng(1).normal(size=n).astype(np.float32)
    # a non-finite entry on a pixel the box covers makes every gradient NaN
    img = ctx.render(cam).reshape(-1, 4)
    hit = int(np.flatnonzero(img[:, 3] > 0)[0])
    g[4 * hit + 1] = np.inf
    dtf, dsig = ctx.render_adjoint(cam, g)
    assert np.all(np.isnan(dtf)) and np.all(np.isnan(dsig))


@pytest.mark.parametrize("grad,illum", [(False, None), (True, vv.ILLUM_ZOECKLER)])
def test_render_record_image_and_reused_forward_pass(ctx, grad, illum):
    """vv_render_record: its image and sample count are vv_render's bit for
    bit, and the adjoint that reuses its forward pass gives the bits of an
    adjoint that runs its own, with and without the opacity-scale gradient."""
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, n=24, W=96, H=80, sigma_seed=4, grad=grad,
                                                    step=1.0 / 256.0, illum=illum)
    dimg = np.random.default_rng(8).normal(size=96 * 80 * 4).astype(np.float32)
    ref_img, ref_ns = ctx.render(cam, count_samples=True)
    ref_tf, ref_sg = ctx.render_adjoint(cam, dimg)
    ref_tf_only, _ = ctx.render_adjoint(cam, dimg, want_sigma=False)
    img, ns = ctx.render_record(cam, count_samples=True)
    np.testing.assert_array_equal(img, ref_img)
    assert ns == ref_ns and ns > 0
    # per-ray sample counts of the record: the oracle's termination indices
    counts = ctx.record_ray_samples(cam)
    ldims, lch = ctx.lic_info()
    _, ref_counts, ref_total = orc.raycast(c["f4"], c["dims"], ctx.download_lic(), lch, ldims, rgba, op,
                                           oparams, view, sigma=sigma)
    assert int(counts.sum()) == ns == ref_total
    np.testing.assert_array_equal(counts, ref_counts)
    with pytest.raises(vv.VVError):
        ctx.record_ray_samples(vv.default_camera(96, 80, ctx.field_geometry()[2]))  # another camera
    tf, sg = ctx.render_adjoint(cam, dimg)
    np.testing.assert_array_equal(tf, ref_tf)
    np.testing.assert_array_equal(sg, ref_sg)
    # this view's record keeps a column per segment, which the adjoint leaves as it was
    np.testing.assert_array_equal(ctx.record_ray_samples(cam), counts)
    tf_again, _ = ctx.render_adjoint(cam, dimg)
    np.testing.assert_array_equal(tf_again, ref_tf)
    # device image, then the TF-only gradient from the record
    dev = torch.empty(96 * 80 * 4, dtype=torch.float32, device="cuda")
    ctx.render_record(cam, out=dev)
    ctx.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), ref_img)
    tf2, _ = ctx.render_adjoint(cam, dimg, want_sigma=False)
    np.testing.assert_array_equal(tf2, ref_tf_only)
    # a second adjoint after one that consumed the record runs its own forward pass
    tf3, sg3 = ctx.render_adjoint(cam, dimg)
    np.testing.assert_array_equal(tf3, ref_tf)
    np.testing.assert_array_equal(sg3, ref_sg)


def test_render_record_is_invalidated_by_scene_and_camera_changes(ctx):
    """After a record, a changed transfer function, opacity scale, LIC volume or
    camera makes the adjoint run its own forward pass: every result equals a
    plain adjoint of the new state (a stale record would give the old
    scene's gradients)."""
    c, cam, rgba, op, sigma, oparams, view = _scene(ctx, n=24, W=64, H=56, sigma_seed=4)
    dimg = np.random.default_rng(9).normal(size=64 * 56 * 4).astype(np.float32)
    rgba2, op2 = fields.tf_random(5)

    def plain():
        return ctx.render_adjoint(cam, dimg)

    def recorded(change):
        ctx.render_record(cam)
        change()
        return ctx.render_adjoint(cam, dimg)

    stale_tf, stale_sg = plain()
    cases = [
        ("tf", lambda: ctx.set_tf(rgba2, op2), lambda: ctx.set_tf(rgba, op)),
        ("sigma", lambda: ctx.set_opacity_scale(sigma * np.float32(0.5)), lambda: ctx.set_opacity_scale(sigma)),
    ]
    for name, change, restore in cases:
        tf, sg = recorded(change)
        want_tf, want_sg = plain()
        assert not np.array_equal(want_tf, stale_tf), name
        np.testing.assert_array_equal(tf, want_tf, err_msg=name)
        np.testing.assert_array_equal(sg, want_sg, err_msg=name)
        restore()
    # another camera: the record of the first does not apply, and the other
    # camera's forward pass replaces it (the first camera's adjoint then runs
    # its own again)
    cam2 = vv.default_camera(64, 56, ctx.field_geometry()[2])
    ctx.render_record(cam)
    tf, sg = ctx.render_adjoint(cam2, dimg)
    want_tf, want_sg = ctx.render_adjoint(cam2, dimg)
    assert not np.array_equal(want_tf, stale_tf)
    np.testing.assert_array_equal(tf, want_tf)
    np.testing.assert_array_equal(sg, want_sg)
    with pytest.raises(vv.VVError):
        ctx.record_ray_samples(cam)
    tf, sg = ctx.render_adjoint(cam, dimg)
    np.testing.assert_array_equal(tf, stale_tf)
    np.testing.assert_array_equal(sg, stale_sg)
    # unchanged TF tables set again keep the record
    ctx.render_record(cam)
    ctx.set_tf(rgba, op)
    tf, sg = ctx.render_adjoint(cam, dimg)
    np.testing.assert_array_equal(tf, stale_tf)
    np.testing.assert_array_equal(sg, stale_sg)
    # a new LIC volume (which also drops the opacity scale)
    ldims, lch = ctx.lic_info()
    lic = ctx.download_lic()
    ctx.set_opacity_scale(None)
    base_tf, _ = ctx.render_adjoint(cam, dimg, want_sigma=False)
    ctx.render_record(cam)
    ctx.upload_lic(lic * np.float32(0.75), lch, ldims)
    tf, _ = ctx.render_adjoint(cam, dimg, want_sigma=False)
    want_tf, _ = ctx.render_adjoint(cam, dimg, want_sigma=False)
    assert not np.array_equal(want_tf, base_tf)
    np.testing.assert_array_equal(tf, want_tf)
