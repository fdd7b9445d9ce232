"""Batched LIC + ray casting (vv_batch): timesteps of different sizes and
views of different sizes in one captured launch.

Every LIC volume and every image of a batch must be bit-identical to the
same timestep / view computed on its own through vv_compute_lic and
vv_render (which the parity tests tie to the oracle); one timestep is also
checked against the oracle directly.  Replays of the captured graph, a
re-run after autotuning, and FP16 storage are covered as well.
"""
import numpy as np
import pytest
import torch

import vectorvisualization_amd as vv
from vectorvisualization_amd import fields
from oracle import oracle as orc

from conftest import rel_err

pytestmark = pytest.mark.gpu

L = 12
STEPS = [  # dims, spacing, noise seed
    ((24, 24, 24), (1.0, 1.0, 1.0), 11),
    ((32, 20, 16), (1.0, 1.0, 2.0), 12),
    ((40, 40, 40), (1.0, 1.0, 1.0), 13),
]
VIEWS = [  # timestep, width, height, yaw, pitch
    (0, 64, 64, 0.0, 0.0),
    (1, 96, 48, 0.7, 0.2),
    (2, 80, 80, -1.1, 0.4),
    (0, 33, 57, 2.5, -0.3),
    (2, 128, 96, 3.0, 0.1),
    # views smaller than one pixel tile of the batch caster
    (1, 1, 1, 0.3, 0.1),
    (2, 3, 1, -0.4, 0.0),
    (0, 5, 2, 1.2, 0.2),
]


def _field(i, dims):
    rng = np.random.default_rng(100 + i)
    n = max(dims)
    v = fields.rankine_vortices(n, seed=i + 1).reshape(n, n, n, 3)[: dims[2], : dims[1], : dims[0]]
    v = v + 0.05 * rng.normal(size=v.shape)
    return np.ascontiguousarray(v, np.float32)


def _setup(ctx):
    lut, inv = vv.make_filter(vv.FILTER_TRIANGLE, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = L
    p.step_size_lic = 0.01
    p.step_size_vol = 1.0 / 128.0
    p.gradient_scale = 3.0
    p.tf_index = vv.TF_INDEX_MAG
    ctx.set_lic_params(p)
    rgba, op = fields.tf_magnitude(0.05, 0.3)
    ctx.set_tf(rgba, op)
    return lut, inv, p, rgba, op


def _single(ctx, storage):
    """Each timestep and view computed on its own (vv_compute_lic + vv_render)."""
    lics, centers, imgs, ns = [], [], [], []
    for i, (dims, sp, seed) in enumerate(STEPS):
        ctx.upload_field(_field(i, dims), dims, spacing=sp, storage=storage)
        ctx.generate_noise(seed, vv.NOISE_UNIFORM, 0.0, dims)
        ctx.compute_lic()
        lics.append(ctx.download_lic())
        centers.append(ctx.field_geometry()[2])
        for t, w, h, yaw, pitch in VIEWS:
            if t != i:
                continue
            img, n = ctx.render(fields.orbit_camera(w, h, centers[i], yaw, pitch), count_samples=True)
            imgs.append((VIEWS.index((t, w, h, yaw, pitch)), img))
            ns.append(n)
    imgs = [im for _, im in sorted(imgs, key=lambda x: x[0])]
    return lics, centers, imgs, sum(ns)


def _batch(ctx, storage, centers):
    b = vv.Batch(ctx)
    for i, (dims, sp, seed) in enumerate(STEPS):
        b.add_timestep(_field(i, dims), dims, spacing=sp, storage=storage, noise_seed=seed,
                       noise_kind=vv.NOISE_UNIFORM)
    outs = []
    for t, w, h, yaw, pitch in VIEWS:
        img = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
        b.add_view(t, fields.orbit_camera(w, h, centers[t], yaw, pitch), img)
        outs.append(img)
    return b, outs


def _check(b, outs, lics, imgs, total):
    for i, (dims, _, _) in enumerate(STEPS):
        np.testing.assert_array_equal(b.download_lic(i, dims), lics[i])
    for v, img in enumerate(outs):
        np.testing.assert_array_equal(img.cpu().numpy(), imgs[v])
    assert b.samples() == total


@pytest.mark.parametrize("storage", [vv.STORE_FP32, vv.STORE_FP16, vv.STORE_UNORM8])
def test_batch_matches_single(ctx, storage):
    _setup(ctx)
    lics, centers, imgs, total = _single(ctx, storage)
    assert all(np.max(im) > 0.01 for im in imgs)
    b, outs = _batch(ctx, storage, centers)
    assert b.info() == (len(STEPS), len(VIEWS))
    b.run()
    torch.cuda.synchronize()
    _check(b, outs, lics, imgs, total)
    # replays of the captured graph give the same results
    for o in outs:
        o.zero_()
    for _ in range(3):
        b.run()
    _check(b, outs, lics, imgs, total)
    lat = b.view_latency_ms()
    span = b.view_span_ms()
    print("view latency ms:", np.round(lat, 4), "span ms:", np.round(span, 4), flush=True)
    assert lat.shape == (len(VIEWS),) and span.shape == (len(VIEWS),)
    assert np.all(lat > 0) and np.all(lat < 1e4)
    # a view's own render span lies inside the batch, before its end stamp
    assert np.all(span > 0) and np.all(span <= lat)
    b.close()


def test_batch_timestep_against_oracle(ctx):
    lut, inv, p, rgba, op = _setup(ctx)
    dims, sp, seed = STEPS[1]
    b = vv.Batch(ctx)
    v = _field(1, dims)
    b.add_timestep(v, dims, spacing=sp, noise_seed=seed, noise_kind=vv.NOISE_UNIFORM)
    ctx.upload_field(v, dims, spacing=sp)
    ext, sc, cen = ctx.field_geometry()
    cam = fields.orbit_camera(48, 40, cen, 0.7, 0.2)
    img = torch.zeros(48 * 40 * 4, dtype=torch.float32, device="cuda")
    b.add_view(0, cam, img)
    b.run()
    got = b.download_lic(0, dims)
    f4 = orc.encode_field(v, dims)
    noise = orc.gen_noise(seed, vv.NOISE_UNIFORM, 0.0, dims)
    opp = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.01, inv_filter_area=inv,
                     gradient_scale=3.0, step_size_vol=1.0 / 128.0, tf_index=1)
    ref = orc.lic(f4, dims, noise, 1, dims, lut, opp, scale=sc)
    assert rel_err(got, ref) <= 1e-5
    view = orc.view_setup(48, 40, quat=[float(q) for q in cam.quat], center=cen, extent=ext, scale=sc)
    np.testing.assert_array_equal(ctx.view_block(cam), view.as_array())
    opp.alpha_correction = float(np.float32(1.0 / 128.0) * np.float32(128.0))
    opp.light[0], opp.light[1], opp.light[2] = list(cam.light_pos)
    rimg, _, _ = orc.raycast(f4, dims, ref, 1, dims, rgba, op, opp, view)
    assert np.max(rimg) > 0.01
    assert rel_err(img.cpu().numpy(), rimg) <= 1e-5
    b.close()


def test_batch_autotune_keeps_results(ctx):
    _setup(ctx)
    lics, centers, imgs, total = _single(ctx, vv.STORE_FP32)
    b, outs = _batch(ctx, vv.STORE_FP32, centers)
    b.autotune()
    for i in range(len(STEPS)):
        tu = b.tuning(i)
        print(f"timestep {i}: tile {tu[:3]} brick tiles {tu[3:6]} render tile {tu[6:]}", flush=True)
        assert tu[0] * tu[1] * tu[2] == 256
        assert tu[6] * tu[7] % 64 == 0
    b.run()
    _check(b, outs, lics, imgs, total)
    b.close()


def test_batch_errors(ctx):
    _setup(ctx)
    b = vv.Batch(ctx)
    with pytest.raises(vv.VVError):
        b.run()  # no timesteps
    b.add_timestep(_field(0, (8, 8, 8)), (8, 8, 8))
    with pytest.raises(vv.VVError):
        b.add_timestep(_field(0, (8, 8, 8)), (8, 8, 8), storage=vv.STORE_FP16)  # mixed storage
    with pytest.raises(vv.VVError):
        b.add_view(3, vv.default_camera(8, 8), torch.zeros(256, device="cuda"))
    with pytest.raises(ValueError):
        b.add_view(0, vv.default_camera(8, 8), np.zeros(256, np.float32))
    b.close()


def test_batch_line_illumination_matches_single(ctx):
    _setup(ctx)
    p = ctx.get_lic_params()
    p.illum_mode = vv.ILLUM_ZOECKLER
    ctx.set_lic_params(p)
    lics, centers, imgs, total = _single(ctx, vv.STORE_FP32)
    b, outs = _batch(ctx, vv.STORE_FP32, centers)
    b.run()
    _check(b, outs, lics, imgs, total)
    b.close()


def test_batch_views_finish_with_their_timestep(ctx):
    # time steps run smallest first, each followed by its views: the views of
    # the small time step finish before those of the large one
    _setup(ctx)
    b = vv.Batch(ctx)
    b.add_timestep(_field(2, (48, 48, 48)), (48, 48, 48), noise_seed=3)
    b.add_timestep(_field(0, (16, 16, 16)), (16, 16, 16), noise_seed=4)
    outs = []
    for t in (0, 1, 0, 1):
        img = torch.zeros(64 * 64 * 4, dtype=torch.float32, device="cuda")
        b.add_view(t, fields.orbit_camera(64, 64, (0.5, 0.5, 0.5), 0.3 * t, 0.1), img)
        outs.append(img)
    b.run()
    lat = b.view_latency_ms()
    assert max(lat[1], lat[3]) < min(lat[0], lat[2])
    with pytest.raises(vv.VVError):
        b.add_view(0, fields.orbit_camera(8, 8, (0.5, 0.5, 0.5), 0.0), torch.zeros(256, device="cuda"))
        b.view_latency_ms()  # the batch changed since its last run
    b.close()


def test_batch_runs_follow_changed_settings(ctx):
    """A run whose descriptors equal the previous run's only enqueues the
    graph.  A new transfer function (same device table, same descriptors)
    and new LIC parameters (new descriptors, uploaded) must both show in the
    next run, equal to single calls with the same settings."""
    _setup(ctx)
    lics, centers, imgs, total = _single(ctx, vv.STORE_FP32)
    b, outs = _batch(ctx, vv.STORE_FP32, centers)
    b.run()
    b.run()
    _check(b, outs, lics, imgs, total)
    rgba, op = fields.tf_magnitude(0.1, 0.5)
    ctx.set_tf(rgba, op)
    lics2, _, imgs2, total2 = _single(ctx, vv.STORE_FP32)
    assert any(not np.array_equal(a, c) for a, c in zip(imgs, imgs2))
    b.run()
    _check(b, outs, lics2, imgs2, total2)
    p = ctx.get_lic_params()
    p.step_size_lic = 0.007
    ctx.set_lic_params(p)
    lics3, _, imgs3, total3 = _single(ctx, vv.STORE_FP32)
    assert not np.array_equal(lics3[0], lics2[0])
    b.run()
    b.run()
    _check(b, outs, lics3, imgs3, total3)
    b.close()
