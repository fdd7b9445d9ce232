"""The file loaders feeding the engine: the uploaded field, noise, gradients,
filter and gate volume must be what the oracle computes from the same file
contents."""
import numpy as np
import pytest
from PIL import Image

import vectorvisualization_amd as vv
from vectorvisualization_amd import _lib as L
from vectorvisualization_amd import fields
from oracle import oracle as orc

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _dat(path, raw, dims, fmt, extra=()):
    path.write_text("\n".join([f"ObjectFileName: {raw}", "Resolution: %d %d %d" % tuple(dims),
                               f"Format: {fmt}", *extra]) + "\n")


def test_load_field_dat_float_and_time_lerp(ctx, tmp_path):
    dims = (12, 10, 8)
    rng = np.random.default_rng(4)
    v = [rng.normal(size=(8, 10, 12, 3)).astype(np.float32) for _ in range(3)]
    for t in range(3):
        (tmp_path / f"v{t}.raw").write_bytes(v[t].tobytes())
    _dat(tmp_path / "v.dat", "v%d.raw", dims, "FLOAT3", ["TimeDependent: 0 2", "SliceThickness: 1 1 2"])
    ctx.load_field_dat(tmp_path / "v.dat", 1)
    np.testing.assert_array_equal(ctx.download_field(dims), orc.encode_field(v[1], dims))
    ext, sc, cen = ctx.field_geometry()
    assert ext[2] == 1.0 and ext[0] == pytest.approx(12 / 16)  # z spans 8 x 2 = 16
    ctx.load_field_dat(tmp_path / "v.dat", 1, 2, lerp_t=0.3)
    np.testing.assert_array_equal(ctx.download_field(dims), orc.encode_field(v[1], dims, v1=v[2], t=0.3))


def test_load_field_dat_uchar(ctx, tmp_path):
    dims = (6, 6, 6)
    u = np.random.default_rng(5).integers(0, 256, (6, 6, 6, 3), dtype=np.uint8)
    (tmp_path / "u.raw").write_bytes(u.tobytes())
    _dat(tmp_path / "u.dat", "u.raw", dims, "UCHAR3")
    ctx.load_field_dat(tmp_path / "u.dat")
    ref = orc.encode_field(u.astype(np.float32) - 128.0, dims)
    np.testing.assert_array_equal(ctx.download_field(dims), ref)


def test_load_noise_file_and_gradient_cache(ctx, tmp_path):
    dims = (16, 12, 10)
    noise = orc.gen_noise(7, vv.NOISE_BINARY, 0.2, dims)
    f = tmp_path / "noise.raw"
    L.noise_file_write(f, noise, dims)
    ctx.load_noise_file(f)
    np.testing.assert_array_equal(ctx.download_noise(), noise)
    # gradients computed, cached next to the noise file, equal to the oracle's bytes
    ctx.load_noise_file(f, with_gradients=True)
    grd = (tmp_path / "noise.raw.grd").read_bytes()
    g = orc.noise_gradients(noise, dims)
    assert grd == g.tobytes()
    got = ctx.download_noise().reshape(-1, 4)
    np.testing.assert_array_equal(got[:, :3].ravel(), g)
    np.testing.assert_array_equal(got[:, 3], noise)
    # the cache is read when present: plant different bytes and see them come back
    planted = (np.arange(3 * noise.size) % 251).astype(np.uint8)
    (tmp_path / "noise.raw.grd").write_bytes(planted.tobytes())
    ctx.load_noise_file(f, with_gradients=True)
    np.testing.assert_array_equal(ctx.download_noise().reshape(-1, 4)[:, :3].ravel(), planted)
    # a cache of the wrong size is ignored and rewritten
    (tmp_path / "noise.raw.grd").write_bytes(b"\0" * 10)
    ctx.load_noise_file(f, with_gradients=True)
    assert (tmp_path / "noise.raw.grd").read_bytes() == g.tobytes()


def test_load_filter_png_and_lic(ctx, tmp_path):
    # a 100-wide triangle: centred in 128 entries (shift 14), first channel of the first row
    w = 100
    tri = np.rint(255 * (1 - np.abs(np.linspace(-1, 1, w)))).astype(np.uint8)
    img = np.zeros((3, w, 3), np.uint8)
    img[0, :, 0] = tri
    img[0, :, 1] = 255 - tri  # other channels and rows are ignored
    img[1:] = 77
    Image.fromarray(img, "RGB").save(tmp_path / "filter.png")
    ctx.load_filter_png(tmp_path / "filter.png")
    lut = np.zeros(128, np.uint8)
    lut[14:14 + w] = tri
    n = 16
    v = fields.abc_flow(n)
    ctx.upload_field(v, (n, n, n))
    ctx.generate_noise(3, vv.NOISE_BINARY, 0.1, (n, n, n))
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = 12
    ctx.set_lic_params(p)
    np.testing.assert_array_equal(ctx.lic_weights(), orc.lic_weights(lut, 12, 12))
    ctx.compute_lic()
    inv = float(np.float32(0.5 * 128 * 255.0) / np.float32(lut.astype(np.float32).sum()))
    op = orc.params(steps_forward=12, steps_backward=12, inv_filter_area=inv)
    ref = orc.lic(orc.encode_field(v, (n, n, n)), (n, n, n), orc.gen_noise(3, vv.NOISE_BINARY, 0.1, (n, n, n)),
                  1, (n, n, n), lut, op)
    assert rel_err(ctx.download_lic(), ref) <= 1e-5


def test_load_scalar_dat_gates_lic(ctx, tmp_path):
    n = 16
    dims = (n, n, n)
    s = np.random.default_rng(6).uniform(-0.2, 1.2, (n, n, n)).astype(np.float32)
    (tmp_path / "s.raw").write_bytes(s.tobytes())
    _dat(tmp_path / "s.dat", "s.raw", dims, "FLOAT")
    v = fields.abc_flow(n)
    ctx.upload_field(v, dims)
    ctx.generate_noise(2, vv.NOISE_UNIFORM, 0.0, dims)
    ctx.load_scalar_dat(tmp_path / "s.dat")
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = 8
    p.gate_enable = 1
    ctx.set_lic_params(p)
    ctx.compute_lic()
    su8 = np.rint(np.clip(s, 0, 1) * 255).astype(np.uint8).ravel()
    op = orc.params(steps_forward=8, steps_backward=8, gate_enable=1)
    ref = orc.lic(orc.encode_field(v, dims), dims, orc.gen_noise(2, vv.NOISE_UNIFORM, 0.0, dims), 1, dims,
                  np.full(256, 255, np.uint8), op, scalar=su8, sdims=dims)
    assert np.any(ref == 0) and np.any(ref != 0)
    assert rel_err(ctx.download_lic(), ref) <= 1e-5


def test_noise_fallback_generates_white_noise(ctx, tmp_path):
    # NoiseDataSet::loadData: an unreadable noise file gives 256^3 binary
    # noise at density 1/6 (the reference's floor(0.6 u + 0.5) * 255)
    ref = orc.gen_noise(9, vv.NOISE_BINARY, 1.0 / 6.0, (256, 256, 256))
    assert abs(np.mean(ref == 255) - 1.0 / 6.0) < 1e-3
    for path in (None, tmp_path / "missing.raw"):
        assert ctx.load_noise_or_generate(path, seed=9) is True
        np.testing.assert_array_equal(ctx.download_noise(), ref)
    short = tmp_path / "short.raw"
    short.write_bytes(np.array([10, 10, 10], np.int32).tobytes() + bytes(5))
    assert ctx.load_noise_or_generate(short, seed=9) is True
    assert ctx.noise_info()[0] == [256, 256, 256]
    # a readable file is used as is
    dims = (12, 10, 8)
    noise = orc.gen_noise(2, vv.NOISE_UNIFORM, 0.0, dims)
    good = tmp_path / "good.raw"
    L.noise_file_write(good, noise, dims)
    assert ctx.load_noise_or_generate(good, seed=9) is False
    np.testing.assert_array_equal(ctx.download_noise(), noise)


def test_failed_filter_and_tf_loads_keep_the_defaults(ctx, tmp_path):
    # a missing filter PNG leaves the box filter (VV/3DLIC.cpp:725-729), a
    # missing TF file set leaves the caller's table untouched
    p = vv.default_lic_params()
    p.steps_forward = p.steps_backward = 6
    ctx.set_lic_params(p)
    with pytest.raises(vv.VVError):
        ctx.load_filter_png(tmp_path / "nofilter.png")
    np.testing.assert_array_equal(ctx.lic_weights(), np.ones(13, np.float32))
    tf5 = np.arange(1280, dtype=np.uint8) % 251
    with pytest.raises(vv.VVError):
        L.tf_load_png(tmp_path / "notf", tf5)
    np.testing.assert_array_equal(tf5, np.arange(1280, dtype=np.uint8) % 251)
