"""Pins the oracle (CPU, no GPU needed).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a transposed axis, a wrong wrap or sign) fails it, and none of
them retypes the oracle's own formula: they use published vectors, closed
forms, invariants and independent library routines (numpy float16,
scipy.ndimage).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as orc

# ---- half precision --------------------------------------------------------


def test_f16_to_f32_all_halves_match_numpy():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = orc.f16_to_f32(h)
    ref = h.view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_f32_to_f16_round_to_nearest_even_matches_numpy():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.uniform(-70000, 70000, 20000),
        rng.uniform(-1, 1, 20000),
        rng.uniform(-1e-4, 1e-4, 20000),           # half denormal range
        rng.uniform(-1e-7, 1e-7, 2000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf,
                  2 ** -24, 2 ** -25, 2 ** -25 * 1.0001, 3 * 2 ** -26, 0.5 + 2 ** -12]),
    ]).astype(np.float32)
    # exact ties between neighbouring halves
    base = rng.integers(0x0400, 0x7b00, 2000).astype(np.uint16).view(np.float16).astype(np.float32)
    ulp = np.spacing(base.astype(np.float16)).astype(np.float32)
    vals = np.concatenate([vals, base + ulp / 2, base - ulp / 2])
    got = orc.f32_to_f16(vals)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, ref)


# ---- noise ------------------------------------------------------------------


def test_splitmix64_published_vectors():
    # Vigna's splitmix64 reference output for seed 1234567
    ref = [6457827717110365317, 3203168211198807973, 9817491932198370423,
           4593380528125082431, 16408922859458223821]
    assert [orc.splitmix64(1234567, i) for i in range(5)] == ref


def test_noise_binary_density_and_values():
    dims = (64, 64, 32)
    n = orc.gen_noise(11, 0, 0.05, dims)
    assert set(np.unique(n)) <= {0, 255}
    p = np.mean(n == 255)
    N = n.size
    assert abs(p - 0.05) < 5 * math.sqrt(0.05 * 0.95 / N)
    assert not np.any(orc.gen_noise(11, 0, 0.0, dims))
    assert np.all(orc.gen_noise(11, 0, 1.0, dims) == 255)
    # seeds give different streams, same seed same stream
    assert not np.array_equal(n, orc.gen_noise(12, 0, 0.05, dims))
    assert np.array_equal(n, orc.gen_noise(11, 0, 0.05, dims))


def test_noise_uniform_chi_square():
    n = orc.gen_noise(5, 1, 0.0, (64, 64, 64))
    counts = np.bincount(n, minlength=256)
    e = n.size / 256
    chi2 = np.sum((counts - e) ** 2 / e)
    # 255 dof: mean 255, sd ~22.6; 6 sd bound
    assert chi2 < 255 + 6 * 22.6


def test_noise_is_top_byte_of_splitmix():
    n = orc.gen_noise(99, 1, 0.0, (4, 3, 2))
    ref = [orc.splitmix64(99, i) >> 56 for i in range(24)]
    assert list(n) == ref


# ---- noise gradients (VV/gradient.cpp) -------------------------------------
def _w3():
    # filterGradients weights over i,j,k in [-2,0] (VV/gradient.cpp:394-409)
    g = np.exp(-np.arange(-2, 1) ** 2 / 5.0)
    return g, g.sum() ** 3


def test_gradients_constant_volume_quantize_to_127():
    dims = (9, 8, 7)
    n = np.full(np.prod(dims), 200, np.uint8)
    g = orc.noise_gradients(n, dims)
    # zero gradient -> (0+1)/2*255 = 127.5 -> 127 (VV/gradient.cpp:468-477)
    assert np.all(g == 127)


def test_gradients_sobel_interior_matches_scipy():
    dims = (10, 9, 8)
    rng = np.random.default_rng(1)
    vol = rng.integers(0, 256, dims[::-1]).astype(np.uint8)
    raw = orc.noise_gradients_raw(vol.ravel(), dims).reshape(dims[2], dims[1], dims[0], 3)
    # the reference's 3x3x3 weight table: index [dir][i][j][k] multiplies voxel
    # (x+i-1, y+j-1, z+k-1); as an array over (z, y, x) it is the transpose
    w = np.array([
        [[[-1, -3, -1], [-3, -6, -3], [-1, -3, -1]], [[0] * 3] * 3, [[1, 3, 1], [3, 6, 3], [1, 3, 1]]],
        [[[-1, -3, -1], [0, 0, 0], [1, 3, 1]], [[-3, -6, -3], [0, 0, 0], [3, 6, 3]],
         [[-1, -3, -1], [0, 0, 0], [1, 3, 1]]],
        [[[-1, 0, 1], [-3, 0, 3], [-1, 0, 1]], [[-3, 0, 3], [-6, 0, 6], [-3, 0, 3]],
         [[-1, 0, 1], [-3, 0, 3], [-1, 0, 1]]]], np.float64)
    for d in range(3):
        k_zyx = np.transpose(w[d], (2, 1, 0))
        ref = ndimage.correlate(vol.astype(np.float64), k_zyx, mode="constant") / 2.0
        np.testing.assert_array_equal(raw[1:-1, 1:-1, 1:-1, d], ref[1:-1, 1:-1, 1:-1])
    # faces: one-sided differences (VV/gradient.cpp:264-307)
    f = vol.astype(np.float64)
    np.testing.assert_array_equal(raw[0, :, :, 2], f[1] - f[0])
    np.testing.assert_array_equal(raw[-1, :, :, 2], f[-1] - f[-2])
    np.testing.assert_array_equal(raw[:, :, 0, 0], f[:, :, 1] - f[:, :, 0])


def test_gradients_linear_ramp_gives_axis_normal():
    dims = (12, 11, 10)
    for axis in range(3):
        idx = np.indices(dims[::-1])[2 - axis]
        vol = (idx * 9).astype(np.uint8)
        raw = orc.noise_gradients_raw(vol.ravel(), dims).reshape(dims[2], dims[1], dims[0], 3)
        inner = raw[1:-1, 1:-1, 1:-1]
        # sum over the 3x3 face of the weight table is 22, /2 per unit slope*2
        assert np.all(inner[..., axis] == 9 * 22)
        assert np.all(inner[..., [a for a in range(3) if a != axis]] == 0)
        q = orc.noise_gradients(vol.ravel(), dims).reshape(dims[2], dims[1], dims[0], 3)
        deep = q[2:-2, 2:-2, 2:-2]
        assert np.all(deep[..., axis] == 255)
        assert np.all(deep[..., [a for a in range(3) if a != axis]] == 127)


def test_filter_gradients_window_quirks():
    dims = (9, 9, 9)
    g = np.zeros((9, 9, 9, 3), np.float32)
    g[...] = (1.0, -2.0, 0.5)
    out = orc.filter_gradients(g.ravel(), dims).reshape(9, 9, 9, 3)
    _, S = _w3()
    # deep interior (fw = 2): normalised weights sum to one
    np.testing.assert_allclose(out[4, 4, 4], (1.0, -2.0, 0.5), rtol=1e-6)
    # one voxel from the border (fw = 1): single tap weighted w(-2,-2,-2)/S
    w1 = math.exp(-12.0 / 5.0) / S
    np.testing.assert_allclose(out[1, 4, 4], np.array((1.0, -2.0, 0.5)) * w1, rtol=1e-6)
    # border layer (fw = 0): no taps
    assert np.all(out[0] == 0) and np.all(out[:, :, -1] == 0)


# ---- filter LUT ------------------------------------------------------------


def test_box_filter():
    lut, inv = orc.make_filter(0, 20)
    assert lut.size == 256 and np.all(lut == 255) and inv == 0.5


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_filters_symmetric_and_normalised(kind):
    lut, inv = orc.make_filter(kind, 40)
    assert np.array_equal(lut, lut[::-1])
    assert lut[127] == 255 or lut[127] >= 254
    np.testing.assert_allclose(inv, 0.5 * 256 * 255 / lut.astype(np.float64).sum(), rtol=1e-6)


def test_triangle_area_closed_form():
    lut, inv = orc.make_filter(1, 30)
    # area of a unit triangle over the unit interval is 1/2 -> inv ~ 1
    assert abs(inv - 1.0) < 0.01
    g, inv_g = orc.make_filter(3, 40)
    # Gaussian sigma = L/4 over s in [-L, L]: area = sqrt(2 pi)/8 of the interval
    assert abs(1.0 / inv_g / 2.0 - math.sqrt(2 * math.pi) / 8.0) < 0.01


# ---- field encode -----------------------------------------------------------


def test_encode_closed_form_and_zero_vector():
    rng = np.random.default_rng(2)
    dims = (5, 4, 3)
    v0 = rng.normal(size=(3, 4, 5, 3)).astype(np.float32)
    v1 = rng.normal(size=(3, 4, 5, 3)).astype(np.float32)
    v0[0, 0, 0] = 0.0
    v1[0, 0, 0] = 0.0
    t = 0.3
    v = v0.astype(np.float64) + t * (v1.astype(np.float64) - v0.astype(np.float64))
    ln = np.linalg.norm(v, axis=-1)
    ref = np.empty((3, 4, 5, 4))
    ref[..., :3] = 0.5 * v / np.where(ln > 0, ln, 1)[..., None] + 0.5
    ref[0, 0, 0, :3] = 0.5
    ref[..., 3] = ln / ln.max()
    got64 = orc.encode_field(v0, dims, v1=v1, t=t, f64=True).reshape(3, 4, 5, 4)
    got32 = orc.encode_field(v0, dims, v1=v1, t=t).reshape(3, 4, 5, 4)
    np.testing.assert_allclose(got64, ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(got32, ref, rtol=0, atol=2e-6)
    assert np.all(got32[0, 0, 0] == [0.5, 0.5, 0.5, 0.0])


def test_encode_raw_magnitude_keeps_the_length():
    # fillTexDataFloat: rgb as the interpolated encoding, a = |v| unnormalised
    v = np.array([[3, 4, 0], [0, 0, 0], [0, 0.5, 0], [1, 2, 2]], np.float32)
    for f64 in (False, True):
        raw = orc.encode_field(v, (4, 1, 1), f64=f64, raw_magnitude=True).reshape(4, 4)
        ref = orc.encode_field(v, (4, 1, 1), f64=f64).reshape(4, 4)
        np.testing.assert_array_equal(raw[:, :3], ref[:, :3])
        np.testing.assert_array_equal(raw[:, 3], [5.0, 0.0, 0.5, 3.0])
        np.testing.assert_array_equal(ref[:, 3], np.array([1.0, 0.0, 0.1, 0.6], raw.dtype))


def test_encode_u8_known_answers():
    # fillTexDataCharInterp: byte = (unsigned char)(v/|v|*127.0 + 128.0), truncated
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 0], [3, 4, 0], [-3, -4, 0],
                  [-1e-6, 1, 0], [-1e-9, 1, 0], [0, 0, 0.5],
                  [-2.0 ** -53, 1, 0], [-2.0 ** -54, 1, 0], [-1e-40, 1, 0]], np.float32)
    got = orc.encode_field_u8(v, (11, 1, 1)).reshape(11, 4)
    # directions: 0.6f*127+128 = 204.2, 0.8f*127+128 = 229.6 -> 229 (not rounded
    # to 230), -0.6f -> 51.8 -> 51, -0.8f -> 26.4 -> 26; a component just below
    # zero gives 127 even where float arithmetic would round 127q+128 to 128,
    # but 128 once 127|q| <= 2^-47, where the reference's double sum rounds
    # to 128.0 (127 * 2^-53 > 2^-47 > 127 * 2^-54; 1e-40 is subnormal)
    np.testing.assert_array_equal(got[:, :3], [[255, 128, 128], [1, 128, 128], [128, 128, 128],
                                              [204, 229, 128], [51, 26, 128], [127, 255, 128],
                                              [127, 255, 128], [128, 128, 255],
                                              [127, 255, 128], [128, 255, 128], [128, 255, 128]])
    # magnitudes: |v|/5 * 255 truncated: 1 -> 51, 5 -> 255, 0 -> 0, 0.5 -> 25.5 -> 25
    np.testing.assert_array_equal(got[:, 3], [51, 51, 0, 255, 255, 51, 51, 25, 51, 51, 51])


def test_encode_u8_lerp_and_degenerate_fields():
    v0 = np.array([[1, 0, 0], [0, 0, 2]], np.float32)
    v1 = np.array([[0, 1, 0], [0, 0, 2]], np.float32)
    got = orc.encode_field_u8(v0, (2, 1, 1), v1=v1, t=0.5).reshape(2, 4)
    # (0.5, 0.5, 0): q = 0.70710677 -> 127q + 128 = 217.8 -> 217; |v| = 0.7071 of max 2
    np.testing.assert_array_equal(got[0], [217, 217, 128, 90])
    np.testing.assert_array_equal(got[1], [128, 128, 255, 255])
    # an all-zero field: rgb 128, magnitude 0 (0/0 in the reference)
    z = orc.encode_field_u8(np.zeros((5, 3), np.float32), (5, 1, 1)).reshape(5, 4)
    assert np.all(z == [128, 128, 128, 0])


def test_encode_u8_within_one_byte_of_the_direction():
    rng = np.random.default_rng(11)
    v = rng.normal(size=(4000, 3)).astype(np.float32)
    b = orc.encode_field_u8(v, (4000, 1, 1)).reshape(-1, 4).astype(np.float64)
    q = v / np.linalg.norm(v.astype(np.float64), axis=1, keepdims=True)
    e = b[:, :3] - (127.0 * q + 128.0)   # truncation: in (-1, 0]
    assert np.all(e <= 1e-6) and np.all(e > -1.0 - 1e-6)
    m = np.linalg.norm(v.astype(np.float64), axis=1)
    em = b[:, 3] - 255.0 * m / m.max()
    assert np.all(em <= 1e-4) and np.all(em > -1.0 - 1e-4)


def test_unorm8_sampling_equals_prescaled_field():
    # sampling bytes with field_unorm8 (blend, then x 1/255) and sampling the
    # pre-scaled field b/255 differ only by rounding, in every use of the
    # field: LIC directions and the ray caster's TF index (.b and .a) and
    # line-illumination tangents
    n, L = 10, 6
    dims = (n, n, n)
    v = np.random.default_rng(12).normal(size=(n, n, n, 3)).astype(np.float32)
    b = orc.encode_field_u8(v, dims)
    fb = orc.field_bytes_as_float(b, f64=True)
    fs = b.astype(np.float64) / 255.0
    noise = orc.gen_noise(3, 1, 0.0, dims)
    lut, inv = orc.make_filter(0, L)
    pu = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.02, inv_filter_area=inv,
                    field_unorm8=1, gradient_scale=3.0)
    pf = orc.params(steps_forward=L, steps_backward=L, step_size_lic=0.02, inv_filter_area=inv,
                    gradient_scale=3.0)
    a = orc.lic(fb, dims, noise, 1, dims, lut, pu)
    c = orc.lic(fs, dims, noise, 1, dims, lut, pf)
    np.testing.assert_allclose(a, c, rtol=1e-12, atol=1e-12)
    assert np.max(np.abs(a - orc.lic(fb, dims, noise, 1, dims, lut, pf))) > 0.1  # the flag matters
    rgba, op = np.random.default_rng(13).uniform(0, 1, 1024), np.linspace(0, 1, 256)
    luts = orc.illum_luts()
    view = orc.view_setup(24, 24, quat=(0.2, 0.3, 0.1, 0.927))
    for tf_index, illum in ((0, 0), (1, 0), (0, 2), (1, 3)):
        for q in (pu, pf):
            q.tf_index, q.illum_mode = tf_index, illum
            orc.attach_luts(q, luts)
        iu, nu, _ = orc.raycast(fb, dims, a, 1, dims, rgba, op, pu, view, f64=True)
        ifl, nf, _ = orc.raycast(fs, dims, a, 1, dims, rgba, op, pf, view, f64=True)
        np.testing.assert_array_equal(nu, nf)
        np.testing.assert_allclose(iu, ifl, rtol=0, atol=1e-12)


# ---- LIC -------------------------------------------------------------------


def _case(n=8, F=5, B=7, h=0.01, kind=0):
    lut, inv = orc.make_filter(kind, F)
    p = orc.params(steps_forward=F, steps_backward=B, step_size_lic=h, inv_filter_area=inv)
    return lut, inv, p


def test_lic_constant_noise_closed_form():
    n = 8
    rng = np.random.default_rng(3)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    for kind in (0, 3):
        lut, inv, p = _case(kind=kind)
        b = 200
        noise = np.full(n ** 3, b, np.uint8)
        for f64 in (False, True):
            f4 = orc.encode_field(v, (n, n, n), f64=f64)
            out = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p)
            w = orc.lic_weights(lut, 5, 7, f64=True)
            ref = b / 255.0 * w.sum() * inv / 12.0 * 30.0
            np.testing.assert_allclose(out, ref, rtol=1e-6 if not f64 else 1e-12)
    # box weights are all exactly one
    assert np.all(orc.lic_weights(np.full(256, 255, np.uint8), 9, 4) == 1.0)


def test_lic_zero_noise_is_zero():
    n = 8
    lut, inv, p = _case()
    f4 = orc.encode_field(np.random.default_rng(0).normal(size=(n, n, n, 3)), (n, n, n))
    out = orc.lic(f4, (n, n, n), np.zeros(n ** 3, np.uint8), 1, (n, n, n), lut, p)
    assert np.all(out == 0)


def test_lic_uniform_field_straight_streamline():
    # v = (1,0,0): encoded rgb (1, .5, .5); each Heun step moves x by exactly h_e
    n, F, B, h = 16, 9, 6, 0.02
    lut, inv, p = _case(n, F, B, h)
    v = np.zeros((n, n, n, 3), np.float32)
    v[..., 0] = 1.0
    f4 = orc.encode_field(v, (n, n, n), f64=True)
    rng = np.random.default_rng(5)
    line = rng.integers(0, 256, n).astype(np.uint8)   # noise varies along x only
    noise = np.broadcast_to(line, (n, n, n)).astype(np.uint8).ravel()
    out = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p).reshape(n, n, n)
    he = np.float64(np.float32(h)) * np.float64(np.float32(0.3))  # f64 oracle: h_e in double
    xs = (np.arange(n) + 0.5) / n
    lp = np.concatenate([line, line[:1]]).astype(np.float64)  # periodic
    def N(u):  # REPEAT linear interpolation along x, independent of the oracle
        x = u * n - 0.5
        i = np.floor(x)
        f = x - i
        i = np.mod(i, n).astype(int)
        return ((1 - f) * lp[i] + f * lp[i + 1]) / 255.0
    ks = np.concatenate([[0], -np.arange(1, B + 1), np.arange(1, F + 1)])
    ref = sum(N(xs + k * float(he)) for k in ks) * inv / (F + B) * 30.0
    for z in (0, 7):
        for y in (3, 15):
            np.testing.assert_allclose(out[z, y], ref, rtol=1e-9)


def test_lic_f64_matches_scipy_brute_force():
    n, F, B, h = 8, 6, 6, 0.03
    lut, inv, p = _case(n, F, B, h)
    v = np.random.default_rng(8).normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n), f64=True).reshape(n, n, n, 4)
    noise = orc.gen_noise(4, 1, 0.0, (n, n, n)).reshape(n, n, n).astype(np.float64)
    out = orc.lic(f4.ravel(), (n, n, n), noise.astype(np.uint8).ravel(), 1, (n, n, n), lut, p)

    def samp(vol, pts, mode):  # pts (m,3) texture coords x,y,z
        c = np.stack([pts[:, 2], pts[:, 1], pts[:, 0]]) * n - 0.5
        return ndimage.map_coordinates(vol, c, order=1, mode=mode)

    def F3(pts):
        return np.stack([samp(f4[..., ch], pts, "nearest") for ch in range(3)], 1)

    def Nz(pts):
        return samp(noise, pts, "grid-wrap") / 255.0

    g = (np.indices((n, n, n)).reshape(3, -1)[::-1].T + 0.5) / n
    he = float(np.float64(np.float32(h)) * np.float64(np.float32(0.3)))
    w = orc.lic_weights(lut, F, B, f64=True)
    acc = Nz(g) * w[0]
    for sgn, steps, off in ((-1, B, 1), (1, F, 1 + B)):
        x = g.copy()
        s = F3(x)
        for i in range(steps):
            d1 = (2 * s - 1) * sgn * he
            d2 = (2 * F3(x + d1) - 1) * sgn * he
            x = x + 0.5 * (d1 + d2)
            s = F3(x)
            acc = acc + Nz(x) * w[off + i]
    ref = acc * inv / (F + B) * 30.0
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-12)


def test_lic_f32_close_to_f64():
    n = 12
    lut, inv, p = _case(n, 10, 10, 0.01, kind=3)
    v = np.random.default_rng(9).normal(size=(n, n, n, 3)).astype(np.float32)
    noise = orc.gen_noise(1, 1, 0.0, (n, n, n))
    a = orc.lic(orc.encode_field(v, (n, n, n)), (n, n, n), noise, 1, (n, n, n), lut, p)
    b = orc.lic(orc.encode_field(v, (n, n, n), f64=True), (n, n, n), noise, 1, (n, n, n), lut, p)
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-4


# ---- view ------------------------------------------------------------------
def _quat_to_R(q):
    x, y, z, w = np.array(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _T(t):
    M = np.eye(4)
    M[:3, 3] = t
    return M


def test_view_matches_glu_unproject():
    W, H, fovy, dist = 37, 23, 35.0, 4.0
    q = np.array([0.1, -0.3, 0.2, 0.9])
    q = q / np.linalg.norm(q)
    pos = np.array([0.05, -0.02, 0.3])
    c = np.array([0.5, 0.4, 0.3])
    v = orc.view_setup(W, H, fovy=fovy, dist=dist, pos=pos, quat=q, center=c,
                       extent=(1, 0.8, 0.6))
    Rm = np.eye(4)
    Rm[:3, :3] = _quat_to_R(q)
    MV = _T(pos) @ _T([0, 0, -dist]) @ Rm @ _T(-c)
    f = 1.0 / math.tan(math.radians(fovy) / 2)
    near, far = 0.1, 50.0
    P = np.array([[f / (W / H), 0, 0, 0], [0, f, 0, 0],
                  [0, 0, (far + near) / (near - far), 2 * far * near / (near - far)], [0, 0, -1, 0]])
    inv = np.linalg.inv(P @ MV)
    cam = np.linalg.inv(MV)[:3, 3]
    np.testing.assert_allclose(np.array(v.cam), cam, atol=1e-6)
    for i, j in [(0, 0), (W - 1, 0), (18, 11), (5, 20)]:
        ndc = [(2 * (i + 0.5) / W - 1), (2 * (j + 0.5) / H - 1)]
        a = inv @ np.array([*ndc, -1, 1]); a = a[:3] / a[3]
        b = inv @ np.array([*ndc, 1, 1]); b = b[:3] / b[3]
        ref = (b - a) / np.linalg.norm(b - a)
        d = np.array(v.dw) + ndc[0] * np.array(v.du) + ndc[1] * np.array(v.dv)
        np.testing.assert_allclose(d / np.linalg.norm(d), ref, atol=1e-6)


def test_light_position_default():
    np.testing.assert_allclose(orc.light_position((0.5, 0.5, 0.5)), (0.5, 0.5, 1.5))


# ---- ray cast --------------------------------------------------------------
def _const_scene(n, lval, tfv, opv, W=33, H=33, step=1.0 / 128.0):
    f4 = np.zeros((n, n, n, 4), np.float32)
    f4[...] = (0.5, 0.5, 0.5, 0.5)
    lic = np.full(n ** 3, lval, np.float32)
    rgba = np.tile(np.array(tfv, np.float32), 256)
    op = np.full(256, opv, np.float32)
    p = orc.params(step_size_vol=step)
    v = orc.view_setup(W, H)
    return f4.ravel(), lic, rgba, op, p, v


def test_raycast_constant_volume_closed_form():
    n = 8
    ell = 0.5
    tfv = (0.8, 0.6, 0.4, 0.1)
    f4, lic, rgba, op, p, v = _const_scene(n, ell, tfv, 0.1)
    img, ns, total = orc.raycast(f4, (n, n, n), lic, 1, (n, n, n), rgba, op, p, v, f64=True)
    img = img.reshape(33, 33, 4)
    f32 = lambda x: float(np.float32(x))  # noqa: E731  (tables are float32)
    a = f32(0.1) * f32(0.1)
    k = ns.reshape(33, 33)[16, 16]
    assert k == 129  # z from 1 down to 0 in steps of 1/128
    T = (1 - a) ** k
    np.testing.assert_allclose(img[16, 16, 3], 1 - T, rtol=1e-9)
    np.testing.assert_allclose(img[16, 16, :3], np.float32(tfv[:3]).astype(np.float64) * ell * (1 - T),
                               rtol=1e-9)
    assert np.all(img[0, 0] == 0)  # corner pixels miss the box
    assert total == ns.sum()


def test_raycast_early_termination_index():
    n = 8
    f4, lic, rgba, op, p, v = _const_scene(n, 0.5, (1, 1, 1, 1.0), 0.1)
    img, ns, _ = orc.raycast(f4, (n, n, n), lic, 1, (n, n, n), rgba, op, p, v, f64=True)
    # a = 0.1: stop after the first n with 1 - 0.9^n > 0.95 -> n = 29
    assert ns.reshape(33, 33)[16, 16] == 29
    a = float(np.float32(0.1))
    np.testing.assert_allclose(img.reshape(33, 33, 4)[16, 16, 3], 1 - (1 - a) ** 29, rtol=1e-12)


def test_raycast_alpha_correction_half_step():
    n = 8
    f4, lic, rgba, op, p, v = _const_scene(n, 0.5, (1, 1, 1, 0.2), 0.1, step=1.0 / 256.0)
    img, ns, _ = orc.raycast(f4, (n, n, n), lic, 1, (n, n, n), rgba, op, p, v, f64=True)
    k = ns.reshape(33, 33)[16, 16]
    assert k == 257
    a = 1 - math.sqrt(1 - float(np.float32(0.2)) * float(np.float32(0.1)))  # 1-(1-a)^(128/256)
    np.testing.assert_allclose(img.reshape(33, 33, 4)[16, 16, 3], 1 - (1 - a) ** k, rtol=1e-9)


def test_raycast_empty_tf_black():
    n = 8
    f4, lic, rgba, op, p, v = _const_scene(n, 0.5, (1, 1, 1, 0.0), 0.5)
    img, _, _ = orc.raycast(f4, (n, n, n), lic, 1, (n, n, n), rgba, op, p, v)
    assert np.all(img == 0)


def test_raycast_f32_close_to_f64():
    n = 12
    rng = np.random.default_rng(2)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n))
    lic = rng.uniform(0, 1, n ** 3).astype(np.float32)
    rgba = rng.uniform(0, 1, 1024).astype(np.float32)
    op = rng.uniform(0, 0.3, 256).astype(np.float32)
    p = orc.params()
    view = orc.view_setup(40, 30, quat=(0.1, 0.2, 0.0, 0.97))
    a, _, _ = orc.raycast(f4, (n,) * 3, lic, 1, (n,) * 3, rgba, op, p, view)
    b, _, _ = orc.raycast(f4.astype(np.float64), (n,) * 3, lic, 1, (n,) * 3, rgba, op, p, view)
    assert np.max(np.abs(a - b)) < 1e-4


# ---- adjoint ----------------------------------------------------------------
@pytest.mark.parametrize("lic_ch,illum,step", [(1, 0, 1.0 / 128.0), (4, 1, 1.0 / 256.0),
                                              (1, 2, 1.0 / 128.0), (4, 3, 1.0 / 200.0)])
def test_adjoint_matches_central_differences(lic_ch, illum, step):
    n = 8
    rng = np.random.default_rng(6)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n), f64=True)
    lic = rng.uniform(0.05, 0.6, n ** 3 * lic_ch)
    rgba = rng.uniform(0.05, 0.9, 1024)
    op = rng.uniform(0.01, 0.08, 256)
    sigma = rng.uniform(0.5, 1.5, n ** 3)
    p = orc.params(step_size_vol=step, illum_mode=illum, light=(0.3, 0.8, 1.2))
    luts = orc.illum_luts()
    orc.attach_luts(p, luts)
    view = orc.view_setup(14, 12, quat=(0.2, -0.1, 0.05, 0.97))
    W = rng.normal(size=14 * 12 * 4)

    def loss(tf, sg):
        img, _, _ = orc.raycast(f4, (n,) * 3, lic, lic_ch, (n,) * 3, tf, op, p, view, sigma=sg,
                                f64=True)
        return float(np.dot(W, img))

    dtf, dsg = orc.raycast_adjoint(f4, (n,) * 3, lic, lic_ch, (n,) * 3, rgba, op, p, view, W,
                                   sigma=sigma)
    hh = 1e-6
    used = np.flatnonzero(dtf)
    for idx in rng.choice(used, 12, replace=False):
        e = np.zeros(1024)
        e[idx] = hh
        fd = (loss(rgba + e, sigma) - loss(rgba - e, sigma)) / (2 * hh)
        assert abs(fd - dtf[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, dtf[idx])
    usedv = np.flatnonzero(dsg)
    for idx in rng.choice(usedv, 8, replace=False):
        e = np.zeros(n ** 3)
        e[idx] = hh
        fd = (loss(rgba, sigma + e) - loss(rgba, sigma - e)) / (2 * hh)
        assert abs(fd - dsg[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, dsg[idx])


# ---- fused ray caster ---------------------------------------------------------
def test_fused_equals_volume_caster_for_constant_lic():
    # constant noise makes the LIC constant everywhere: the on-the-fly LIC of
    # every sample then equals the trilinear LIC volume, and with low opacity
    # neither caster terminates early, so the images agree
    n = 10
    rng = np.random.default_rng(8)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n))
    noise = np.full(n ** 3, 200, np.uint8)
    lut, inv = orc.make_filter(1, 6)
    p = orc.params(steps_forward=6, steps_backward=6, inv_filter_area=inv, gradient_scale=1.0)
    lic = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p)
    assert np.ptp(lic) < 1e-5 * np.max(lic)
    t = np.arange(256) / 255.0
    rgba = np.stack([t, 1 - t, 0.5 + 0 * t, 0.01 + 0.01 * t], 1).astype(np.float32).ravel()
    op = t.astype(np.float32)
    view = orc.view_setup(24, 20, quat=(0.1, 0.2, 0.0, 0.97))
    a, na, _ = orc.raycast(f4, (n,) * 3, lic, 1, (n,) * 3, rgba, op, p, view)
    b, nb, _ = orc.raycast_fused(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    assert np.max(a) > 0
    np.testing.assert_array_equal(na, nb)
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(a)


def test_fused_terminates_on_sample_alpha():
    # a constant per-sample alpha of 0.5: the volume caster stops when the
    # accumulated alpha 1 - 0.5^k exceeds 0.95 (k = 5); the fused caster
    # tests the sample's own alpha, never above 0.95, and runs to the exit
    n = 8
    f4 = orc.encode_field(np.ones((n, n, n, 3), np.float32), (n, n, n))
    noise = np.full(n ** 3, 255, np.uint8)
    lut = np.full(256, 255, np.uint8)
    p = orc.params(steps_forward=2, steps_backward=2, inv_filter_area=0.5, gradient_scale=1.0)
    licv = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p)
    rgba = np.tile(np.array([1, 1, 1, 0.5], np.float32), 256)
    op = np.ones(256, np.float32)  # opacity table 1 at any LIC value
    view = orc.view_setup(16, 16)
    _, n_vol, _ = orc.raycast(f4, (n,) * 3, licv, 1, (n,) * 3, rgba, op, p, view)
    _, n_fused, _ = orc.raycast_fused(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    _, n_full, _ = orc.raycast(f4, (n,) * 3, licv, 1, (n,) * 3, rgba * np.tile([1, 1, 1, 0], 256), op,
                               p, view)
    hit = n_full > 0
    assert np.all(n_vol[hit] == np.minimum(5, n_full[hit]))
    np.testing.assert_array_equal(n_fused, n_full)


# ---- view-aligned slicing ---------------------------------------------------------
def _slicing_case(alpha):
    n = 8
    f4 = orc.encode_field(np.ones((n, n, n, 3), np.float32), (n, n, n))
    noise = np.full(n ** 3, 255, np.uint8)
    lut = np.full(256, 255, np.uint8)
    p = orc.params(steps_forward=2, steps_backward=2, inv_filter_area=0.5, gradient_scale=1.0,
                   step_size_vol=1.0 / 128.0)
    rgba = np.tile(np.array([1, 1, 1, alpha], np.float32), 256)
    op = np.ones(256, np.float32)
    view = orc.view_setup(15, 15)  # identity rotation: the slices are z planes
    return f4, noise, lut, p, rgba, op, view, n


def test_slicing_counts_and_fp16_target():
    # n = D / step + 1 = 129 z planes cross the central ray; with a per-sample
    # alpha of 0.01 all are composited
    f4, noise, lut, p, rgba, op, view, n = _slicing_case(0.06)
    op = np.full(256, 0.01 / 0.06, np.float32)
    img, ns, _ = orc.render_slices(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    assert ns[7 * 15 + 7] == 129
    # the render target holds halves
    assert np.array_equal(img, img.astype(np.float16).astype(np.float32))


def test_slicing_stops_at_accumulated_alpha():
    # per-sample alpha 0.5: dest.a = 0.5, 0.75, 0.875, 0.9375, 0.96875, then
    # every later slice is skipped (dest.a >= 0.95)
    f4, noise, lut, p, rgba, op, view, n = _slicing_case(0.5)
    img, ns, _ = orc.render_slices(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    c = 7 * 15 + 7
    assert ns[c] == 5
    assert img[4 * c + 3] == np.float32(0.96875)


def test_slicing_skips_low_tf_alpha():
    # TF alpha at most 0.05: no sample is computed
    f4, noise, lut, p, rgba, op, view, n = _slicing_case(0.05)
    img, ns, total = orc.render_slices(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    assert total == 0 and np.all(img == 0)


def test_fast_build_equals_the_checker():
    """The timed CPU baseline runs liboracle_fast.so (-O3, AVX2/FMA code
    generation, -ffp-contract=off): its LIC must equal the checker's bit for
    bit, in FP32 and FP64."""
    n = 12
    lut, inv, p = _case(n, 9, 11, 0.02, kind=3)
    v = np.random.default_rng(31).normal(size=(n, n, n, 3)).astype(np.float32)
    noise = orc.gen_noise(5, 1, 0.0, (n, n, n))
    for f64 in (False, True):
        f4 = orc.encode_field(v, (n, n, n), f64=f64)
        a = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p)
        b = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p, fast=True)
        np.testing.assert_array_equal(a, b)


def test_lic_speed_of_flow_scales_each_step_by_the_magnitude():
    """SPEED_OF_FLOW (inc_lic.glsl:107-110,120-122): v = (m(z), 0, 0) with
    m(z) = (z + 1)/n, so |v|/max = m(z) on voxel plane z.  A streamline seeded
    on a voxel centre stays in its z plane, where the magnitude channel
    samples exactly m(z), and each Heun step moves x by m(z) h_e (by h_e
    without SPEED_OF_FLOW)."""
    n, F, B, h = 16, 9, 6, 0.02
    lut, inv, p = _case(n, F, B, h)
    p.speed_of_flow = 1
    m = (np.arange(n) + 1.0) / n
    v = np.zeros((n, n, n, 3), np.float32)
    v[..., 0] = m[:, None, None]
    f4 = orc.encode_field(v, (n, n, n), f64=True)
    rng = np.random.default_rng(6)
    line = rng.integers(0, 256, n).astype(np.uint8)
    noise = np.broadcast_to(line, (n, n, n)).astype(np.uint8).ravel()
    out = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p).reshape(n, n, n)
    he = float(np.float64(np.float32(h)) * np.float64(np.float32(0.3)))
    xs = (np.arange(n) + 0.5) / n
    lp = np.concatenate([line, line[:1]]).astype(np.float64)

    def N(u):  # REPEAT linear interpolation along x, independent of the oracle
        x = u * n - 0.5
        i = np.floor(x)
        f = x - i
        i = np.mod(i, n).astype(int)
        return ((1 - f) * lp[i] + f * lp[i + 1]) / 255.0

    ks = np.concatenate([[0], -np.arange(1, B + 1), np.arange(1, F + 1)])
    for z in (0, 5, 15):
        ref = sum(N(xs + k * m[z] * he) for k in ks) * inv / (F + B) * 30.0
        for y in (2, 11):
            np.testing.assert_allclose(out[z, y], ref, rtol=1e-9)
    # magnitude varying along x: each step moves by M(x_k) h_e, both Heun
    # displacements scaled by the magnitude at the step's start x_k
    # (licdir2 *= step.a with step not yet re-read), M the linear CLAMP
    # interpolation of |v|/max along x
    mx = 0.3 + 0.7 * np.arange(n) / (n - 1)
    v2 = np.zeros((n, n, n, 3), np.float32)
    v2[..., 0] = mx[None, None, :]
    f4b = orc.encode_field(v2, (n, n, n), f64=True)
    outb = orc.lic(f4b, (n, n, n), noise, 1, (n, n, n), lut, p).reshape(n, n, n)
    amag = f4b.reshape(n, n, n, 4)[0, 0, :, 3]

    def M(u):
        return np.interp(u * n - 0.5, np.arange(n), amag)

    refb = N(xs) * 1.0
    for sgn, steps in ((-1, B), (1, F)):
        x = xs.copy()
        for _ in range(steps):
            d = sgn * M(x) * he
            x = x + 0.5 * (d + d)
            refb = refb + N(x)
    refb = refb * inv / (F + B) * 30.0
    for z in (0, 9):
        np.testing.assert_allclose(outb[z, 4], refb, rtol=1e-9)
    p.speed_of_flow = 0
    plain = orc.lic(f4, (n, n, n), noise, 1, (n, n, n), lut, p).reshape(n, n, n)
    ref = sum(N(xs + k * he) for k in ks) * inv / (F + B) * 30.0
    np.testing.assert_allclose(plain[5, 2], ref, rtol=1e-9)
    assert np.max(np.abs(plain[5] - out[5])) > 1e-3


def _blend_bytes(src, count):
    """glBlendFunc(GL_ONE_MINUS_DST_ALPHA, GL_ONE) into an 8-bit RGBA
    framebuffer cleared to 0: `count` equal fragments src (premultiplied,
    clamped to [0,1]), then the white background quad; every result stored
    as round(255 x).  Double precision, written from the GL blend equation.
    Returns the pixel and the number of fragments before its alpha reached
    255 (after which a fragment changes nothing)."""
    d = np.zeros(4)
    s = np.clip(np.asarray(src, np.float64), 0.0, 1.0)
    used = 0
    for _ in range(count):
        if d[3] == 255.0:
            break
        d = np.floor(np.clip(s * (1.0 - d[3] / 255.0) + d / 255.0, 0.0, 1.0) * 255.0 + 0.5)
        used += 1
    d = np.floor(np.clip(1.0 * (1.0 - d[3] / 255.0) + d / 255.0, 0.0, 1.0) * 255.0 + 0.5)
    return d.astype(np.uint8), used


@pytest.mark.parametrize("alpha,rgb", [(0.5, (1.0, 1.0, 1.0)), (0.06, (0.9, 0.4, 0.2)), (0.3, (8.0, 0.5, 0.1))])
def test_slicing_blend_under_operator_closed_form(alpha, rgb):
    """lic3d_slicingblend_fragment.glsl blended with GL_ONE_MINUS_DST_ALPHA,
    GL_ONE (VV/renderer.cpp:1149-1157), then the white quad (:1233-1258).
    The scene's LIC is constant, so every slice adds the same fragment: the
    pixel is the closed-form sequence above.  Per-sample alpha 0.5 shows
    there is no 0.95 test (the framebuffer alpha keeps rising to 255); an rgb
    above 1 shows the clamp of the fragment colour."""
    f4, noise, lut, p, rgba, op, view, n = _slicing_case(alpha)
    rgba = np.tile(np.array([*rgb, alpha], np.float32), 256)
    img, ns, total = orc.render_slices_blend(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    # illumLIC: rgb = l * tf.rgb * illum_scale, l = 255/255 * (F+B+1) * 0.5 / (F+B) * 1.0
    lic = 1.0 * 5 * 0.5 / 4 * 1.0
    a = alpha  # opacity table 1, alpha correction 1
    src = [lic * rgb[0] * a, lic * rgb[1] * a, lic * rgb[2] * a, a]
    img = img.reshape(-1, 4)
    # the central ray crosses all 129 slices (no 0.95 test: alpha 0.5 goes on
    # until the framebuffer alpha is 255)
    c = 7 * 15 + 7
    want, used = _blend_bytes(src, 129)
    np.testing.assert_array_equal(img[c], want)
    assert ns[c] == used
    assert used > 5 if alpha == 0.5 else used == 129
    if rgb[0] > 1.0:
        assert src[0] > 1.0  # the fragment colour is clamped before blending
    for q in range(img.shape[0]):
        np.testing.assert_array_equal(img[q], _blend_bytes(src, int(ns[q]))[0])
    assert np.all(img[:, 3] == 255)


def test_slicing_blend_skipped_fragments_leave_the_white_background():
    f4, noise, lut, p, rgba, op, view, n = _slicing_case(0.05)
    img, ns, total = orc.render_slices_blend(f4, (n,) * 3, noise, 1, (n,) * 3, lut, rgba, op, p, view)
    assert total == 0 and np.all(img == 255)


@pytest.mark.parametrize("sg", [0.4, 1.7, 40.0])
def test_raycast_opacity_scale_closed_form(sg):
    """Opacity scale sigma (the adjoint's second input): each sample's alpha
    is ao * tf.a * sigma, clamped to [0, 1], before the alpha correction.  A
    constant volume and a constant sigma give dest.a = 1 - (1 - a sigma)^k,
    or 1 after one sample when a sigma >= 1."""
    n = 8
    tfv = (0.8, 0.6, 0.4, 0.1)
    f4, lic, rgba, op, p, v = _const_scene(n, 0.5, tfv, 0.4)
    sigma = np.full(n ** 3, sg, np.float64)
    img, ns, _ = orc.raycast(f4, (n, n, n), lic, 1, (n, n, n), rgba, op, p, v, sigma=sigma, f64=True)
    a = min(float(np.float32(0.4)) * float(np.float32(0.1)) * sg, 1.0)
    k = ns.reshape(33, 33)[16, 16]
    want_k = 129
    if a >= 1.0:
        want_k = 1
    else:
        for j in range(1, 130):  # early termination once dest.a > 0.95
            if 1 - (1 - a) ** j > 0.95:
                want_k = j
                break
    assert k == want_k
    px = img.reshape(33, 33, 4)[16, 16]
    np.testing.assert_allclose(px[3], 1 - (1 - a) ** k, rtol=1e-9)
    col = np.float32(tfv[:3]).astype(np.float64) * 0.5  # tf.rgb * l, premultiplied by the same alpha
    np.testing.assert_allclose(px[:3], col * (1 - (1 - a) ** k), rtol=1e-9)


def test_lic_gate_and_frequency_match_scipy_brute_force():
    """freqSampling (inc_lic.glsl:71-90): the noise at x * freqScale, and 0
    unless the scalar volume's value at x lies in (0.1, 0.3); the scalar
    volume is sampled with CLAMP_TO_EDGE, the noise with REPEAT.  Brute force
    on scipy.ndimage.map_coordinates, independent of the oracle."""
    n, F, B, h = 8, 5, 6, 0.03
    freq = float(np.float32(1.7))  # the parameter block holds floats
    lo, hi = float(np.float32(0.1)), float(np.float32(0.3))
    lut, inv, p = _case(n, F, B, h, kind=1)
    p.freq_scale = freq
    p.gate_enable = 1
    rng = np.random.default_rng(12)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n), f64=True).reshape(n, n, n, 4)
    nd = (6, 7, 5)
    noise = rng.integers(0, 256, (nd[2], nd[1], nd[0]), dtype=np.uint8)
    scalar = rng.integers(0, 120, (n, n, n), dtype=np.uint8)  # values around the gate
    out = orc.lic(f4.ravel(), (n, n, n), noise.ravel(), 1, nd, lut, p, scalar=scalar.ravel(),
                  sdims=(n, n, n))

    def samp(vol, pts, dims, mode):
        c = np.stack([pts[:, 2] * dims[2], pts[:, 1] * dims[1], pts[:, 0] * dims[0]]) - 0.5
        return ndimage.map_coordinates(vol, c, order=1, mode=mode)

    def F3(pts):
        return np.stack([samp(f4[..., ch], pts, (n, n, n), "nearest") for ch in range(3)], 1)

    def Nz(pts):
        s = samp(scalar.astype(np.float64), pts, (n, n, n), "nearest") / 255.0
        val = samp(noise.astype(np.float64), pts * freq, nd, "grid-wrap") / 255.0
        return np.where((s > lo) & (s < hi), val, 0.0)

    g = (np.indices((n, n, n)).reshape(3, -1)[::-1].T + 0.5) / n
    he = float(np.float64(np.float32(h)) * np.float64(np.float32(0.3)))
    w = orc.lic_weights(lut, F, B, f64=True)
    acc = Nz(g) * w[0]
    for sgn, steps, off in ((-1, B, 1), (1, F, 1 + B)):
        x = g.copy()
        s = F3(x)
        for i in range(steps):
            d1 = (2 * s - 1) * sgn * he
            d2 = (2 * F3(x + d1) - 1) * sgn * he
            x = x + 0.5 * (d1 + d2)
            s = F3(x)
            acc = acc + Nz(x) * w[off + i]
    ref = acc * inv / (F + B) * 30.0
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-12)
    # the gate is neither always open nor always shut here
    assert 0.05 < np.mean(ref > 0) < 0.999


def test_slicing_blend_depth_varying_tf_central_ray():
    """Blended slices along the central ray of a scene whose TF colour
    changes with depth: |v| grows from the front (z = 1) to the back, the
    slicing shaders index the TF with .a = |v|/max, the front half of the TF
    is dim (tf.a 0.1) and the back half bright (rgb up to 40, tf.a 10).  The
    bright fragments' colour exceeds 1 and is clamped before blending, and
    the first one is opaque, so its clamped colour is the final one.  Slice
    k of 129 crosses the central ray at z_k = 1 - (k + 1/2)/129.  Numpy from
    the GL blend equation, within one byte step (the oracle's slice depths
    are computed in float)."""
    n = 8
    m = 1.0 - 0.9 * np.arange(n) / (n - 1)  # per z plane, larger at the back (low z)
    v = np.zeros((n, n, n, 3), np.float32)
    v[...] = m[:, None, None, None]
    f4 = orc.encode_field(v, (n, n, n), f64=True)
    amag = f4.reshape(n, n, n, 4)[:, 0, 0, 3]
    noise = np.full(n ** 3, 255, np.uint8)
    lut = np.full(256, 255, np.uint8)
    p = orc.params(steps_forward=2, steps_backward=2, inv_filter_area=0.5, gradient_scale=1.0,
                   step_size_vol=1.0 / 128.0)
    tf = np.zeros((256, 4), np.float64)
    tf[:128] = (0.2, 0.1, 0.05, 0.1)
    tf[128:] = (40.0, 1.0, 0.2, 10.0)  # alpha 0.1 * 10: the first bright fragment is opaque
    op = np.full(256, 0.1, np.float64)
    view = orc.view_setup(15, 15)
    img, ns, _ = orc.render_slices_blend(f4, (n,) * 3, noise, 1, (n,) * 3, lut, tf.ravel(), op, p, view)
    px = img.reshape(-1, 4)[7 * 15 + 7].astype(int)
    lic = 1.0 * 5 * 0.5 / 4 * 1.0
    centres = (np.arange(256) + 0.5) / 256
    d = np.zeros(4)
    clamped_bright = False
    for k in range(129):
        z = 1.0 - (k + 0.5) / 129
        u = np.interp(z * n - 0.5, np.arange(n), amag)
        t4 = np.array([np.interp(u, centres, tf[:, c]) for c in range(4)])
        if not t4[3] > 0.05 or d[3] == 255:
            continue
        a = np.interp(1.3 * lic, centres, op) * t4[3]
        src = np.concatenate([lic * t4[:3] * a, [a]])
        clamped_bright |= src[0] > 1.0 and d[0] < d[3]
        s = np.clip(src, 0.0, 1.0)
        d = np.floor(np.clip(s * (1.0 - d[3] / 255.0) + d / 255.0, 0.0, 1.0) * 255.0 + 0.5)
    d = np.floor(np.clip(1.0 - d[3] / 255.0 + d / 255.0, 0.0, 1.0) * 255.0 + 0.5)
    assert clamped_bright  # the case exercises the fragment clamp
    assert np.max(np.abs(px - d.astype(int))) <= 1, (px, d)
