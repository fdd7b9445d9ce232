"""Independent pins of the oracle's shading (illumLIC, illumGradient,
illumMallo, illumZoeckler) and of its 4-channel (RGBA noise) LIC.

Shading: a one-sample scene.  The field and the LIC volume are constant, so
every trilinear lookup returns the constant; the view is 1x1, whose one ray
is the pixel-centre ray d = dw of the view; the sample distance (10) is so
large that the ray leaves the box after its first sample.  The pixel then
holds clamp(src) of that sample.  The expected value is numpy code written
from VV/shader/inc_illum.glsl:1-172 (double precision), with the transfer
function and opacity tables looked up by np.interp over texel centres and
the line-illumination tables by scipy.ndimage.map_coordinates (GL LINEAR +
CLAMP_TO_EDGE), both independent of the oracle's own lookups.  The tables
themselves are pinned by tests/test_illum_tables.py.

RGBA LIC: a brute force on scipy.ndimage.map_coordinates following
VV/shader/inc_lic.glsl:61-68,152-202 (freqSamplingGrad: the RGBA noise at
the position itself) and VV/shader/lic3d_fragment.glsl:61-67 (rgb kept as
convolved, alpha scaled by licKernel.b * gradient.r).
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as orc

_CENTRES = (np.arange(256) + 0.5) / 256.0
_LUTS = None


def _luts():
    global _LUTS
    if _LUTS is None:
        _LUTS = orc.illum_luts()
    return _LUTS


def _f32(x):
    return np.asarray(np.float32(x), np.float64)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _tf_lookup(table, u):
    """texture1D with LINEAR + CLAMP_TO_EDGE on a 256-entry table column"""
    return float(np.interp(u, _CENTRES, table))


def _tex2d(table, nc, ch, s, t):
    """texture2D(table, (s, t)).ch with LINEAR + CLAMP_TO_EDGE on a 256x256
    table stored row by row (t selects the row, s the column)"""
    img = np.asarray(table, np.float64).reshape(256, 256, nc)[..., ch]
    c = np.array([[t * 256.0 - 0.5], [s * 256.0 - 0.5]])
    return float(ndimage.map_coordinates(img, c, order=1, mode="nearest")[0])


# ---- the shading functions of inc_illum.glsl, restated in numpy ------------
def _alpha(l, tf, op, corr):
    a = _tf_lookup(op, l * 1.3) * tf[3]                 # :35 / :105 / :148 / :166
    return 1.0 - (1.0 - a) ** corr                       # :38 / :108 / :151 / :169


def illum_lic(l, tf, op, g, corr):                     # inc_illum.glsl:158-172
    return l * tf[:3] * g, _alpha(l, tf, op, corr)


def illum_gradient(illum, tf, op, g, corr, light, dirv, diff, spec, amb, n):  # :1-41
    L = _unit(light)                                   # :8, scaleVolInv = 0
    V = _unit(-dirv)                                   # :9
    N = _unit(-illum[:3])                              # :11
    R = _unit(2.0 * np.dot(L, N) * N - L)              # :13
    s = np.clip(np.dot(R, V), 0.0, 1.0) ** n           # :15-16
    specular = s * illum[3] * spec                     # :18
    color = tf[:3] * illum[3]                          # :22 mix(0, tf, illum.a)
    diffuse = np.clip(np.dot(L, N), 0.0, 1.0) * diff * g  # :24-26
    color = color * (diffuse + 0.3 + amb) + specular   # :29
    color = color * g                                  # :31
    return color, _alpha(illum[3], tf, op, corr)


def illum_mallo(l, tf, op, g, corr, light, dirv, tangent, n, luts):  # :46-113
    _, mdiff, mspec = luts
    L = _unit(light)                                   # :62
    V = _unit(-dirv)                                   # :63
    T = _unit(2.0 * tangent - 1.0)                     # :65
    B = _unit(np.cross(T, V))                          # :67
    N = np.cross(B, T)                                 # :68
    H = _unit(V + L)                                   # :70
    lt = np.array([np.dot(L, N), np.dot(L, T), np.dot(H, N), np.dot(H, T)])  # :72-75
    tmp = np.array([1.0 / np.sqrt(1.0 - lt[1] ** 2), 1.0 / np.sqrt(1.0 - lt[3] ** 2)])  # :78-79
    lt[2], lt[3] = lt[0] * tmp[0], lt[2] * tmp[1]      # :80 lt.zw = lt.xz * tmp
    lt = 0.5 * lt + 0.5                                # :82
    diffuse = np.array([_tex2d(mdiff, 3, c, lt[2], lt[1]) for c in range(3)])   # :85 lt.zy
    specular = np.array([_tex2d(mspec, 3, c, lt[2], lt[3]) for c in range(3)])  # :87 lt.zw
    att = tmp[1] ** (-float(n))                        # :91
    specular = np.clip(specular * att, 0.0, 1.0)       # :92
    diffuse = diffuse * g                              # :96
    color = tf[:3] * l * diffuse + specular            # :99-102
    return color, _alpha(l, tf, op, corr)


def illum_zoeckler(l, tf, op, g, corr, light, dirv, tangent, luts):  # :117-154
    zo = luts[0]
    L = _unit(light)                                   # :124
    V = _unit(dirv)                                    # :125, +dir
    T = _unit(2.0 * tangent - 1.0)                     # :127
    s, t = 0.5 * np.dot(L, T) + 0.5, 0.5 * np.dot(V, T) + 0.5  # :129-131
    # LUMINANCE_ALPHA texture: .rg = (L, L), both from the luminance channel
    lum = _tex2d(zo, 2, 0, s, t)                       # :133
    shading_r, shading_g = lum * g, lum                # :135
    color = tf[:3] * l * shading_r + 0.9 * shading_g   # :141-145
    return color, _alpha(l, tf, op, corr)


# ---- one-sample scenes -----------------------------------------------------
def _one_sample(mode, field_rgba, lic, tf, op, light, quat, g, corr, tf_index=0, diff=1.0, spec=1.0,
                amb=0.0, n=40):
    dims = (2, 2, 2)
    f4 = np.tile(np.asarray(field_rgba, np.float64), 8)
    lic = np.atleast_1d(np.asarray(lic, np.float64))
    lc = lic.size
    licv = np.tile(lic, 8)
    p = orc.params(illum_mode=mode, step_size_vol=10.0, illum_scale=g, tf_index=tf_index,
                   light=tuple(light), light_diffuse=diff, light_specular=spec, light_ambient=amb,
                   spec_exponent=n)
    p.alpha_correction = corr
    if mode >= 2:
        orc.attach_luts(p, _luts())
    q = _unit(quat)
    v = orc.view_setup(1, 1, quat=q, center=(0.5, 0.5, 0.5))
    img, ns, _ = orc.raycast(f4, dims, licv, lc, dims, tf.ravel(), op, p, v, f64=True)
    assert ns[0] == 1
    return img, _unit(np.array(list(v.dw), np.float64))


def _composite_one(color, a):
    # src.rgb *= src.a; dest = clamp((1 - 0) * src + 0, 0, 1)
    return np.clip(np.concatenate([color * a, [a]]), 0.0, 1.0)


def _geometry(rng):
    quat = rng.normal(size=4)
    light = _f32(rng.normal(size=3) + np.array([0.0, 0.0, 1.5]))
    tf = _f32(rng.uniform(0.05, 0.5, (256, 4)))
    op = _f32(rng.uniform(0.2, 0.9, 256))
    u = float(_f32(rng.uniform(0.02, 0.98)))
    g = float(_f32(rng.choice([1.0, 1.25, 0.75])))
    corr = float(_f32(rng.choice([1.0, 0.5, 2.0, 0.7])))
    return quat, light, tf, op, u, g, corr


@pytest.mark.parametrize("seed", range(6))
def test_illum_lic_one_sample(seed):
    rng = np.random.default_rng(100 + seed)
    quat, light, tf, op, u, g, corr = _geometry(rng)
    l = float(rng.uniform(0.2, 0.7))
    img, _ = _one_sample(0, [0.3, 0.7, u, 0.5], l, tf, op, light, quat, g, corr)
    tfs = np.array([_tf_lookup(tf[:, c], u) for c in range(4)])
    color, a = illum_lic(l, tfs, op, g, corr)
    np.testing.assert_allclose(img, _composite_one(color, a), rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("seed", range(10))
def test_illum_gradient_one_sample(seed):
    rng = np.random.default_rng(200 + seed)
    quat, light, tf, op, u, g, corr = _geometry(rng)
    dirv = _one_sample(0, [0.5, 0.5, u, 0.5], 0.5, tf, op, light, quat, g, corr)[1]
    L, V = _unit(light), -dirv
    # normals near the half vector give a specular term well above zero; the
    # odd seeds move further off it
    N = _unit(_unit(L + V) + (0.05 if seed % 2 == 0 else 0.6) * rng.normal(size=3))
    lic = np.concatenate([-N * rng.uniform(0.2, 2.0), [rng.uniform(0.3, 0.8)]])
    diff, spec, amb = (float(_f32(x)) for x in (0.8, 0.6, 0.1))
    img, d2 = _one_sample(1, [0.5, 0.5, u, 0.5], lic, tf, op, light, quat, g, corr, diff=diff, spec=spec,
                          amb=amb)
    tfs = np.array([_tf_lookup(tf[:, c], u) for c in range(4)])
    color, a = illum_gradient(lic, tfs, op, g, corr, light, d2, diff, spec, amb, 40)
    exp = _composite_one(color, a)
    np.testing.assert_allclose(img, exp, rtol=1e-9, atol=1e-12)
    assert np.all(exp[:3] > 0) and np.all(exp[:3] < 1)  # no channel saturates
    # the case exercises the term a sign mistake would change
    assert np.dot(_unit(light), N) > 0.05


@pytest.mark.parametrize("seed", range(10))
def test_illum_mallo_one_sample(seed):
    rng = np.random.default_rng(300 + seed)
    quat, light, tf, op, u, g, corr = _geometry(rng)
    tangent = rng.uniform(0.05, 0.95, 3)
    l = float(rng.uniform(0.3, 0.9))
    img, dirv = _one_sample(2, [tangent[0], tangent[1], tangent[2], u], l, tf, op, light, quat, g, corr,
                            tf_index=1)
    tfs = np.array([_tf_lookup(tf[:, c], u) for c in range(4)])
    color, a = illum_mallo(l, tfs, op, g, corr, light, dirv, tangent, 40, _luts())
    np.testing.assert_allclose(img, _composite_one(color, a), rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("seed", range(10))
def test_illum_zoeckler_one_sample(seed):
    rng = np.random.default_rng(400 + seed)
    quat, light, tf, op, u, g, corr = _geometry(rng)
    tangent = rng.uniform(0.05, 0.95, 3)
    tangent[2] = u  # the TF index (.b) is the tangent's third component
    l = float(rng.uniform(0.3, 0.9))
    img, dirv = _one_sample(3, [tangent[0], tangent[1], u, 0.4], l, tf, op, light, quat, g, corr)
    tfs = np.array([_tf_lookup(tf[:, c], u) for c in range(4)])
    color, a = illum_zoeckler(l, tfs, op, g, corr, light, dirv, tangent, _luts())
    np.testing.assert_allclose(img, _composite_one(color, a), rtol=1e-6, atol=1e-9)


def test_line_illumination_view_dependence_as_written():
    """Reversing the view changes a Mallo sample (its binormal and normal
    come from cross(tangent, -dir)) and leaves a Zoeckler sample as it was:
    the Zoeckler texture is LUMINANCE_ALPHA, .rg reads its luminance twice,
    and the luminance is a function of dot(light, tangent) alone, so the
    view coordinate dot(dir, tangent) selects among equal texels."""
    rng = np.random.default_rng(7)
    quat, light, tf, op, u, g, corr = _geometry(rng)
    tangent = np.array([0.8, 0.35, 0.6])
    lum = np.asarray(_luts()[0]).reshape(256, 256, 2)[..., 0]
    assert np.all(lum == lum[0:1, :]) and np.ptp(lum) > 0.1
    for mode, fn in ((2, illum_mallo), (3, illum_zoeckler)):
        outs, dirs = [], []
        for q in ((0.0, 0.0, 0.0, 1.0), (0.0, 1.0, 0.0, 0.0)):  # a half turn about y reverses the view
            img, dirv = _one_sample(mode, [*tangent, u], 0.6, tf, op, light, q, g, corr, tf_index=1)
            dirs.append(dirv)
            tfs = np.array([_tf_lookup(tf[:, c], u) for c in range(4)])
            args = (0.6, tfs, op, g, corr, light, dirv, tangent)
            color, a = fn(*args, 40, _luts()) if mode == 2 else fn(*args, _luts())
            np.testing.assert_allclose(img, _composite_one(color, a), rtol=1e-6, atol=1e-9)
            outs.append(img)
        assert np.dot(dirs[0], dirs[1]) < -0.999
        change = np.max(np.abs(outs[0] - outs[1]))
        assert change > 1e-4 if mode == 2 else change == 0.0


# ---- RGBA-noise LIC ----------------------------------------------------------
def test_rgba_noise_lic_matches_scipy_brute_force():
    n, F, B, h = 8, 5, 7, 0.03
    nd = (6, 5, 7)  # noise grid (x, y, z) other than the field's: REPEAT over its own size
    lut, inv = orc.make_filter(1, 12)  # triangle
    p = orc.params(steps_forward=F, steps_backward=B, step_size_lic=h, inv_filter_area=inv,
                   gradient_scale=3.0)
    rng = np.random.default_rng(21)
    v = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    f4 = orc.encode_field(v, (n, n, n), f64=True).reshape(n, n, n, 4)
    noise4 = rng.integers(0, 256, (nd[2], nd[1], nd[0], 4), dtype=np.uint8)
    out = orc.lic(f4.ravel(), (n, n, n), noise4.ravel(), 4, nd, lut, p).reshape(-1, 4)

    def samp(vol, pts, dims, mode):  # pts (m,3) texture coords x,y,z
        c = np.stack([pts[:, 2] * dims[2], pts[:, 1] * dims[1], pts[:, 0] * dims[0]]) - 0.5
        return ndimage.map_coordinates(vol, c, order=1, mode=mode)

    def F3(pts):
        return np.stack([samp(f4[..., ch], pts, (n, n, n), "nearest") for ch in range(3)], 1)

    def N4(pts):  # freqSamplingGrad: texture3D(noiseSampler, pos), RGBA
        return np.stack([samp(noise4[..., ch].astype(np.float64), pts, nd, "grid-wrap") for ch in range(4)],
                        1) / 255.0

    g = (np.indices((n, n, n)).reshape(3, -1)[::-1].T + 0.5) / n
    he = float(np.float64(np.float32(h)) * np.float64(np.float32(0.3)))
    w = orc.lic_weights(lut, F, B, f64=True)
    acc = N4(g) * w[0]
    for sgn, steps, off in ((-1, B, 1), (1, F, 1 + B)):
        x = g.copy()
        s = F3(x)
        for i in range(steps):
            d1 = (2 * s - 1) * sgn * he
            d2 = (2 * F3(x + d1) - 1) * sgn * he
            x = x + 0.5 * (d1 + d2)
            s = F3(x)
            acc = acc + N4(x) * w[off + i]
    ref = acc.copy()
    ref[:, 3] *= inv / (F + B) * 3.0  # lic3d_fragment.glsl:67; rgb stays as convolved
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-12)
    assert np.max(np.abs(ref[:, :3])) > 1.0  # unscaled rgb sums of several noise texels
