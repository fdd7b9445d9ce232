"""Line-illumination tables (Zoeckler et al., Mallo et al.): the oracle's
tables against closed forms and an independent Simpson rule, and the
library's tables (host code) against the oracle's, bit for bit."""
import numpy as np
import pytest
from scipy import integrate

import vectorvisualization_amd as vv
from vectorvisualization_amd import _lib as L
from oracle import oracle as orc


def _half(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float16).astype(np.float32)


def _close_in_halves(a, b):
    """equal up to one FP16 step (the references here are evaluated with a
    different but equivalent double expression)"""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    step = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(np.float16)).astype(np.float32)
    return np.all(np.abs(a - b) <= step)


@pytest.fixture(scope="module")
def luts():
    return [t.copy() for t in orc.illum_luts()]


def test_zoeckler_closed_forms(luts):
    z = luts[0].reshape(256, 256, 2)
    x = np.arange(256)
    lt = 2.0 * x / 255.0 - 1.0
    # diffuse: (0.1 + (1 - lt^2)^(diffExp/2) * 0.5) * 0.5, for every row
    ref = (np.float32(0.1) + (1.0 - lt * lt) * 0.5) * 0.5
    for y in (0, 77, 255):
        assert _close_in_halves(z[y, :, 0], ref)
    # edges: only the ambient term, 0.1 * 0.5
    assert np.all(z[:, 0, 0] == _half(0.05)) and np.all(z[:, 255, 0] == _half(0.05))
    # on the anti-diagonal lt = -vt the dot product is -1: specular 0.8 * 0.5 * 0.9
    anti = z[255 - x, x, 1]
    assert np.all(anti == _half(np.float32(0.4) * np.float32(0.9)))
    # specular is |cos(a + b)|^40 with cos a = lt, cos b = vt
    vt = lt
    a = np.arccos(lt)[None, :]
    b = np.arccos(vt)[:, None]
    spec = np.clip(np.abs(np.cos(a + b)) ** 40 * 0.8 * 0.5, 0, 1).astype(np.float32) * np.float32(0.9)
    assert _close_in_halves(z[..., 1], spec)


def test_mallo_diffuse_closed_form(luts):
    md = luts[1].reshape(256, 256, 3)
    s = (np.arange(256) + 0.5) / 256
    alpha = np.arccos(2 * s - 1)[None, :]
    t = s[:, None]
    fd = np.sqrt(1 - (2 * t - 1) ** 2) * (np.sin(alpha) + (np.pi - alpha) * np.cos(alpha)) / 4
    ref = np.clip(fd * 0.5, 0, 1)
    for c in range(3):
        assert _close_in_halves(md[..., c], ref)
    # at alpha = pi/2 (s = 1/2 falls between texels): F_d = sqrt(1-lt^2)/4 * 0.5 varies smoothly
    assert np.all(md[..., 0] >= 0) and md[..., 0].max() <= 0.5 * (np.pi / 4) + 1e-3


def test_mallo_specular_simpson_matches_scipy():
    rng = np.random.default_rng(0)
    for _ in range(20):
        alpha, beta = rng.uniform(0.01, np.pi - 0.01, 2)
        n = float(rng.choice([2.0, 10.0, 40.0]))
        a, b = alpha - np.pi / 2, np.pi / 2
        th = np.linspace(a, b, 21)
        f = np.maximum(np.cos(th - beta), 0.0) ** n * np.cos(th) / 2
        ref = integrate.simpson(f, x=th)
        assert orc.mallo_spec_term(alpha, beta, n) == pytest.approx(ref, rel=1e-12, abs=1e-15)
    # for a smooth integrand the rule is close to the integral itself
    alpha, beta = 1.1, 0.7
    exact, _ = integrate.quad(lambda th: max(np.cos(th - beta), 0.0) ** 2 * np.cos(th) / 2,
                              alpha - np.pi / 2, np.pi / 2)
    assert orc.mallo_spec_term(alpha, beta, 2.0) == pytest.approx(exact, rel=1e-5)


def test_mallo_specular_table(luts):
    ms = luts[2].reshape(256, 256, 3)
    for (y, x) in [(0, 0), (40, 200), (128, 128), (255, 17)]:
        s, t = (x + 0.5) / 256, (y + 0.5) / 256
        v = 3.5 * orc.mallo_spec_term(np.arccos(2 * s - 1), np.arccos(2 * t - 1), 40.0) * 0.8
        assert ms[y, x, 0] == _half(np.clip(v, 0, 1))


def test_tables_are_halves(luts):
    for t in luts:
        assert np.array_equal(t, t.astype(np.float16).astype(np.float32))


def test_library_tables_equal_oracle(luts):
    got = L.make_illum_tables()
    for a, b in zip(got, luts):
        np.testing.assert_array_equal(a, b)
    m = vv.default_line_material()
    m.diffuse[1] = 0.25
    m.specular[2] = 0.3
    m.ambient[0] = 0.2
    m.spec_exp = 12.0
    m.diff_exp = 3.0
    om = orc.line_mat(diffuse=(0.5, 0.25, 0.5, 0.5), specular=(0.8, 0.8, 0.3, 0.8),
                      ambient=(0.2, 0.1, 0.1, 0.1), spec_exp=12.0, diff_exp=3.0)
    for a, b in zip(L.make_illum_tables(m), orc.illum_luts(om)):
        np.testing.assert_array_equal(a, b)


def test_material_validation():
    m = vv.default_line_material()
    m.spec_exp = -1.0
    with pytest.raises(vv.VVError):
        L.make_illum_tables(m)
    m = vv.default_line_material()
    m.diffuse[0] = float("nan")
    with pytest.raises(vv.VVError):
        L.make_illum_tables(m)
