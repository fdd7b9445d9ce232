"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product
(vectorvisualization_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
FAST_LIB_PATH = HERE / "liboracle_fast.so"


def build() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


class Params(C.Structure):
    """orc_params (oracle/vvoracle.h)."""

    _fields_ = [
        ("step_size_vol", C.c_float),
        ("gradient_scale", C.c_float),
        ("illum_scale", C.c_float),
        ("freq_scale", C.c_float),
        ("num_iterations", C.c_int),
        ("steps_forward", C.c_int),
        ("steps_backward", C.c_int),
        ("step_size_lic", C.c_float),
        ("eye_factor", C.c_float),
        ("illum_mode", C.c_int),
        ("tf_index", C.c_int),
        ("gate_enable", C.c_int),
        ("gate_lo", C.c_float),
        ("gate_hi", C.c_float),
        ("inv_filter_area", C.c_float),
        ("alpha_correction", C.c_float),
        ("light", C.c_float * 3),
        ("light_diffuse", C.c_float),
        ("light_specular", C.c_float),
        ("light_ambient", C.c_float),
        ("spec_exponent", C.c_int),
        ("lut_zoeckler", C.c_void_p),
        ("lut_mallo_diff", C.c_void_p),
        ("lut_mallo_spec", C.c_void_p),
        ("field_unorm8", C.c_int),
        ("speed_of_flow", C.c_int),
    ]


class LineMat(C.Structure):
    """orc_line_mat (LineMat, VV/illumination.h:40-62)."""

    _fields_ = [
        ("light_color", C.c_float * 4),
        ("ambient", C.c_float * 4),
        ("diffuse", C.c_float * 4),
        ("specular", C.c_float * 4),
        ("spec_exp", C.c_float),
        ("diff_exp", C.c_float),
    ]


def line_mat(**kw) -> LineMat:
    """The reference's LineMat defaults, overridden by keyword arguments."""
    m = LineMat()
    for i in range(4):
        m.light_color[i] = 1.0
        m.ambient[i] = 0.1
        m.diffuse[i] = 0.5
        m.specular[i] = 0.8
    m.spec_exp = 40.0
    m.diff_exp = 2.0
    for k, v in kw.items():
        if isinstance(getattr(m, k), (int, float)):
            setattr(m, k, v)
        else:
            for i in range(4):
                getattr(m, k)[i] = v[i]
    return m


class View(C.Structure):
    _fields_ = [
        ("cam", C.c_float * 3),
        ("du", C.c_float * 3),
        ("dv", C.c_float * 3),
        ("dw", C.c_float * 3),
        ("inv_w2", C.c_float),
        ("inv_h2", C.c_float),
        ("extent", C.c_float * 3),
        ("scale", C.c_float * 3),
        ("tex_max", C.c_float * 3),
        ("width", C.c_int),
        ("height", C.c_int),
    ]

    def as_array(self) -> np.ndarray:
        vals = (list(self.cam) + list(self.du) + list(self.dv) + list(self.dw)
                + [self.inv_w2, self.inv_h2] + list(self.extent) + list(self.scale)
                + list(self.tex_max) + [0.0])
        return np.array(vals, np.float32)


def _load(path: Path = LIB_PATH) -> C.CDLL:
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
    L = C.c_long
    sig = {
        "orc_f32_to_f16": (C.c_uint16, [F]),
        "orc_illum_luts": (None, [C.POINTER(LineMat), P, P, P]),
        "orc_mallo_spec_term": (D, [D, D, D]),
        "orc_f16_to_f32": (F, [C.c_uint16]),
        "orc_f32_to_f16_array": (None, [P, P, L]),
        "orc_f16_to_f32_array": (None, [P, P, L]),
        "orc_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_gen_noise": (None, [C.c_uint64, I, D, I, I, I, P]),
        "orc_noise_gradients": (I, [P, I, I, I, P, P]),
        "orc_noise_gradients_raw": (None, [P, I, I, I, P, P]),
        "orc_filter_gradients": (None, [I, I, I, P]),
        "orc_quantize_gradients": (None, [L, P, P]),
        "orc_make_filter": (I, [I, I, D, P, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
        "orc_filter_inv_area": (F, [P, I]),
        "orc_view_setup": (None, [D, I, I, D, P, P, P, P, P, C.POINTER(View)]),
        "orc_light_position": (None, [P, P, P, D, P, P]),
        "orc_encode_field_f32": (None, [P, P, F, I, I, I, P]),
        "orc_encode_field_u8": (None, [P, P, F, I, I, I, P]),
        "orc_encode_field_f64": (None, [P, P, D, I, I, I, P]),
        "orc_encode_field_raw_f32": (None, [P, P, F, I, I, I, P]),
        "orc_encode_field_raw_f64": (None, [P, P, D, I, I, I, P]),
        "orc_lic_weights_f32": (None, [P, I, I, I, P]),
        "orc_lic_weights_f64": (None, [P, I, I, I, P]),
        "orc_lic_f32": (None, [P, I, I, I, P, P, I, P, P, P, P, I, C.POINTER(Params), I, I, I, P]),
        "orc_lic_f64": (None, [P, I, I, I, P, P, I, P, P, P, P, I, C.POINTER(Params), I, I, I, P]),
        "orc_lic_box_f32": (None, [P, I, I, I, P, P, I, P, P, P, P, I, C.POINTER(Params), P, I, P]),
        "orc_lic_box_f64": (None, [P, I, I, I, P, P, I, P, P, P, P, I, C.POINTER(Params), P, I, P]),
        "orc_raycast_fused_f32": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                            C.POINTER(View), I, P, P]),
        "orc_raycast_fused_f64": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                            C.POINTER(View), I, P, P]),
        "orc_render_slices_f32": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                            C.POINTER(View), I, P, P]),
        "orc_render_slices_f64": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                            C.POINTER(View), I, P, P]),
        "orc_render_slices_blend_f32": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                                  C.POINTER(View), I, P, P]),
        "orc_render_slices_blend_f64": (C.c_long, [P, P, P, I, P, P, P, P, I, P, P, P, P, C.POINTER(Params),
                                                  C.POINTER(View), I, P, P]),
        "orc_raycast_f32": (C.c_long, [P, P, P, I, P, P, P, P, C.POINTER(Params), C.POINTER(View),
                                       I, P, P]),
        "orc_raycast_f64": (C.c_long, [P, P, P, I, P, P, P, P, C.POINTER(Params), C.POINTER(View),
                                       I, P, P]),
        "orc_raycast_adjoint_f64": (None, [P, P, P, I, P, P, P, P, C.POINTER(Params),
                                           C.POINTER(View), P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()
_FAST = None


def fast_lib() -> C.CDLL:
    """The -O3 AVX2/FMA build of the same source (the timed CPU baseline)."""
    global _FAST
    if _FAST is None:
        _FAST = _load(FAST_LIB_PATH)
    return _FAST


def _p(a: np.ndarray | None):
    """Address of a C-contiguous array.  The caller must keep `a` referenced
    until the C call returns (no temporaries)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _i3(d):
    return np.ascontiguousarray(np.array(d, np.int32))


def params(**kw) -> Params:
    """Reference defaults (VV/types.h:93-98) overridden by keyword arguments."""
    p = Params()
    p.step_size_vol = 1.0 / 128.0
    p.gradient_scale = 30.0
    p.illum_scale = 1.0
    p.freq_scale = 1.0
    p.num_iterations = 255
    p.steps_forward = 32
    p.steps_backward = 32
    p.step_size_lic = 0.01
    p.eye_factor = 0.3
    p.illum_mode = 0
    p.tf_index = 0
    p.gate_enable = 0
    p.gate_lo = 0.1
    p.gate_hi = 0.3
    p.inv_filter_area = 0.5
    p.light_diffuse = 1.0
    p.light_specular = 1.0
    p.light_ambient = 0.0
    p.spec_exponent = 40
    light = kw.pop("light", (0.5, 0.5, 1.5))
    for k, v in kw.items():
        setattr(p, k, v)
    p.light[0], p.light[1], p.light[2] = light
    # alphaCorrection = stepSizeVol*128 (VV/renderer.cpp:988-989), in float
    p.alpha_correction = float(np.float32(p.step_size_vol) * np.float32(128.0))
    return p


def f32_to_f16(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(a.shape, np.uint16)
    LIB.orc_f32_to_f16_array(_p(a), _p(out), a.size)
    return out


def f16_to_f32(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16)
    out = np.empty(a.shape, np.float32)
    LIB.orc_f16_to_f32_array(_p(a), _p(out), a.size)
    return out


def splitmix64(seed: int, i: int) -> int:
    return int(LIB.orc_splitmix64(C.c_uint64(seed), C.c_uint64(i)))


def gen_noise(seed: int, kind: int, density: float, dims) -> np.ndarray:
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, np.uint8)
    LIB.orc_gen_noise(C.c_uint64(seed), kind, density, nx, ny, nz, _p(out))
    return out


def noise_gradients(noise: np.ndarray, dims, spacing=(1.0, 1.0, 1.0)) -> np.ndarray:
    nx, ny, nz = dims
    noise = np.ascontiguousarray(noise, np.uint8)
    sd = np.array(spacing, np.float32)
    out = np.empty(3 * nx * ny * nz, np.uint8)
    LIB.orc_noise_gradients(_p(noise), nx, ny, nz, _p(sd), _p(out))
    return out


def noise_gradients_raw(noise: np.ndarray, dims, spacing=(1.0, 1.0, 1.0)) -> np.ndarray:
    nx, ny, nz = dims
    noise = np.ascontiguousarray(noise, np.uint8)
    sd = np.array(spacing, np.float32)
    out = np.empty(3 * nx * ny * nz, np.float32)
    LIB.orc_noise_gradients_raw(_p(noise), nx, ny, nz, _p(sd), _p(out))
    return out


def filter_gradients(g: np.ndarray, dims) -> np.ndarray:
    g = np.array(g, np.float32, copy=True)
    LIB.orc_filter_gradients(dims[0], dims[1], dims[2], _p(g))
    return g


def make_filter(kind: int, length: int = 0, sigma: float = 0.0):
    lut = np.zeros(256, np.uint8)
    w = C.c_int()
    inv = C.c_float()
    st = LIB.orc_make_filter(kind, length, sigma, _p(lut), C.byref(w), C.byref(inv))
    if st:
        raise ValueError("bad filter")
    return lut[: w.value].copy(), inv.value


def view_setup(width, height, fovy=35.0, dist=4.0, pos=(0, 0, 0), quat=(0, 0, 0, 1),
               center=(0.5, 0.5, 0.5), extent=(1, 1, 1), scale=(1, 1, 1)) -> View:
    v = View()
    arrs = [np.array(x, np.float64) for x in (pos, quat, center, extent, scale)]
    LIB.orc_view_setup(fovy, width, height, dist, *[_p(a) for a in arrs], C.byref(v))
    return v


def light_position(center=(0.5, 0.5, 0.5), cam_pos=(0, 0, 0), cam_quat=(0, 0, 0, 1),
                   light_dist=1.0, light_quat=(0, 0, 0, 1)) -> np.ndarray:
    out = np.zeros(3, np.float64)
    arrs = [np.array(x, np.float64) for x in (center, cam_pos, cam_quat)]
    lq = np.array(light_quat, np.float64)
    LIB.orc_light_position(_p(arrs[0]), _p(arrs[1]), _p(arrs[2]), light_dist, _p(lq), _p(out))
    return out


def encode_field(v0: np.ndarray, dims, v1=None, t=0.0, f64=False, raw_magnitude=False) -> np.ndarray:
    """fillTexDataFloatInterp; raw_magnitude: fillTexDataFloat (a = |v|)."""
    nx, ny, nz = dims
    v0 = np.ascontiguousarray(v0, np.float32).ravel()
    v1 = None if v1 is None else np.ascontiguousarray(v1, np.float32).ravel()
    out = np.empty(4 * nx * ny * nz, np.float64 if f64 else np.float32)
    if raw_magnitude:
        fn = LIB.orc_encode_field_raw_f64 if f64 else LIB.orc_encode_field_raw_f32
    else:
        fn = LIB.orc_encode_field_f64 if f64 else LIB.orc_encode_field_f32
    fn(_p(v0), _p(v1), t, nx, ny, nz, _p(out))
    return out


def encode_field_u8(v0: np.ndarray, dims, v1=None, t=0.0) -> np.ndarray:
    """UNORM8 field bytes (4 per voxel): fillTexDataCharInterp.  Pass
    ``field_bytes_as_float(...)`` with ``params(field_unorm8=1)`` to the
    sampling functions."""
    nx, ny, nz = dims
    v0 = np.ascontiguousarray(v0, np.float32).ravel()
    v1 = None if v1 is None else np.ascontiguousarray(v1, np.float32).ravel()
    out = np.empty(4 * nx * ny * nz, np.uint8)
    LIB.orc_encode_field_u8(_p(v0), _p(v1), t, nx, ny, nz, _p(out))
    return out


def field_bytes_as_float(b: np.ndarray, f64=False) -> np.ndarray:
    """byte values of a UNORM8 field as the float array the sampling
    functions take (they scale blended values by 1/255 when field_unorm8)"""
    return np.ascontiguousarray(b, np.uint8).astype(np.float64 if f64 else np.float32)


def lic_weights(lut: np.ndarray, F: int, B: int, f64=False) -> np.ndarray:
    lut = np.ascontiguousarray(lut, np.uint8)
    out = np.empty(1 + F + B, np.float64 if f64 else np.float32)
    fn = LIB.orc_lic_weights_f64 if f64 else LIB.orc_lic_weights_f32
    fn(_p(lut), lut.size, F, B, _p(out))
    return out


def lic(field4: np.ndarray, dims, noise: np.ndarray, noise_ch: int, ndims, lut: np.ndarray,
        p: Params, scale=(1, 1, 1), scalar=None, sdims=None, z0=0, z1=None, nthreads=0,
        f64=None, fast=False) -> np.ndarray:
    """LIC volume (slab z0..z1) of an encoded field; precision follows field4's dtype.
    fast: run the -O3 AVX2/FMA build (same results)."""
    nx, ny, nz = dims
    if f64 is None:
        f64 = field4.dtype == np.float64
    field4 = np.ascontiguousarray(field4, np.float64 if f64 else np.float32)
    noise = np.ascontiguousarray(noise, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    z1 = nz if z1 is None else z1
    out = np.zeros((z1 - z0) * ny * nx * noise_ch, np.float64 if f64 else np.float32)
    sc = np.array(scale, np.float32)
    nd = _i3(ndims)
    sd = _i3(sdims) if sdims is not None else None
    scalar = None if scalar is None else np.ascontiguousarray(scalar, np.uint8)
    lib = fast_lib() if fast else LIB
    fn = lib.orc_lic_f64 if f64 else lib.orc_lic_f32
    fn(_p(field4), nx, ny, nz, _p(sc), _p(noise), noise_ch, _p(nd), _p(scalar), _p(sd), _p(lut),
       lut.size, C.byref(p), z0, z1, nthreads, _p(out))
    return out


def lic_box(field4: np.ndarray, dims, noise: np.ndarray, noise_ch: int, ndims, lut: np.ndarray,
            p: Params, box, scale=(1, 1, 1), nthreads=0) -> np.ndarray:
    """LIC of the voxels in box = (x0, x1, y0, y1, z0, z1); precision follows field4."""
    nx, ny, nz = dims
    f64 = field4.dtype == np.float64
    field4 = np.ascontiguousarray(field4)
    noise = np.ascontiguousarray(noise, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    b = np.array(box, np.int32)
    out = np.zeros((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]) * noise_ch,
                   np.float64 if f64 else np.float32)
    fn = LIB.orc_lic_box_f64 if f64 else LIB.orc_lic_box_f32
    sc = np.array(scale, np.float32)
    nd = _i3(ndims)
    fn(_p(field4), nx, ny, nz, _p(sc), _p(noise), noise_ch,
       _p(nd), None, None, _p(lut), lut.size, C.byref(p), _p(b), nthreads, _p(out))
    return out


def raycast(field4, fdims, licv, lic_ch, ldims, tf_rgba, tf_opac, p: Params, v: View,
            sigma=None, nthreads=0, f64=None):
    if f64 is None:
        f64 = field4.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    field4 = np.ascontiguousarray(field4, dt)
    licv = np.ascontiguousarray(licv, dt)
    tf_rgba = np.ascontiguousarray(tf_rgba, dt)
    tf_opac = np.ascontiguousarray(tf_opac, dt)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dt)
    img = np.zeros(v.width * v.height * 4, dt)
    ns = np.zeros(v.width * v.height, np.int32)
    fn = LIB.orc_raycast_f64 if f64 else LIB.orc_raycast_f32
    fd, ld = _i3(fdims), _i3(ldims)  # keep the arrays alive across the call
    total = fn(_p(field4), _p(fd), _p(licv), lic_ch, _p(ld), _p(tf_rgba),
               _p(tf_opac), _p(sigma), C.byref(p), C.byref(v), nthreads, _p(img), _p(ns))
    return img, ns, int(total)


def raycast_fused(field4, fdims, noise, noise_ch, ndims, lut, tf_rgba, tf_opac, p: Params, v: View,
                  scalar=None, sdims=None, sigma=None, sigma_dims=None, nthreads=0, f64=None):
    """Fused ray caster: the LIC of every sample computed on the fly."""
    if f64 is None:
        f64 = field4.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    field4 = np.ascontiguousarray(field4, dt)
    noise = np.ascontiguousarray(noise, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    scalar = None if scalar is None else np.ascontiguousarray(scalar, np.uint8)
    tf_rgba = np.ascontiguousarray(tf_rgba, dt)
    tf_opac = np.ascontiguousarray(tf_opac, dt)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dt)
    img = np.zeros(v.width * v.height * 4, dt)
    ns = np.zeros(v.width * v.height, np.int32)
    fn = LIB.orc_raycast_fused_f64 if f64 else LIB.orc_raycast_fused_f32
    fd, nd = _i3(fdims), _i3(ndims)
    sd = _i3(sdims) if sdims is not None else None
    gd = _i3(sigma_dims) if sigma_dims is not None else None
    total = fn(_p(field4), _p(fd), _p(noise), noise_ch, _p(nd), _p(scalar), _p(sd), _p(lut), lut.size,
               _p(tf_rgba), _p(tf_opac), _p(sigma), _p(gd), C.byref(p), C.byref(v), nthreads, _p(img),
               _p(ns))
    return img, ns, int(total)


def render_slices(field4, fdims, noise, noise_ch, ndims, lut, tf_rgba, tf_opac, p: Params, v: View,
                  scalar=None, sdims=None, sigma=None, sigma_dims=None, nthreads=0, f64=None):
    """View-aligned slicing with the LIC of every slice sample on the fly."""
    if f64 is None:
        f64 = field4.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    field4 = np.ascontiguousarray(field4, dt)
    noise = np.ascontiguousarray(noise, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    scalar = None if scalar is None else np.ascontiguousarray(scalar, np.uint8)
    tf_rgba = np.ascontiguousarray(tf_rgba, dt)
    tf_opac = np.ascontiguousarray(tf_opac, dt)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dt)
    img = np.zeros(v.width * v.height * 4, dt)
    ns = np.zeros(v.width * v.height, np.int32)
    fn = LIB.orc_render_slices_f64 if f64 else LIB.orc_render_slices_f32
    fd, nd = _i3(fdims), _i3(ndims)
    sd = _i3(sdims) if sdims is not None else None
    gd = _i3(sigma_dims) if sigma_dims is not None else None
    total = fn(_p(field4), _p(fd), _p(noise), noise_ch, _p(nd), _p(scalar), _p(sd), _p(lut), lut.size,
               _p(tf_rgba), _p(tf_opac), _p(sigma), _p(gd), C.byref(p), C.byref(v), nthreads, _p(img),
               _p(ns))
    return img, ns, int(total)


def render_slices_blend(field4, fdims, noise, noise_ch, ndims, lut, tf_rgba, tf_opac, p: Params, v: View,
                        scalar=None, sdims=None, sigma=None, sigma_dims=None, nthreads=0, f64=None):
    """The same slices blended into an 8-bit RGBA framebuffer over a white
    background (lic3d_slicingblend_fragment.glsl): (uint8 image, samples, total)."""
    if f64 is None:
        f64 = field4.dtype == np.float64
    dt = np.float64 if f64 else np.float32
    field4 = np.ascontiguousarray(field4, dt)
    noise = np.ascontiguousarray(noise, np.uint8)
    lut = np.ascontiguousarray(lut, np.uint8)
    scalar = None if scalar is None else np.ascontiguousarray(scalar, np.uint8)
    tf_rgba = np.ascontiguousarray(tf_rgba, dt)
    tf_opac = np.ascontiguousarray(tf_opac, dt)
    sigma = None if sigma is None else np.ascontiguousarray(sigma, dt)
    img = np.zeros(v.width * v.height * 4, np.uint8)
    ns = np.zeros(v.width * v.height, np.int32)
    fn = LIB.orc_render_slices_blend_f64 if f64 else LIB.orc_render_slices_blend_f32
    fd, nd = _i3(fdims), _i3(ndims)
    sd = _i3(sdims) if sdims is not None else None
    gd = _i3(sigma_dims) if sigma_dims is not None else None
    total = fn(_p(field4), _p(fd), _p(noise), noise_ch, _p(nd), _p(scalar), _p(sd), _p(lut), lut.size,
               _p(tf_rgba), _p(tf_opac), _p(sigma), _p(gd), C.byref(p), C.byref(v), nthreads, _p(img),
               _p(ns))
    return img, ns, int(total)


def raycast_adjoint(field4, fdims, licv, lic_ch, ldims, tf_rgba, tf_opac, p: Params, v: View,
                    dimage, sigma=None, want_sigma=True):
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)  # noqa: E731
    field4, licv, tf_rgba, tf_opac, sigma, dimage = map(f, (field4, licv, tf_rgba, tf_opac, sigma,
                                                            dimage))
    dtf = np.zeros(1024, np.float64)
    nv = int(np.prod(ldims))
    dsig = np.zeros(nv, np.float64) if want_sigma else None
    fd, ld = _i3(fdims), _i3(ldims)  # keep the arrays alive across the call
    LIB.orc_raycast_adjoint_f64(_p(field4), _p(fd), _p(licv), lic_ch, _p(ld),
                                _p(tf_rgba), _p(tf_opac), _p(sigma), C.byref(p), C.byref(v),
                                _p(dimage), _p(dtf), _p(dsig))
    return dtf, dsig


def illum_luts(mat: LineMat | None = None):
    """(zoeckler 256x256x2, mallo_diff 256x256x3, mallo_spec 256x256x3) float32."""
    m = line_mat() if mat is None else mat
    z = np.empty(256 * 256 * 2, np.float32)
    md = np.empty(256 * 256 * 3, np.float32)
    ms = np.empty(256 * 256 * 3, np.float32)
    LIB.orc_illum_luts(C.byref(m), _p(z), _p(md), _p(ms))
    return z, md, ms


def attach_luts(p: Params, luts) -> None:
    """Point p at illumination tables (the caller keeps `luts` alive)."""
    z, md, ms = luts
    p.lut_zoeckler = z.ctypes.data
    p.lut_mallo_diff = md.ctypes.data
    p.lut_mallo_spec = ms.ctypes.data


def mallo_spec_term(alpha: float, beta: float, n: float) -> float:
    return LIB.orc_mallo_spec_term(alpha, beta, n)
