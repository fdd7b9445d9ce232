"""Shared small workloads for the parity tests: the same inputs are handed to
the HIP path (through the C ABI) and to the oracle."""
from __future__ import annotations

import numpy as np

import vectorvisualization_amd as vv
from vectorvisualization_amd import fields
from oracle import oracle as orc


def field(kind: str, n: int, seed: int = 1) -> np.ndarray:
    if kind == "abc":
        return fields.abc_flow(n)
    if kind == "rankine":
        return fields.rankine_vortices(n, seed=seed)
    if kind == "gyre":
        return fields.double_gyre_3d(n, t=1.0)
    if kind == "uniform":
        v = np.zeros((n, n, n, 3), np.float32)
        v[..., 0] = 1.0
        return v
    raise ValueError(kind)


def oracle_noise(seed: int, kind: int, density: float, dims, with_gradients: bool) -> np.ndarray:
    """Noise texture as the engine builds it: 1 channel, or RGBA = (grad.xyz, noise)."""
    n = orc.gen_noise(seed, kind, density, dims)
    if not with_gradients:
        return n
    g = orc.noise_gradients(n, dims).reshape(-1, 3)
    out = np.empty((n.size, 4), np.uint8)
    out[:, :3] = g
    out[:, 3] = n
    return out.ravel()


def oracle_field(v: np.ndarray, dims, storage: int, v1=None, t=0.0) -> np.ndarray:
    """The field array the oracle samples for a storage type: encoded floats
    (FP32), the same rounded to halves (FP16), or UNORM8 byte values (to be
    sampled with params(field_unorm8=1))."""
    if storage == vv.STORE_UNORM8:
        return orc.field_bytes_as_float(orc.encode_field_u8(v, dims, v1=v1, t=t))
    f4 = orc.encode_field(v, dims, v1=v1, t=t)
    if storage == vv.STORE_FP16:
        f4 = orc.f16_to_f32(orc.f32_to_f16(f4))
    return f4


def float_field_params(op):
    """copy of oracle params for an unquantised float field (the FP64 error
    reference of a UNORM8 run)"""
    q = orc.Params.from_buffer_copy(op)
    q.field_unorm8 = 0
    return q


def lic_setup(ctx: vv.Context, *, n=32, kind="abc", noise_kind=vv.NOISE_BINARY, density=0.05,
              grad=False, filt=vv.FILTER_BOX, L=20, h=0.01, storage=vv.STORE_FP32, seed=1,
              gradient_scale=30.0, freq=1.0, gate=None, speed=False):
    """Loads a case into ctx and returns what the oracle needs for the same case."""
    dims = (n, n, n)
    v = field(kind, n, seed)
    ctx.upload_field(v, dims, storage=storage)
    ctx.generate_noise(seed, noise_kind, density, dims, with_gradients=grad)
    lut, inv = vv.make_filter(filt, L)
    ctx.set_filter(lut, inv)
    p = vv.default_lic_params()
    p.steps_forward = L
    p.steps_backward = L
    p.step_size_lic = h
    p.gradient_scale = gradient_scale
    p.freq_scale = freq
    p.illum_mode = vv.ILLUM_GRADIENT if grad else vv.ILLUM_LIC
    scalar = None
    if gate is not None:
        scalar = gate
        ctx.upload_scalar(scalar, dims)
        p.gate_enable = 1
    p.speed_of_flow = 1 if speed else 0
    ctx.set_lic_params(p)
    f4 = oracle_field(v, dims, storage)
    noise = oracle_noise(seed, noise_kind, density, dims, grad)
    op = orc.params(steps_forward=L, steps_backward=L, step_size_lic=h, inv_filter_area=inv,
                    gradient_scale=gradient_scale, freq_scale=freq,
                    illum_mode=1 if grad else 0, gate_enable=1 if gate is not None else 0,
                    field_unorm8=1 if storage == vv.STORE_UNORM8 else 0, speed_of_flow=1 if speed else 0)
    return dict(v=v, dims=dims, f4=f4, noise=noise, nch=4 if grad else 1, lut=lut, inv=inv,
                op=op, params=p, scalar=scalar)


def oracle_lic(case, f64=False, **kw):
    f4 = case["f4"]
    op = case["op"]
    if f64:
        f4 = orc.encode_field(case["v"], case["dims"], f64=True)
        op = float_field_params(op)
    return orc.lic(f4, case["dims"], case["noise"], case["nch"], case["dims"], case["lut"],
                   op, scalar=case["scalar"],
                   sdims=case["dims"] if case["scalar"] is not None else None, **kw)
