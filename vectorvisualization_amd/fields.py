"""Synthetic vector fields of the benchmark workloads (host numpy, float32).

Each generator returns an array of shape (n, n, n, 3) in z, y, x order, i.e.
the x-fastest interleaved layout of a DAT/RAW FLOAT3 volume
(VV/reader.cpp:290-295), sampled at voxel centres.
"""
from __future__ import annotations

import numpy as np


def _centres(n: int, lo: float, hi: float) -> np.ndarray:
    return lo + (np.arange(n, dtype=np.float64) + 0.5) / n * (hi - lo)


def abc_flow(n: int, A: float = np.sqrt(3.0), B: float = np.sqrt(2.0), C: float = 1.0) -> np.ndarray:
    """Arnold-Beltrami-Childress flow on [0, 2pi]^3:
    u = A sin z + C cos y, v = B sin x + A cos z, w = C sin y + B cos x."""
    t = _centres(n, 0.0, 2.0 * np.pi)
    z, y, x = np.meshgrid(t, t, t, indexing="ij")
    out = np.empty((n, n, n, 3), np.float32)
    out[..., 0] = A * np.sin(z) + C * np.cos(y)
    out[..., 1] = B * np.sin(x) + A * np.cos(z)
    out[..., 2] = C * np.sin(y) + B * np.cos(x)
    return out


def rankine_vortices(n: int, count: int = 12, seed: int = 1, background=(0.1, 0.0, 0.0),
                     rmin: float = 0.05, rmax: float = 0.2, chunk: int = 32) -> np.ndarray:
    """Sum of `count` Rankine vortices in the unit box plus a uniform flow.

    Vortex k has a seeded centre c_k in [0,1]^3, unit axis a_k, core radius
    R_k in [rmin, rmax] and strength s_k in +-[0.5, 1.5].  At distance rho
    from its axis it adds s_k * (rho/R_k if rho < R_k else R_k/rho) along
    a_k x (p - c_k) / |a_k x (p - c_k)|.
    """
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (count, 3))
    axes = rng.normal(size=(count, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    radii = rng.uniform(rmin, rmax, count)
    strength = rng.uniform(0.5, 1.5, count) * rng.choice([-1.0, 1.0], count)
    t = _centres(n, 0.0, 1.0)
    out = np.empty((n, n, n, 3), np.float32)
    y2, x2 = np.meshgrid(t, t, indexing="ij")
    for z0 in range(0, n, chunk):
        zs = t[z0:z0 + chunk]
        p = np.stack(np.broadcast_arrays(zs[:, None, None], y2[None], x2[None]), -1)[..., ::-1]
        v = np.zeros(p.shape, np.float64)
        v[...] = background
        for k in range(count):
            d = p - centres[k]
            cr = np.cross(axes[k], d)
            rho = np.linalg.norm(cr, axis=-1)
            R = radii[k]
            mag = np.where(rho < R, rho / R, R / np.maximum(rho, 1e-12))
            with np.errstate(invalid="ignore", divide="ignore"):
                unit = np.where(rho[..., None] > 1e-12, cr / rho[..., None], 0.0)
            v += strength[k] * mag[..., None] * unit
        out[z0:z0 + chunk] = v
    return out


def double_gyre_3d(n: int, t: float, A: float = 0.1, eps: float = 0.25,
                   omega: float = 2.0 * np.pi / 10.0, rot: float = 0.5) -> np.ndarray:
    """Double gyre extended to 3D, rotating about the box's z axis.

    In the (x', y') frame rotated by rot*t about (0.5, 0.5) the in-plane
    flow is the classic double gyre on [0,2]x[0,1] (x' stretched by 2):
    psi = A sin(pi f) sin(pi y'), f = eps sin(omega t) x'^2 + (1 - 2 eps sin(omega t)) x',
    u = -dpsi/dy', v = dpsi/dx'.  The z component is
    w = 0.25 A pi sin(2 pi z) cos(pi y'), so the gyres wobble in z.
    """
    c = _centres(n, 0.0, 1.0)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    th = rot * t
    ct, st = np.cos(th), np.sin(th)
    xr = ct * (x - 0.5) + st * (y - 0.5) + 0.5
    yr = -st * (x - 0.5) + ct * (y - 0.5) + 0.5
    X = 2.0 * xr
    a = eps * np.sin(omega * t)
    b = 1.0 - 2.0 * a
    f = a * X * X + b * X
    dfdx = 2.0 * a * X + b
    u = -np.pi * A * np.sin(np.pi * f) * np.cos(np.pi * yr)
    v = np.pi * A * np.cos(np.pi * f) * np.sin(np.pi * yr) * dfdx
    w = 0.25 * A * np.pi * np.sin(2.0 * np.pi * z) * np.cos(np.pi * yr)
    out = np.empty((n, n, n, 3), np.float32)
    out[..., 0] = ct * u - st * v
    out[..., 1] = st * u + ct * v
    out[..., 2] = w
    return out


def tf_linear_ramp():
    """256-entry RGBA + opacity ramp in the LIC value (rgb = a = opacity = i/255)."""
    r = np.arange(256, dtype=np.float32) / 255.0
    rgba = np.stack([r, r, r, r], 1).astype(np.float32).ravel()
    return rgba, r.copy()


def tf_random(seed: int):
    """Seeded random 256x4 RGBA TF in [0,1] and a fixed opacity ramp."""
    rng = np.random.default_rng(seed)
    rgba = rng.uniform(0.0, 1.0, (256, 4)).astype(np.float32).ravel()
    op = (np.arange(256, dtype=np.float32) / 255.0)
    return rgba, op
