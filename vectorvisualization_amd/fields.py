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
                     rmin: float = 0.05, rmax: float = 0.2, device="cpu"):
    """Sum of `count` Rankine vortices in the unit box plus a uniform flow.

    Vortex k has a seeded centre c_k in [0,1]^3, unit axis a_k, core radius
    R_k in [rmin, rmax] and strength s_k in +-[0.5, 1.5].  At distance rho
    from its axis it adds s_k * (rho/R_k if rho < R_k else R_k/rho) along
    a_k x (p - c_k) / |a_k x (p - c_k)|.  Returns a numpy array (computed
    with rankine_vortices_torch on `device`).
    """
    t = rankine_vortices_torch(n, count, seed, background, rmin, rmax, device)
    return t.cpu().numpy()


def rankine_vortices_torch(n: int, count: int = 12, seed: int = 1, background=(0.1, 0.0, 0.0),
                           rmin: float = 0.05, rmax: float = 0.2, device="cpu"):
    """rankine_vortices() as a contiguous (n, n, n, 3) float32 torch tensor on
    `device`: seeded vortex parameters, component-wise float32 arithmetic."""
    import torch

    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 1.0, (count, 3))
    axes = rng.normal(size=(count, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    radii = rng.uniform(rmin, rmax, count)
    strength = rng.uniform(0.5, 1.5, count) * rng.choice([-1.0, 1.0], count)
    t = torch.tensor(_centres(n, 0.0, 1.0), dtype=torch.float32, device=device)
    z = t.view(n, 1, 1)
    y = t.view(1, n, 1)
    x = t.view(1, 1, n)
    out = torch.empty((n, n, n, 3), dtype=torch.float32, device=device)
    u = torch.full((n, n, n), float(background[0]), dtype=torch.float32, device=device)
    v = torch.full((n, n, n), float(background[1]), dtype=torch.float32, device=device)
    w = torch.full((n, n, n), float(background[2]), dtype=torch.float32, device=device)
    for k in range(count):
        ax, ay, az = (float(a) for a in axes[k])
        dx, dy, dz = x - float(centres[k, 0]), y - float(centres[k, 1]), z - float(centres[k, 2])
        cx = ay * dz - az * dy
        cy = az * dx - ax * dz
        cz = ax * dy - ay * dx
        rho = torch.sqrt(cx * cx + cy * cy + cz * cz)
        R = float(radii[k])
        mag = torch.where(rho < R, rho / R, R / torch.clamp(rho, min=1e-12))
        sc = float(strength[k]) * mag / torch.clamp(rho, min=1e-12)
        sc = torch.where(rho > 1e-12, sc, torch.zeros_like(sc))
        u += sc * cx
        v += sc * cy
        w += sc * cz
    out[..., 0] = u
    out[..., 1] = v
    out[..., 2] = w
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


def double_gyre_3d_torch(n: int, t: float, A: float = 0.1, eps: float = 0.25,
                         omega: float = 2.0 * np.pi / 10.0, rot: float = 0.5, device="cuda"):
    """double_gyre_3d evaluated with torch on `device` (float64 math, float32
    result of shape (n, n, n, 3)), for volumes too large to build quickly on
    the host."""
    import torch

    c = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) / n
    z = c.view(n, 1, 1)
    y = c.view(1, n, 1)
    x = c.view(1, 1, n)
    th = rot * t
    ct, st = float(np.cos(th)), float(np.sin(th))
    xr = ct * (x - 0.5) + st * (y - 0.5) + 0.5
    yr = -st * (x - 0.5) + ct * (y - 0.5) + 0.5
    X = 2.0 * xr
    a = eps * float(np.sin(omega * t))
    b = 1.0 - 2.0 * a
    f = a * X * X + b * X
    dfdx = 2.0 * a * X + b
    u = -np.pi * A * torch.sin(np.pi * f) * torch.cos(np.pi * yr)
    v = np.pi * A * torch.cos(np.pi * f) * torch.sin(np.pi * yr) * dfdx
    w = 0.25 * A * np.pi * torch.sin(2.0 * np.pi * z) * torch.cos(np.pi * yr)
    out = torch.empty((n, n, n, 3), dtype=torch.float32, device=device)
    out[..., 0] = ct * u - st * v
    out[..., 1] = st * u + ct * v
    out[..., 2] = w.expand(n, n, n)
    return out


def tf_linear_ramp():
    """256-entry RGBA + opacity ramp in the LIC value (rgb = a = opacity = i/255)."""
    r = np.arange(256, dtype=np.float32) / 255.0
    rgba = np.stack([r, r, r, r], 1).astype(np.float32).ravel()
    return rgba, r.copy()


def tf_magnitude(alpha_lo: float = 0.02, alpha_hi: float = 0.12):
    """"Vector-length" TF for an index = |v|/max|v|: rgb runs blue -> green
    -> red with the magnitude, alpha ramps from alpha_lo to alpha_hi; the
    LIC-opacity table is the linear ramp i/255."""
    t = np.arange(256, dtype=np.float64) / 255.0
    rgba = np.empty((256, 4), np.float64)
    rgba[:, 0] = np.clip(2.0 * t - 0.5, 0.0, 1.0)
    rgba[:, 1] = 1.0 - np.abs(2.0 * t - 1.0)
    rgba[:, 2] = np.clip(1.5 - 2.0 * t, 0.0, 1.0)
    rgba[:, 3] = alpha_lo + (alpha_hi - alpha_lo) * t
    return rgba.astype(np.float32).ravel(), t.astype(np.float32)


def tf_random(seed: int):
    """Seeded random 256x4 RGBA TF in [0,1] and a fixed opacity ramp."""
    rng = np.random.default_rng(seed)
    rgba = rng.uniform(0.0, 1.0, (256, 4)).astype(np.float32).ravel()
    op = (np.arange(256, dtype=np.float32) / 255.0)
    return rgba, op


def orbit_camera(width: int, height: int, center, yaw: float, pitch: float = 0.0, dist: float = 4.0):
    """Camera on an orbit around `center`: rotation by `yaw` about the y axis
    followed by `pitch` about the x axis (radians), at distance `dist`; the
    light keeps the default offset along the view axis."""
    from ._lib import default_camera

    cam = default_camera(width, height, center)
    qy = np.array([0.0, np.sin(0.5 * yaw), 0.0, np.cos(0.5 * yaw)])
    qx = np.array([np.sin(0.5 * pitch), 0.0, 0.0, np.cos(0.5 * pitch)])
    # Hamilton product qy * qx, (x, y, z, w) order
    x1, y1, z1, w1 = qy
    x2, y2, z2, w2 = qx
    q = (w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
         w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
         w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2)
    for i in range(4):
        cam.quat[i] = float(q[i])
    cam.dist = float(dist)
    return cam
