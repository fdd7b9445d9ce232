"""Host-side workload generators: the torch and numpy versions agree, and
orbit cameras are proper rotations."""
import numpy as np

from vectorvisualization_amd import fields


def test_double_gyre_torch_matches_numpy():
    for t in (0.0, 1.7):
        a = fields.double_gyre_3d(12, t)
        b = fields.double_gyre_3d_torch(12, t, device="cpu").numpy()
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-6)


def test_double_gyre_is_divergence_free_in_plane_at_t0():
    # at t = 0 (no rotation, a = 0) psi = A sin(pi X) sin(pi y) with X = 2x, u = -dpsi/dy,
    # v = dpsi/dX: the flow is divergence-free in (X, y), i.e. du/dx / 2 + dv/dy = 0
    n = 64
    v = fields.double_gyre_3d(n, 0.0).astype(np.float64)
    h = 1.0 / n
    dudx = (v[:, 1:-1, 2:, 0] - v[:, 1:-1, :-2, 0]) / (2 * h)
    dvdy = (v[:, 2:, 1:-1, 1] - v[:, :-2, 1:-1, 1]) / (2 * h)
    assert np.max(np.abs(0.5 * dudx + dvdy)) < 0.01 * np.max(np.abs(dvdy))


def test_orbit_camera_quaternion():
    cam = fields.orbit_camera(32, 16, (0.5, 0.5, 0.5), 0.0, 0.0)
    assert list(cam.quat) == [0.0, 0.0, 0.0, 1.0]
    cam = fields.orbit_camera(32, 16, (0.5, 0.5, 0.5), 1.2, -0.4, dist=3.0)
    q = np.array(cam.quat, np.float64)
    assert abs(np.linalg.norm(q) - 1.0) < 1e-6
    # yaw then pitch: q = qy * qx, so q rotates e_x like R_y(1.2) R_x(-0.4) e_x = R_y(1.2) e_x
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    np.testing.assert_allclose(R @ [1, 0, 0], [np.cos(1.2), 0, -np.sin(1.2)], atol=1e-6)
    assert cam.dist == 3.0 and cam.width == 32 and cam.height == 16
