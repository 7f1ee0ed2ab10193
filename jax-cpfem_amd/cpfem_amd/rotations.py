"""Orientation helpers: quaternion and Euler-angle rotations, uniform SO(3) sampling."""
import numpy as np


def get_rot_mat(q):
    """Scalar-first quaternion(s) -> rotation matrix, not normalised (models_copper.py:37-45)."""
    q = np.asarray(q, dtype=np.float64)
    q0, q1, q2, q3 = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([
        np.stack([q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * q1 * q2 - 2 * q0 * q3, 2 * q1 * q3 + 2 * q0 * q2], -1),
        np.stack([2 * q1 * q2 + 2 * q0 * q3, q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * q2 * q3 - 2 * q0 * q1], -1),
        np.stack([2 * q1 * q3 - 2 * q0 * q2, 2 * q2 * q3 + 2 * q0 * q1, q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3], -1),
    ], -2)
    return R


def euler_zyx(angles_deg):
    """Extrinsic z-y-x Euler angles (degrees) -> R = Rx(c) Ry(b) Rz(a); the scipy
    ``from_euler('zyx', ..., degrees=True)`` of poly_sensitivity_rot_mesh2.py:41-53."""
    a = np.deg2rad(np.asarray(angles_deg, dtype=np.float64))
    ca, sa = np.cos(a[..., 0]), np.sin(a[..., 0])
    cb, sb = np.cos(a[..., 1]), np.sin(a[..., 1])
    cc, sc = np.cos(a[..., 2]), np.sin(a[..., 2])
    z = np.zeros_like(ca)
    o = np.ones_like(ca)
    Rz = np.stack([np.stack([ca, -sa, z], -1), np.stack([sa, ca, z], -1), np.stack([z, z, o], -1)], -2)
    Ry = np.stack([np.stack([cb, z, sb], -1), np.stack([z, o, z], -1), np.stack([-sb, z, cb], -1)], -2)
    Rx = np.stack([np.stack([o, z, z], -1), np.stack([z, cc, -sc], -1), np.stack([z, sc, cc], -1)], -2)
    return Rx @ Ry @ Rz


def uniform_quaternions(n, rng):
    """Uniform random rotations (Shoemake 1992), scalar-first unit quaternions."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1)
