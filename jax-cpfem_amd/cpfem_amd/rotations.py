"""Orientation helpers: quaternion and Euler-angle rotations, uniform SO(3) sampling.

``rotate_theta_transfer`` and ``alpha_rotate_tensor_rank_2`` accept torch tensors and then
stay differentiable under autograd (the orientation-design path differentiates the loss
with respect to the Euler angles through them)."""
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


def rotate_theta_transfer(angles_deg):
    """(..., 3) extrinsic z-y-x Euler angles in degrees -> (..., 3, 3) rotation matrices
    (poly_sensitivity_rot_mesh2.py:41-53, scipy ``from_euler('zyx', degrees=True)``).
    Torch input gives a differentiable torch result; anything else goes through NumPy."""
    try:
        import torch
    except ImportError:  # pragma: no cover
        torch = None
    if torch is None or not isinstance(angles_deg, torch.Tensor):
        return euler_zyx(angles_deg)
    a = torch.deg2rad(angles_deg)
    ca, sa = torch.cos(a[..., 0]), torch.sin(a[..., 0])
    cb, sb = torch.cos(a[..., 1]), torch.sin(a[..., 1])
    cc, sc = torch.cos(a[..., 2]), torch.sin(a[..., 2])
    z, o = torch.zeros_like(ca), torch.ones_like(ca)
    Rz = torch.stack([torch.stack([ca, -sa, z], -1), torch.stack([sa, ca, z], -1), torch.stack([z, z, o], -1)], -2)
    Ry = torch.stack([torch.stack([cb, z, sb], -1), torch.stack([z, o, z], -1), torch.stack([-sb, z, cb], -1)], -2)
    Rx = torch.stack([torch.stack([o, z, z], -1), torch.stack([z, cc, -sc], -1), torch.stack([z, sc, cc], -1)], -2)
    return Rx @ Ry @ Rz


def alpha_rotate_tensor_rank_2(R_a, R_b):
    """R_a @ R_b per point: a newly applied rotation composed with the existing
    orientation (poly_sensitivity_rot_mesh2.py:60-65, applied to params[3] at :203)."""
    return R_a @ R_b
