"""Orientation helpers of the host package against scipy (closed form) and autograd."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_euler_zyx_matches_scipy_and_is_differentiable():
    from scipy.spatial.transform import Rotation
    from cpfem_amd.rotations import alpha_rotate_tensor_rank_2, euler_zyx, get_rot_mat, rotate_theta_transfer
    rng = np.random.default_rng(0)
    ang = rng.uniform(-180, 180, size=(7, 3))
    ref = Rotation.from_euler("zyx", ang, degrees=True).as_matrix()
    np.testing.assert_allclose(euler_zyx(ang), ref, atol=1e-14)
    np.testing.assert_allclose(rotate_theta_transfer(ang), ref, atol=1e-14)
    t = torch.tensor(ang, dtype=torch.float64, requires_grad=True)
    np.testing.assert_allclose(rotate_theta_transfer(t).detach().numpy(), ref, atol=1e-14)
    assert torch.autograd.gradcheck(lambda x: rotate_theta_transfer(x), (t,), eps=1e-6, atol=1e-8)
    q = rng.normal(size=(7, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    Rq = get_rot_mat(q)
    np.testing.assert_allclose(Rq, Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix(), atol=1e-14)
    np.testing.assert_allclose(alpha_rotate_tensor_rank_2(ref, Rq), ref @ Rq, atol=0)
