"""Unit pins of the oracle's local update (a1-a10): closed forms, scipy conventions and
central finite differences, each chosen so that a plausible slip (a transposed rotation,
a dropped term, a wrong sign) fails it."""
import os

import numpy as np
from scipy.spatial.transform import Rotation

from oracle import cp_local
from conftest import GOLDEN

RNG = np.random.default_rng(0)


def _quats(n):
    q = RNG.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def test_quat_matches_scipy():
    q = _quats(16)
    R = cp_local.rot_mat_from_quat(q)
    Rs = Rotation.from_quat(np.c_[q[:, 1:], q[:, :1]]).as_matrix()  # scipy is scalar-last
    np.testing.assert_allclose(R, Rs, atol=1e-14)


def test_quat_no_normalisation():
    q = 2.0 * _quats(4)
    R = cp_local.rot_mat_from_quat(q)
    np.testing.assert_allclose(R, 4.0 * cp_local.rot_mat_from_quat(q / 2.0), atol=1e-13)


def test_euler_zyx_matches_scipy():
    ang = RNG.uniform(-180, 180, size=(16, 3))
    R = cp_local.euler_zyx_to_rot(ang)
    np.testing.assert_allclose(R, Rotation.from_euler("zyx", ang, degrees=True).as_matrix(), atol=1e-14)


def test_schmid_tables():
    for name in ("slip_fcc12.txt", "slip_bcc12.txt", "slip_bcc24.txt"):
        P0 = cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, name)))
        np.testing.assert_allclose(np.trace(P0, axis1=1, axis2=2), 0.0, atol=1e-15)
        np.testing.assert_allclose(np.linalg.norm(P0, axis=(1, 2)), 1.0, atol=1e-15)


def test_latent_matrix():
    q = cp_local.latent_matrix(12, 1.4)
    for i in range(12):
        for j in range(12):
            assert q[i, j] == (1.0 if i // 3 == j // 3 else 1.4)
    assert np.all(cp_local.latent_matrix(12, 1.4, self_block=False) == 1.4)


def test_rotate_rank4_properties():
    C = cp_local.cubic_stiffness(168.4e3, 121.4e3, 75.4e3)
    np.testing.assert_array_equal(cp_local.rotate_rank4(np.eye(3), C), C)
    R = cp_local.rot_mat_from_quat(_quats(1))[0]
    Cr = cp_local.rotate_rank4(R, C)
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        np.testing.assert_allclose(Cr, Cr.transpose(perm), atol=1e-9)
    Rz = cp_local.euler_zyx_to_rot(np.array([90.0, 0, 0]))
    C2 = cp_local.cubic_stiffness(1.0, 2.0, 3.0) + 0.5 * np.einsum("i,j,k,l->ijkl", *[np.eye(3)[0]] * 4)
    C2r = cp_local.rotate_rank4(Rz, C2)
    assert abs(C2r[1, 1, 1, 1] - C2[0, 0, 0, 0]) < 1e-14 and abs(C2r[0, 0, 0, 0] - C2[1, 1, 1, 1]) < 1e-14
    # isotropic stiffness is rotation invariant
    lam, mu = 100.0, 40.0
    Ci = cp_local.cubic_stiffness(lam + 2 * mu, lam, mu)
    np.testing.assert_allclose(cp_local.rotate_rank4(R, Ci), Ci, atol=1e-11)
    # the rotated stiffness acts as R (C : (R^T E R)) R^T
    E = RNG.normal(size=(3, 3))
    E = E + E.T
    lhs = np.einsum("ijkl,kl->ij", Cr, E)
    rhs = R @ np.einsum("ijkl,kl->ij", C, R.T @ E @ R) @ R.T
    np.testing.assert_allclose(lhs, rhs, atol=1e-8)


def _material(table="slip_fcc12.txt", **kw):
    P0 = cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, table)))
    base = dict(xm=0.02, gss_a=2.0, h=500.0, t_sat=300.0, r=1.4, C11=204.6e3, C12=137.7e3, C44=126.2e3)
    base.update(kw)
    return cp_local.Material(P0, **base)


def _batch(Q, scale=1e-3, g=88.0, N=12):
    """Random states whose elastic trial stress is O(g): |u_grad| ~ scale, Fp ~ I."""
    H = scale * RNG.normal(size=(Q, 3, 3))
    Fp = np.eye(3) + 0.2 * scale * RNG.normal(size=(Q, 3, 3))
    gg = g * (1 + 0.1 * RNG.random((Q, N)))
    slip = 1e-3 * RNG.normal(size=(Q, N))
    R = cp_local.rot_mat_from_quat(_quats(Q))
    return H, Fp, gg, slip, R


def test_elastic_limit_is_stvenant_kirchhoff():
    mat = _material()
    H, Fp, g, slip, R = _batch(8)
    Fp[:] = np.eye(3)
    g[:] = 1e12  # no slip at all
    out = cp_local.local_update(H, Fp, g, slip, R, mat, 1.0)
    F = np.eye(3) + H
    E = 0.5 * (np.swapaxes(F, 1, 2) @ F - np.eye(3))
    Cr = cp_local.rotate_rank4(R, np.broadcast_to(mat.C, (8, 3, 3, 3, 3)))
    S = np.einsum("qijkl,qkl->qij", Cr, E)
    np.testing.assert_allclose(out["S"], S, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(out["P"], F @ S, rtol=1e-11, atol=1e-8)


def test_uniaxial_elastic_closed_form():
    """sigma_zz after a 1.25e-3 single-element tension step in the elastic limit equals
    lambda_z S_zz / lambda_x^2 with S_zz = E001 * E_zz (E001 = (C11-C12)(C11+2C12)/(C11+C12))."""
    C11, C12 = 1.684e5, 1.214e5
    E001 = (C11 - C12) * (C11 + 2 * C12) / (C11 + C12)
    nu = C12 / (C11 + C12)
    ez = 1.25e-3
    # homogeneous state with free lateral stress: solve lateral stretch so S_xx = 0
    from scipy.optimize import brentq
    C = cp_local.cubic_stiffness(C11, C12, 75.4e3)

    def sxx(lx):
        F = np.diag([lx, lx, 1 + ez])
        E = 0.5 * (F.T @ F - np.eye(3))
        return np.einsum("ijkl,kl->ij", C, E)

    lx = brentq(lambda l: sxx(l)[0, 0], 0.99, 1.0)
    S = sxx(lx)
    Ezz = 0.5 * ((1 + ez) ** 2 - 1)
    Exx = 0.5 * (lx * lx - 1)
    assert abs(Exx / Ezz + nu) < 1e-12
    assert abs(S[2, 2] - E001 * Ezz) / S[2, 2] < 1e-12
    mat = cp_local.Material(cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, "slip_fcc12.txt"))),
                            xm=0.1, gss_a=2.5, h=541.5, t_sat=109.8, r=1.0, C11=C11, C12=C12, C44=75.4e3)
    H = (np.diag([lx, lx, 1 + ez]) - np.eye(3))[None]
    out = cp_local.local_update(H, np.eye(3)[None], np.full((1, 12), 1e12), np.zeros((1, 12)),
                                np.eye(3)[None], mat, 0.125)
    sig = out["sigma"][0]
    assert abs(sig[2, 2] - (1 + ez) * S[2, 2] / lx ** 2) / sig[2, 2] < 1e-12
    assert abs(sig[2, 2] - 83.604) < 1e-2  # the elastic estimate behind the Cu curve's first point


def _ctx(H, Fp, g, slip, R, mat, dt):
    return cp_local._Ctx(H, Fp, g, slip, R, mat, dt)


def test_residual_jacobian_fd():
    mat = _material()
    H, Fp, g, slip, R = _batch(6, scale=1e-3)
    ctx = _ctx(H, Fp, g, slip, R, mat, 0.5)
    S = 60.0 * RNG.normal(size=(6, 3, 3))
    S = S + np.swapaxes(S, 1, 2)
    _, aux = cp_local.residual(ctx, S)
    J = cp_local.jac_S(ctx, S, aux)
    eps = 1e-4
    for m in range(9):
        dS = np.zeros((6, 3, 3))
        dS[:, m // 3, m % 3] = eps
        rp, _ = cp_local.residual(ctx, S + dS)
        rm, _ = cp_local.residual(ctx, S - dS)
        np.testing.assert_allclose((rp - rm) / (2 * eps), J[:, :, m], rtol=2e-6, atol=2e-6)
    JH = cp_local.jac_H(ctx, S, aux)
    for m in range(9):
        Hp, Hm = H.copy(), H.copy()
        Hp[:, m // 3, m % 3] += 1e-7
        Hm[:, m // 3, m % 3] -= 1e-7
        rp, _ = cp_local.residual(_ctx(Hp, Fp, g, slip, R, mat, 0.5), S)
        rm, _ = cp_local.residual(_ctx(Hm, Fp, g, slip, R, mat, 0.5), S)
        np.testing.assert_allclose((rp - rm) / 2e-7, JH[:, :, m], rtol=1e-6, atol=1e-3)


def test_newton_converges_symmetric():
    mat = _material()
    H, Fp, g, slip, R = _batch(32, scale=1.2e-3)
    out = cp_local.local_update(H, Fp, g, slip, R, mat, 0.5)
    assert not out["capped"].any()
    assert np.all(np.linalg.norm(out["residual"], axis=1) <= 1e-8)
    np.testing.assert_allclose(out["S"], np.swapaxes(out["S"], 1, 2), atol=1e-9)
    assert out["iters"].max() >= 3  # genuinely plastic points


def test_tangent_fd():
    mat = _material(xm=0.05)
    H, Fp, g, slip, R = _batch(6, scale=1e-3)
    out = cp_local.local_update(H, Fp, g, slip, R, mat, 0.5, tol=1e-12)
    A = out["A"]
    eps = 1e-7
    for m in range(9):
        Hp, Hm = H.copy(), H.copy()
        Hp[:, m // 3, m % 3] += eps
        Hm[:, m // 3, m % 3] -= eps
        Pp = cp_local.pk1_only(Hp, Fp, g, slip, R, mat, 0.5, tol=1e-12)
        Pm = cp_local.pk1_only(Hm, Fp, g, slip, R, mat, 0.5, tol=1e-12)
        fd = ((Pp - Pm) / (2 * eps)).reshape(6, 9)
        np.testing.assert_allclose(fd, A[:, :, m], rtol=2e-5, atol=2e-2)


def test_hardening_sign_and_block():
    mat = _material(r=1.4)
    H, Fp, g, slip, R = _batch(4, scale=1.2e-3)
    g[:] = 88.0
    out = cp_local.local_update(H, Fp, g, slip, R, mat, 1.0, want_tangent=False)
    dgam = out["slip_new"] - slip
    tmp = 500.0 * np.abs(dgam) * (1 - 88.0 / 300.0) ** 2.0
    q = cp_local.latent_matrix(12, 1.4)
    np.testing.assert_allclose(out["g_new"] - g, tmp @ q.T, rtol=1e-12, atol=1e-14)
