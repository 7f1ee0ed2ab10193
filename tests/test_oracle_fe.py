"""Unit pins of the oracle's FE half (a13-a17): hex8 closed forms, element stiffness vs FD,
elastic symmetry / rigid-body modes, and BiCGSTAB vs a direct solve."""
import os

import numpy as np
import scipy.sparse.linalg as spla

from oracle import cp_local, fe
from conftest import GOLDEN

RNG = np.random.default_rng(1)


def _distorted_mesh():
    pts, cells = fe.box_mesh(2, 2, 2, 1.0, 0.8, 1.2)
    pts = pts + 0.03 * RNG.normal(size=pts.shape)
    return pts, cells


def test_hex8_volume_and_partition_of_unity():
    pts, cells = _distorted_mesh()
    sg, JxW = fe.shape_grads(pts, cells)
    assert abs(JxW.sum() - 1.0 * 0.8 * 1.2) < 0.05
    # exact: sum JxW = volume of the (trilinear) cells; compare with the undistorted box
    p0, c0 = fe.box_mesh(2, 2, 2, 1.0, 0.8, 1.2)
    assert abs(fe.shape_grads(p0, c0)[1].sum() - 0.96) < 1e-14
    np.testing.assert_allclose(sg.sum(axis=2), 0.0, atol=1e-12)
    N, _ = fe.hex8_shape(np.array([0.3, 0.6, 0.2]))
    assert abs(N.sum() - 1.0) < 1e-15


def test_patch_test():
    pts, cells = _distorted_mesh()
    sg, _ = fe.shape_grads(pts, cells)
    H = RNG.normal(size=(3, 3))
    u = pts @ H.T + RNG.normal(size=3)
    ug = fe.u_grads(u, cells, sg)
    np.testing.assert_allclose(ug, np.broadcast_to(H, ug.shape), atol=1e-12)


def _problem(pts, cells, plastic=True):
    P0 = cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, "slip_fcc12.txt")))
    mat = cp_local.Material(P0, xm=0.05, gss_a=2.0, h=500.0, t_sat=300.0, r=1.4,
                            C11=204.6e3, C12=137.7e3, C44=126.2e3)
    q = RNG.normal(size=(len(cells), 4))
    R = cp_local.rot_mat_from_quat(q / np.linalg.norm(q, axis=1, keepdims=True))
    return fe.FEProblem(pts, cells, mat, R, 40.0 if plastic else 1e12, dt=0.5)


def test_element_stiffness_fd():
    pts, cells = fe.box_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    prob = _problem(pts, cells)
    prob.tol = 1e-12
    u = 1e-3 * RNG.normal(size=prob.ndof)
    prob.newton_update(u)
    K = prob.V.reshape(24, 24)
    eps = 1e-7
    for j in range(24):
        up, um = u.copy(), u.copy()
        up[prob.cell_dofs[0, j]] += eps
        um[prob.cell_dofs[0, j]] -= eps
        fd = (prob.compute_residual(up) - prob.compute_residual(um)) / (2 * eps)
        np.testing.assert_allclose(fd[prob.cell_dofs[0]], K[:, j], rtol=1e-5, atol=1e-2)


def test_elastic_stiffness_symmetric_with_rigid_modes():
    pts, cells = fe.box_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    prob = _problem(pts, cells, plastic=False)
    prob.newton_update(np.zeros(prob.ndof))
    K = prob.V.reshape(24, 24)
    np.testing.assert_allclose(K, K.T, atol=1e-6 * np.abs(K).max())
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert np.sum(np.abs(ev) < 1e-8 * ev.max()) == 6


def test_bicgstab_matches_direct_solve():
    pts, cells = fe.box_mesh(3, 3, 3, 1.0, 1.0, 1.0)
    prob = _problem(pts, cells)
    F = fe.faces(pts, (1.0, 1.0, 1.0))
    prob.set_dirichlet([F["left"], F["front"], F["bottom"], F["top"]], [0, 1, 2, 2], [0, 0, 0, 1e-3])
    u = np.zeros(prob.ndof)
    res = prob.apply_bc_vec(prob.newton_update(u), u)
    A = prob.get_A()
    x0 = np.zeros(prob.ndof)
    x0[prob.bc_dofs] = prob.bc_vals
    x, k = fe.jax_solve(A, -res, x0)
    xd = spla.spsolve(A.tocsc(), -res)
    assert k > 0
    np.testing.assert_allclose(x, xd, rtol=1e-7, atol=1e-12 * np.abs(xd).max())
    # BC rows: identity
    np.testing.assert_allclose(x[prob.bc_dofs], prob.bc_vals, atol=1e-14)
