"""The C++/OpenMP CPU restatement that the benchmark times (oracle/cpu) computes the reference
algorithm: its assembled residual and CSR Jacobian (with the Dirichlet row elimination), its
point pass (Cauchy stress, history update) and its Jacobi-BiCGSTAB agree with the NumPy
oracle on a small polycrystal, and its per-point VJP agrees with central differences of the
oracle's P."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import slip_table

pytest.importorskip("scipy")


def _setup(n=3, seed=0, strain=1e-3):
    from oracle import cp_local, fe
    from oracle.cpu import build, cpu_ref
    if not os.path.exists(build.LIBS["base"]):
        build.build()
    rng = np.random.default_rng(seed)
    pts, cells = fe.box_mesh(n, n, n, 1.0, 1.0, 1.0)
    nc, Q = len(cells), 8 * len(cells)
    table = slip_table("slip_fcc12.txt")
    d = table[:, 3:6] / np.linalg.norm(table[:, 3:6], axis=1)[:, None]
    nrm = table[:, 0:3] / np.linalg.norm(table[:, 0:3], axis=1)[:, None]
    q = rng.normal(size=(4, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    Rc = cp_local.rot_mat_from_quat(q)[rng.integers(0, 4, size=nc)]
    mp = dict(xm=0.05, gss_a=2.0, h=500.0, t_sat=300.0, r=1.4, C11=204.6e3, C12=137.7e3, C44=126.2e3)
    om = cp_local.Material(cp_local.schmid_tensors(table), **mp)
    g0 = 88.0 * (1 + 0.1 * rng.random(nc))
    prob = fe.FEProblem(pts, cells, om, Rc, g0, dt=1.0)
    F = fe.faces(pts, (1.0, 1.0, 1.0))
    prob.set_dirichlet([F["left"], F["front"], F["bottom"], F["top"]], [0, 1, 2, 2], [0.0, 0.0, 0.0, strain])
    u = strain * pts[:, 2:3] * np.array([[-0.3, -0.3, 1.0]]) + 1e-4 * strain * rng.normal(size=pts.shape)
    u = u.ravel()
    soa = lambda a, comps: np.ascontiguousarray(a.reshape((Q,) + comps).reshape(Q, -1).T)
    mat8 = np.array([[mp[k]] * Q for k in ("xm", "gss_a", "h", "t_sat", "r", "C11", "C12", "C44")])
    model = cpu_ref.Model(pts, cells, d, nrm, soa(prob.Fp_inv, (3, 3)), soa(prob.g, (12,)), soa(prob.R, (3, 3)), mat8,
                          1e-3, 1e-8, 5, 250, 1, prob.bc_dofs, prob.bc_vals)
    return prob, model, u


def test_assembly_matches_oracle():
    prob, model, u = _setup()
    res_ref = prob.apply_bc_vec(prob.newton_update(u), u)
    A_ref = prob.get_A()
    res, vals, counts = model.assemble(u, 1.0)
    A = model.csr(vals)
    assert counts[0] > 0 and counts[1] > counts[0]  # plastic points iterate; evaluations include trials
    assert np.abs(res - res_ref).max() <= 1e-10 * np.abs(res_ref).max()
    D = (A - A_ref).tocsr()
    assert abs(D).max() <= 1e-10 * abs(A_ref).max()
    # same sparsity (the oracle's COO sum keeps exact zeros of BC rows as stored entries)
    assert A.nnz == A_ref.nnz


def test_point_pass_matches_oracle():
    from oracle import fe
    prob, model, u = _setup(n=2)
    sig, Fn, gn = model.point_pass(u, 1.0)
    out, _ = prob.local(u, want_tangent=False)
    Q = 8 * prob.nc
    np.testing.assert_allclose(sig.T.reshape(Q, 3, 3), out["sigma"], rtol=0, atol=1e-10 * np.abs(out["sigma"]).max())
    np.testing.assert_allclose(Fn.T.reshape(Q, 3, 3), out["Fp_inv_new"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(gn.T, out["g_new"], rtol=1e-12)
    del fe


def test_bicgstab_matches_oracle_and_spsolve():
    import scipy.sparse.linalg as spla
    from oracle import fe
    prob, model, u = _setup()
    res, vals, _ = model.assemble(u, 1.0)
    A = model.csr(vals)
    b = -res
    x, k = model.bicgstab(vals, b, np.zeros_like(b), tol=1e-12, atol=1e-20)
    xr, kr = fe.bicgstab_jax(A, b, np.zeros_like(b), M=lambda v: v / A.diagonal(), tol=1e-12, atol=1e-20)
    xs = spla.spsolve(A.tocsc(), b)
    assert k > 10
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xs)
    assert abs(k - kr) <= max(3, kr // 5)


def test_vjp_pass_matches_oracle_fd():
    """Hbar = (dP/dH)^T Pbar per point against central differences of the oracle's P."""
    from oracle import cp_local, fe
    prob, model, u = _setup(n=1)
    Q = 8 * prob.nc
    Pbar = np.random.default_rng(3).normal(size=(9, Q))
    Hb, pb = model.vjp_pass(u, 1.0, Pbar)
    ug = fe.u_grads(u.reshape(-1, 3), prob.cells, prob.sg).reshape(Q, 3, 3)
    dH = np.random.default_rng(4).normal(size=(Q, 3, 3))
    eps = 1e-7
    f = lambda H: cp_local.local_update(H, prob._flat(prob.Fp_inv), prob._flat(prob.g), prob._flat(prob.slip),
                                        prob._flat(prob.R), prob.mat, 1.0, tol=1e-13, want_tangent=False)["P"]
    dP = (f(ug + eps * dH) - f(ug - eps * dH)) / (2 * eps)
    lhs = np.einsum("qij,qij->q", Hb.T.reshape(Q, 3, 3), dH)
    rhs = np.einsum("qij,qij->q", Pbar.T.reshape(Q, 3, 3), dP)
    assert np.abs(lhs - rhs).max() <= 1e-6 * np.abs(rhs).max()
    assert np.isfinite(pb).all()
