"""Direct-solve option ('umfpack_solver', solver.py:50-60, :100-101) on the device:
cpfem_solve_direct against scipy's sparse LU (SuperLU, the factorisation spsolve runs when
UMFPACK is absent) on the exported CSR of the same Newton matrix, with the mesh's own node
order and with shuffled node numbers (reverse Cuthill-McKee order); then whole load steps and
an adjoint gradient with 'umfpack_solver' against the same runs with 'jax_solver'."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(n=6, shuffle=False):
    import cpfem_amd as C
    from cpfem_amd import configs
    w = configs.config2(n=n)
    mesh = w.mesh
    if shuffle:
        perm = np.random.default_rng(0).permutation(len(mesh.points))  # old node k -> new perm[k]
        pts = np.empty_like(mesh.points)
        pts[perm] = mesh.points
        mesh = C.Mesh(pts, perm[np.asarray(mesh.cells)])
        w = w.__class__(**{**w.__dict__, "mesh": mesh})
    prob = configs.build_problem(w)
    return w, prob


def _newton_matrix(w, prob):
    from cpfem_amd.solver import bc_x0
    info = w.dirichlet_bc_info(w.disps[1])
    prob.fes[0].update_Dirichlet_boundary_conditions(info)
    prob.dt = w.ts[1] - w.ts[0]
    prob.set_params(prob.internal_vars)
    dofs = torch.zeros(prob.fes[0].num_total_nodes * 3, dtype=torch.float64, device="cuda")
    dofs = bc_x0(prob, dofs)  # Dirichlet values in place: a plastic state at the first step
    res, mat = prob.assemble(dofs)
    return res, mat


@pytest.mark.parametrize("shuffle", [False, True])
def test_direct_solve_matches_sparse_lu(shuffle):
    import scipy.sparse.linalg as sla
    from cpfem_amd import BlockEllMatrix
    from cpfem_amd.solver import direct_solve
    w, prob = _problem(6, shuffle)
    res, mat = _newton_matrix(w, prob)
    b = -res
    A = BlockEllMatrix(prob.ctx, mat).to_scipy().tocsc()
    bh = b.cpu().numpy()
    lu = sla.splu(A)
    xr = lu.solve(bh)
    xr += lu.solve(bh - A @ xr)  # one refinement step of the reference solve as well
    for refine, tol in ((0, 1e-9), (2, 1e-11)):
        x = direct_solve(prob, mat, b, {"umfpack_solver": {"refine": refine}}).cpu().numpy()
        err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        rel_res = prob.solver_info[-1]["residual"] / prob.solver_info[-1]["b_norm"]
        print(f"shuffle={shuffle} refine={refine}: |x - x_lu|/|x_lu| = {err:.2e}, |Ax - b|/|b| = {rel_res:.2e}")
        assert err < tol, (shuffle, refine, err)
        assert rel_res < 1e-10, (shuffle, refine, rel_res)


def test_load_steps_with_direct_solver_match_krylov():
    from cpfem_amd import configs
    outs = []
    for opt in ({"jax_solver": {}}, {"umfpack_solver": {}}):
        w, prob = _problem(4)
        sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
        _, _, sig = configs.run_steps(prob, w, prob.internal_vars, sol, 0, 4, dict(opt))
        outs.append(np.array(sig))
    assert any(i.get("direct") for i in prob.solver_info)
    np.testing.assert_allclose(outs[1], outs[0], rtol=0, atol=1e-7 * np.abs(outs[0]).max())


def test_adjoint_with_direct_solver_matches_krylov():
    """implicit_vjp with adjoint_solver_options {'umfpack_solver': {}} (as the steel
    calibration drivers pass, calibration_case4_UQ_polyCrystalSteel_1D_GB.py:90)."""
    import cpfem_amd as C
    m = C.box_mesh(2, 2, 2, 1.0, 1.0, 1.0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    # symmetry BCs (a single pinned corner would leave the rotation about z free: a singular
    # matrix that Krylov tolerates and an unpivoted LU does not)
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    front = lambda p: np.isclose(p[1], 0.0, atol=1e-5)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    top = lambda p: np.isclose(p[2], 1.0, atol=1e-5)
    zero = lambda p: 0.0
    grads = []
    for opt in ({"jax_solver": {}}, {"umfpack_solver": {}}):
        info = [[left, front, bottom, top], [0, 1, 2, 2], [zero, zero, zero, lambda p: 2e-3]]
        prob = C.CrystalPlasticity(mesh, dirichlet_bc_info=info,
                                   additional_info=(np.array([[1.0, 0, 0, 0]]), np.zeros(8)),
                                   material=C.CALIB_COPPER.replace(tol=1e-10))
        prob.dt = 0.2
        fwd = C.ad_wrapper(prob, {**opt, "tol": 1e-9, "rel_tol": 1e-13}, dict(opt))
        alpha = torch.full((2,), 1.3, dtype=torch.float64, device="cuda", requires_grad=True)
        params = prob.internal_vars
        params[1] = alpha[0] * params[1]
        params[-5] = alpha[1] * params[-5]
        sol = fwd(params)[0]
        s = prob.compute_avg_stress(sol, params)[:, 2, 2].mean()
        s.backward()
        grads.append(alpha.grad.cpu().numpy())
    np.testing.assert_allclose(grads[1], grads[0], rtol=1e-7)
