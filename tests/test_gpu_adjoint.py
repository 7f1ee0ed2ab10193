"""Hand-written adjoint vs central finite differences of the same forward.

(1) ``simulate`` (whole history, checkpointed) on a 2x2x2 polycrystal: gradient of a random
    linear functional of the averaged stresses w.r.t. scales of (g0, h0, g_sat, a, m, r, C11,
    C12, C44) and w.r.t. per-grain z-y-x Euler angles.
(2) The drop-in AD path of the reference's calibration driver (calibration_case1_
    singleCrystalCopper_GB.py:176-246: ad_wrapper + compute_avg_stress + update_int_vars_gp
    under autograd) for the Cu single-element objective at alpha = 1.4: gradient w.r.t. the six
    scale coefficients; and the L-BFGS-B loop of that driver (cpfem_amd.optimize) against the
    points the reference's calibration run logged.
Central differences use relative step 1e-6 (1e-4 degree for angles) with the Newton
tolerances tightened so the forward is smooth to ~1e-10 (the local tolerance 1e-10 |r| sits just
above the FP64 rounding floor of the residual at these stress levels; at 1e-11 some points
stall there and hit the local iteration cap, which is an error).  The Krylov tolerance is 1e-12
relative (atol negligible): near 1e-13 Jacobi-BiCGSTAB on these tiny systems reaches the
rounding floor of its recurrence and can break down (rhat.q = 0) in the reference semantics
itself, whichever way the rounding falls."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TIGHT = {"tol": 1e-9, "rel_tol": 1e-13, "krylov_tol": 1e-12, "krylov_atol": 1e-20}
ADJ = {"krylov_tol": 1e-12, "krylov_atol": 1e-20}


def _poly(n_steps=3, k=2):
    import cpfem_amd as C
    from cpfem_amd.rotations import uniform_quaternions
    rng = np.random.default_rng(5)
    m = C.box_mesh(2, 2, 2, 1.0, 1.0, 1.0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    quat = uniform_quaternions(2, rng)
    inds = np.array([0, 1, 1, 0, 1, 0, 0, 1])
    mat = C.Material(slip="fcc12", g0=88.0, h=500.0, t_sat=300.0, gss_a=2.0, xm=0.05, r=1.4,
                     C11=204.6e3, C12=137.7e3, C44=126.2e3, tol=1e-10)
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    front = lambda p: np.isclose(p[1], 0.0, atol=1e-5)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    top = lambda p: np.isclose(p[2], 1.0, atol=1e-5)
    zero = lambda p: 0.0
    bc = lambda d: [[left, front, bottom, top], [0, 1, 2, 2], [zero, zero, zero, (lambda v: (lambda p: v))(d)]]
    prob = C.CrystalPlasticity(mesh, dirichlet_bc_info=bc(0.0), additional_info=(quat, inds), material=mat)
    load = C.LoadHistory(np.linspace(0, 3e-3, n_steps + 1), np.linspace(0, 3.0, n_steps + 1), bc)
    return prob, load, quat, inds, mat


def _inputs(prob, mat, theta, euler, quat, inds):
    """Per-point (g0, mat8, R) from 9 scales and per-grain Euler angles (differentiable)."""
    from cpfem_amd.rotations import alpha_rotate_tensor_rank_2, get_rot_mat, rotate_theta_transfer
    nc = prob.fes[0].num_cells
    base = torch.tensor([mat.xm, mat.gss_a, mat.h, mat.t_sat, mat.r, mat.C11, mat.C12, mat.C44],
                        dtype=torch.float64, device="cuda")
    # theta order: g0, h, t_sat, a, m, r, C11, C12, C44
    scale = torch.stack([theta[4], theta[3], theta[1], theta[2], theta[5], theta[6], theta[7], theta[8]])
    mat8 = (base * scale).expand(nc, 8, 8)
    g0 = (mat.g0 * theta[0]).expand(nc, 8, 12)
    # poly_sensitivity_rot_mesh2.py:41-65,203: R = R_euler(theta) @ R_quat per grain
    Re = rotate_theta_transfer(euler)
    Rq = torch.as_tensor(get_rot_mat(quat), device="cuda")
    R = alpha_rotate_tensor_rank_2(Re, Rq)[torch.as_tensor(inds, device="cuda")]
    return g0, mat8, R.unsqueeze(1).expand(nc, 8, 3, 3)


def test_simulate_gradient_matches_fd():
    import cpfem_amd as C
    prob, load, quat, inds, mat = _poly()
    w = torch.as_tensor(np.random.default_rng(1).normal(size=(load.n_steps, 3, 3)), device="cuda")

    def loss(theta, euler):
        g0, m8, R = _inputs(prob, mat, theta, euler, quat, inds)
        sig = C.simulate(prob, load, g0, m8, R, checkpoint_every=2, solver_options=TIGHT,
                         adjoint_solver_options=ADJ)
        return (sig * w).sum()

    theta = torch.ones(9, dtype=torch.float64, device="cuda", requires_grad=True)
    euler = torch.tensor([[10.0, 20.0, 30.0], [-40.0, 15.0, 70.0]], dtype=torch.float64, device="cuda",
                         requires_grad=True)
    L = loss(theta, euler)
    L.backward()
    gt, ge = theta.grad.cpu().numpy(), euler.grad.cpu().numpy()
    with torch.no_grad():
        fd = np.zeros(9)
        for k in range(9):
            h = 1e-6
            tp, tm = theta.detach().clone(), theta.detach().clone()
            tp[k] += h
            tm[k] -= h
            fd[k] = (loss(tp, euler.detach()) - loss(tm, euler.detach())).item() / (2 * h)
        fde = np.zeros((2, 3))
        for g in range(2):
            for j in range(3):
                h = 1e-4
                ep, em = euler.detach().clone(), euler.detach().clone()
                ep[g, j] += h
                em[g, j] -= h
                fde[g, j] = (loss(theta.detach(), ep) - loss(theta.detach(), em)).item() / (2 * h)
    scale = np.abs(fd).max()
    assert np.abs(gt - fd).max() / scale < 1e-5, (gt, fd)
    assert np.abs(ge - fde).max() / np.abs(fde).max() < 1e-5, (ge, fde)


def _cu_calibration():
    """The objective of calibration_case1_singleCrystalCopper_GB.py:176-246 through the
    drop-in AD path, as a function of the six scale coefficients."""
    import cpfem_amd as C
    ref = torch.as_tensor(np.loadtxt(os.path.join(GOLDEN, "cu_tension_szz.txt"))[:20], device="cuda")
    m = C.box_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    corner = lambda p: np.isclose(p[0], 0.0, atol=1e-5) & np.isclose(p[1], 0.0, atol=1e-5) & np.isclose(p[2], 0.0, atol=1e-5)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    top = lambda p: np.isclose(p[2], 1.0, atol=1e-5)
    zero = lambda p: 0.0
    disps, ts = np.linspace(0, 0.025, 21), np.linspace(0, 2.5, 21)
    info = [[corner, corner, bottom, top], [0, 1, 2, 2], [zero, zero, zero, lambda p: 0.0]]
    prob = C.CrystalPlasticity(mesh, dirichlet_bc_info=info, additional_info=(np.array([[1.0, 0, 0, 0]]), np.zeros(1)),
                               material=C.CALIB_COPPER.replace(tol=1e-10))
    fwd_pred = C.ad_wrapper(prob, {"jax_solver": {}, **TIGHT}, ADJ)
    idx = np.arange(0, 20, 4)

    def simulation(alpha):  # calibration_case1_singleCrystalCopper_GB.py:176-246
        prob.custom_init(np.array([[1.0, 0, 0, 0]]), np.zeros(1))
        prob.set_initial_guess([torch.zeros(8, 3, dtype=torch.float64, device="cuda")])
        params = prob.internal_vars
        params[1] = alpha[0] * params[1]
        params[-5] = alpha[1] * params[-5]
        params[-4] = alpha[2] * params[-4]
        params[-3] = alpha[3] * params[-3]
        params[-2] = alpha[4] * params[-2]
        params[-1] = alpha[5] * params[-1]
        out = []
        for i in range(20):
            prob.dt = ts[i + 1] - ts[i]
            info[-1][-1] = (lambda d: (lambda p: d))(disps[i + 1])
            prob.fes[0].update_Dirichlet_boundary_conditions(info)
            sol = fwd_pred(params)[0]
            out.append(prob.compute_avg_stress(sol, params)[0, 2, 2])
            params = prob.update_int_vars_gp(sol, params)
        s = torch.stack(out)
        return 1e6 * torch.sum((s[idx] - ref[idx]) ** 2) / torch.sum(ref[idx] ** 2)

    return simulation


def test_dropin_calibration_gradient_matches_fd():
    simulation = _cu_calibration()
    alpha = torch.full((6,), 1.4, dtype=torch.float64, device="cuda", requires_grad=True)
    obj = simulation(alpha)
    expect = float(np.loadtxt(os.path.join(GOLDEN, "cu_objective.txt")))
    assert abs(obj.item() - expect) / expect < 1e-8
    obj.backward()
    g = alpha.grad.cpu().numpy()
    fd = np.zeros(6)
    with torch.no_grad():
        for k in range(6):
            h = 1.4e-6
            ap, am = alpha.detach().clone(), alpha.detach().clone()
            ap[k] += h
            am[k] -= h
            fd[k] = (simulation(ap) - simulation(am)).item() / (2 * h)
    assert np.abs(g - fd).max() / np.abs(fd).max() < 1e-5, (g, fd)


def test_orientation_design_gradient_bcc24_dual_phase():
    """Orientation-design scenario in small (poly_sensitivity_rot_mesh2.py:41-65,203-239 with the
    dual-phase BCC24 model of models_DPsteel_inhomo.py): loss = (final mean szz - target)^2,
    gradient w.r.t. per-grain z-y-x Euler angles through the whole history (checkpoint every 2
    steps), FP64 and mixed precision, against central differences of the FP64 forward."""
    import cpfem_amd as C
    from cpfem_amd.rotations import alpha_rotate_tensor_rank_2, get_rot_mat, rotate_theta_transfer
    rng = np.random.default_rng(7)
    m = C.box_mesh(3, 3, 3, 1.0, 1.0, 1.0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    nc, G = len(mesh.cells), 4
    inds = rng.integers(0, G, size=nc)
    phase = (inds == 3).astype(np.int64)  # one hard grain of four
    soft = C.Material(slip="bcc24", g0=100.0, h=400.0, t_sat=250.0, gss_a=2.0, xm=0.05, r=1.4,
                      C11=231.0e3, C12=135.0e3, C44=116.0e3, tol=1e-10)
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    front = lambda p: np.isclose(p[1], 0.0, atol=1e-5)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    top = lambda p: np.isclose(p[2], 1.0, atol=1e-5)
    zero = lambda p: 0.0
    bc = lambda d: [[left, front, bottom, top], [0, 1, 2, 2], [zero, zero, zero, (lambda v: (lambda p: v))(d)]]
    quat = C.rotations.uniform_quaternions(G, rng)
    prob = C.CrystalPlasticity(mesh, dirichlet_bc_info=bc(0.0), additional_info=(quat, inds), material=soft)
    load = C.LoadHistory(np.linspace(0, 3e-3, 4), np.linspace(0, 3.0, 4), bc)
    ph = torch.as_tensor(phase, device="cuda")
    g0 = torch.where(ph[:, None, None] == 1, 4 * soft.g0, soft.g0).to(torch.float64).expand(nc, 8, 24).contiguous()
    base = torch.tensor([soft.xm, soft.gss_a, soft.h, soft.t_sat, soft.r, soft.C11, soft.C12, soft.C44],
                        dtype=torch.float64, device="cuda")
    mat8 = base.expand(nc, 8, 8).clone()
    mat8[ph == 1, :, 2] *= 2.0  # h0 x2 in the hard phase
    Rq = torch.as_tensor(get_rot_mat(quat), device="cuda")
    idx = torch.as_tensor(inds, device="cuda")

    def final_szz(euler, precision="fp64"):
        R = alpha_rotate_tensor_rank_2(rotate_theta_transfer(euler), Rq)[idx].unsqueeze(1).expand(nc, 8, 3, 3)
        o = dict(TIGHT, precision=precision)
        return C.simulate(prob, load, g0, mat8, R, checkpoint_every=2, solver_options=o,
                          adjoint_solver_options=dict(ADJ, precision=precision))[-1, 2, 2]

    euler0 = torch.as_tensor(rng.uniform(-30, 30, size=(G, 3)), device="cuda")
    with torch.no_grad():
        target = final_szz(euler0) * 1.01
    grads = {}
    for prec in ("fp64", "mixed"):
        e = euler0.clone().requires_grad_(True)
        ((final_szz(e, prec) - target) ** 2).backward()
        grads[prec] = e.grad.cpu().numpy()
    fd = np.zeros((G, 3))
    with torch.no_grad():
        for g in range(G):
            for j in range(3):
                h = 1e-4
                ep, em = euler0.clone(), euler0.clone()
                ep[g, j] += h
                em[g, j] -= h
                fd[g, j] = (((final_szz(ep) - target) ** 2) - ((final_szz(em) - target) ** 2)).item() / (2 * h)
    scale = np.abs(fd).max()
    assert np.abs(grads["fp64"] - fd).max() / scale < 1e-5, (grads["fp64"], fd)
    assert np.abs(grads["mixed"] - grads["fp64"]).max() / scale < 1e-6, (grads["mixed"], grads["fp64"])


def test_calibration_loop_replays_reference_first_step(tmp_path):
    """cpfem_amd.optimize.minimize_lbfgsb around the same objective, from the driver's start
    point 1.4 with bounds [0.4, 3.0] (calibration_case1_singleCrystalCopper_GB.py:388-409),
    logs the points the reference's run logged (tests/golden/cu_calib_alpha.txt): the first,
    and the second that L-BFGS-B picks from the first gradient (a box corner set by the
    gradient's signs).  The reference stored no objective for the second point, so the run
    here stops when that point is handed to the objective."""
    from cpfem_amd.optimize import minimize_lbfgsb
    simulation = _cu_calibration()
    calls = []

    class Enough(Exception):
        pass

    def sim(x):
        calls.append(x.detach().cpu().numpy().copy())
        if len(calls) == 2:
            raise Enough
        return simulation(x)

    with pytest.raises(Enough):
        minimize_lbfgsb(sim, [1.4] * 6, ([0.4] * 6, [3.0] * 6), maxiter=10, obj_stop=1.0, log_dir=str(tmp_path))
    ref_alpha = np.loadtxt(os.path.join(GOLDEN, "cu_calib_alpha.txt"))
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "alpha.txt"), ref_alpha)
    expect = float(np.loadtxt(os.path.join(GOLDEN, "cu_objective.txt")))
    assert abs(float(np.loadtxt(tmp_path / "obj_val.txt")) - expect) / expect < 1e-8
