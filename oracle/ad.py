"""Oracle adjoint: the reference's AD gradient restated on the CPU with torch autograd (FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference obtains gradients from JAX reverse mode through three pieces, restated here in
the same structure (and independent of the hand-written adjoint of the device path, which
propagates dual numbers through a 6-unknown form and sweeps checkpointed histories):

* the local Newton as a ``custom_jvp`` (singlecrystal_copper/models_copper.py:204-259): the
  solve itself is not differentiated; its JVP rule is the implicit-function theorem
  dy = -J_y^-1 J_x dx with J_x, J_y from ``jacfwd`` of ``implicit_residual`` (:195-201).
  Here ``_LocalSolve``: forward = the oracle's local Newton (oracle/cp_local.py, pinned to
  the reference-held curves); backward = the transposed rule x_bar = -J_x^T J_y^-T y_bar with
  J_y from torch.func.jacrev of the residual below and J_x^T by autograd;
* the global solve as a ``custom_vjp`` (``ad_wrapper`` / ``implicit_vjp``,
  crystal_plasticity_OR_design/solver.py:801-874): forward = the oracle's Newton solver
  (oracle/fe.py); backward = A^T lambda = v with the row-eliminated Jacobian A at the
  converged solution (get_A, :279-304), then -(dc/dp)^T lambda, c the residual with
  apply_bc_vec (its Dirichlet rows, u - ubar, do not depend on p).  Here ``_GlobalSolve``
  with scipy's sparse LU for the transposed solve;
* everything else by ordinary AD: P (first_PK_stress :155-162), the per-cell Cauchy average
  (compute_avg_stress :297-319), the history update (update_int_vars :164-169 with the
  hardening of helper :172-192) and the drivers' loss loops
  (calibration/calibration_case1_singleCrystalCopper_GB.py:176-246,
  crystal_plasticity_OR_design/poly_sensitivity_rot_mesh2.py:179-239).

The unknown is the 9-component S of the reference (row-major), the rotated stiffness is
rotate_tensor_rank_4 (:21-26) of the crystal-frame cubic C at every residual, as in the
reference.  Per-point material parameters follow the ABI order (xm, gss_a, h, t_sat, r, C11,
C12, C44).

Pins (tests/test_oracle_adjoint.py): the gradient of whole load histories w.r.t. hardening
and elastic parameters, per-point initial resistances and Euler angles equals central
finite differences of the oracle forward (oracle/fe.py), which itself is pinned to the
reference-held curves.
"""
import numpy as np
import scipy.sparse.linalg as spla
import torch

from . import cp_local, fe

T = torch.float64


# --------------------------------------------------------------------------- point model
def cubic_C(C11, C12, C44):
    """(Q,3,3,3,3) crystal-frame cubic stiffness from per-point constants (differentiable);
    the tensor of cp_local.cubic_stiffness (models_copper.py:92-130)."""
    e = torch.eye(3, dtype=T)
    d4 = torch.zeros(3, 3, 3, 3, dtype=T)
    for a in range(3):
        d4[a, a, a, a] = 1.0
    dd = torch.einsum("ij,kl->ijkl", e, e)
    sym = torch.einsum("ik,jl->ijkl", e, e) + torch.einsum("il,jk->ijkl", e, e)
    c11, c12, c44 = (x.reshape(-1, 1, 1, 1, 1) for x in (C11, C12, C44))
    return c12 * dd + c44 * sym + (c11 - c12 - 2.0 * c44) * d4


def euler_zyx(angles_deg):
    """(..., 3) extrinsic z-y-x Euler angles in degrees -> R = Rx(c) Ry(b) Rz(a), differentiable;
    the rotation of cp_local.euler_zyx_to_rot (scipy ``from_euler('zyx', degrees=True)``,
    poly_sensitivity_rot_mesh2.py:41-53)."""
    a = torch.deg2rad(angles_deg)
    ca, sa, cb, sb, cc, sc = (f(a[..., k]) for k in range(3) for f in (torch.cos, torch.sin))
    z, o = torch.zeros_like(ca), torch.ones_like(ca)
    Rz = torch.stack([ca, -sa, z, sa, ca, z, z, z, o], -1).reshape(a.shape[:-1] + (3, 3))
    Ry = torch.stack([cb, z, sb, z, o, z, -sb, z, cb], -1).reshape(a.shape[:-1] + (3, 3))
    Rx = torch.stack([o, z, z, z, cc, -sc, z, sc, cc], -1).reshape(a.shape[:-1] + (3, 3))
    return Rx @ Ry @ Rz


def _prot(R, P0):
    """R P0_a R^T for all slip systems: rotate_tensor_rank_2_vmap (models_copper.py:29-32,173)."""
    return torch.einsum("qij,ajl,qkl->qaik", R, P0, R)


def _helper(S, H, Fpi_old, g_old, R, xm, P0, ao, dt):
    """tau, gamma_inc, Fp_inv_new, Fe, F of helper (models_copper.py:172-192)."""
    Prot = _prot(R, P0)
    tau = torch.einsum("qij,qaij->qa", S, Prot)
    gamma_inc = ao * dt * torch.abs(tau / g_old) ** (1.0 / xm[:, None]) * torch.sign(tau)
    Lp = torch.einsum("qa,qaij->qij", gamma_inc, Prot)
    I = torch.eye(3, dtype=T)
    Fpi_new = Fpi_old @ (I - Lp)
    F = H + I
    Fe = F @ Fpi_new
    return tau, gamma_inc, Fpi_new, Fe, F


def residual(S, H, Fpi_old, g_old, R, xm, C, P0, ao, dt):
    """implicit_residual (models_copper.py:195-201): S - rot4(R, C) : 1/2 (Fe^T Fe - I), (Q, 9)."""
    _, _, _, Fe, _ = _helper(S, H, Fpi_old, g_old, R, xm, P0, ao, dt)
    I = torch.eye(3, dtype=T)
    E = 0.5 * (Fe.transpose(1, 2) @ Fe - I)
    Crot = torch.einsum("qij,qkl,qmn,qop,qjlnp->qikmo", R, R, R, R, C)
    return (S - torch.einsum("qijkl,qkl->qij", Crot, E)).reshape(-1, 9)


class Consts:
    """Per-run constants of the point model: Schmid tensors, latent-matrix rule, ao, tol."""

    def __init__(self, schmid, self_block=True, ao=cp_local.AO, tol=1e-8, max_sub_step=5):
        self.P0 = torch.as_tensor(np.asarray(schmid, dtype=np.float64))
        self.N = self.P0.shape[0]
        self.self_block = self_block
        self.ao = ao
        self.tol = tol
        self.max_sub_step = max_sub_step

    def oracle_material(self, xm, C):
        # only xm and C enter the local solve; the hardening constants are placeholders
        return cp_local.Material(self.P0.numpy(), xm=xm, gss_a=1.0, h=0.0, t_sat=1.0, r=1.0, C=C,
                                 self_block=self.self_block, ao=self.ao)


class _LocalSolve(torch.autograd.Function):
    """S = newton_solver(x) with the IFT rule of f_jvp (models_copper.py:251-259) transposed."""

    @staticmethod
    def forward(ctx, H, Fpi_old, g_old, R, xm, C, consts, dt):
        npf = lambda t: t.detach().numpy()
        mat = consts.oracle_material(npf(xm), npf(C))
        Q = H.shape[0]
        out = cp_local.local_update(npf(H), npf(Fpi_old), npf(g_old), np.zeros((Q, consts.N)), npf(R), mat, dt,
                                    tol=consts.tol, max_sub_step=consts.max_sub_step, want_tangent=False)
        if out["capped"].any():
            raise RuntimeError(f"oracle local Newton: {int(out['capped'].sum())} points did not reach |r| <= {consts.tol}")
        S = torch.as_tensor(out["S"])
        ctx.save_for_backward(S, H, Fpi_old, g_old, R, xm, C)
        ctx.consts, ctx.dt = consts, dt
        return S

    @staticmethod
    def backward(ctx, Sbar):
        S, *x = ctx.saved_tensors
        c, dt = ctx.consts, ctx.dt

        def r_point(Sq, Hq, Fq, gq, Rq, xq, Cq):
            return residual(Sq[None], Hq[None], Fq[None], gq[None], Rq[None], xq[None], Cq[None], c.P0, c.ao, dt)[0]

        Jy = torch.func.vmap(torch.func.jacrev(r_point, argnums=0))(S, *x).reshape(-1, 9, 9)
        mu = torch.linalg.solve(Jy.transpose(1, 2), Sbar.reshape(-1, 9, 1))[..., 0]
        with torch.enable_grad():
            xs = [t.detach().requires_grad_(True) for t in x]
            r = residual(S, *xs, c.P0, c.ao, dt)
            grads = torch.autograd.grad(r, xs, grad_outputs=-mu, allow_unused=True)
        return tuple(grads) + (None, None)


def point_outputs(H, Fpi_old, g_old, slip_old, R, mat8, consts, dt):
    """P, Cauchy stress, Fp_inv_new, g_new, slip_new per point (differentiable in every input):
    first_PK_stress (:155-162), update_int_vars (:164-169) with the Kalidindi hardening of
    helper (:176-187), and P F^T / det F of compute_avg_stress (:308-313)."""
    xm, gss_a, h, t_sat, r = (mat8[:, k] for k in range(5))
    C = cubic_C(mat8[:, 5], mat8[:, 6], mat8[:, 7])
    S = _LocalSolve.apply(H, Fpi_old, g_old, R, xm, C, consts, dt)
    tau, gamma_inc, Fpi_new, Fe, F = _helper(S, H, Fpi_old, g_old, R, xm, consts.P0, consts.ao, dt)
    sigma_e = Fe @ S @ Fe.transpose(1, 2) / torch.linalg.det(Fe)[:, None, None]
    P = torch.linalg.det(F)[:, None, None] * sigma_e @ torch.linalg.inv(F).transpose(1, 2)
    cauchy = P @ F.transpose(1, 2) / torch.linalg.det(F)[:, None, None]
    x = 1.0 - g_old / t_sat[:, None]
    tmp = h[:, None] * torch.abs(gamma_inc) * torch.abs(x) ** gss_a[:, None] * torch.sign(x)
    N = consts.N
    q = r[:, None, None] * torch.ones(N, N, dtype=T)
    if consts.self_block:
        blk = torch.zeros(N, N, dtype=torch.bool)
        for i in range(N):
            blk[i, (i // 3) * 3:(i // 3) * 3 + 3] = True
        q = torch.where(blk, torch.ones((), dtype=T), q)
    g_new = g_old + torch.einsum("qab,qb->qa", q, tmp)
    return dict(P=P, sigma=cauchy, Fp_inv_new=Fpi_new, g_new=g_new, slip_new=slip_old + gamma_inc)


# --------------------------------------------------------------------------- FE layer
class Mesh:
    """Hex8 mesh geometry (oracle/fe.py) and Dirichlet data of the current load step."""

    def __init__(self, points, cells):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells, dtype=np.int64)
        self.nc = len(self.cells)
        self.ndof = 3 * len(self.points)
        sg, jxw = fe.shape_grads(self.points, self.cells)
        self.sg, self.jxw = torch.as_tensor(sg), torch.as_tensor(jxw)
        self.cell_dofs = torch.as_tensor((3 * self.cells[:, :, None] + np.arange(3)).reshape(self.nc, 24))
        self.bc_dofs = np.zeros(0, dtype=np.int64)
        self.bc_vals = np.zeros(0)

    def set_dirichlet(self, node_lists, comps, vals):
        d = {}
        for nodes, c, v in zip(node_lists, comps, vals):
            for n in np.asarray(nodes, dtype=np.int64):
                d[3 * int(n) + int(c)] = float(v)
        self.bc_dofs = np.array(sorted(d), dtype=np.int64)
        self.bc_vals = np.array([d[k] for k in self.bc_dofs])

    def u_grads(self, u):
        """(Q, 3, 3) u_grad = sum_a u_a (x) dN_a (models_copper.py:277-278)."""
        uc = u.reshape(-1, 3)[torch.as_tensor(self.cells)]
        return torch.einsum("cai,cqaj->cqij", uc, self.sg).reshape(-1, 3, 3)

    def residual(self, u, state, consts, dt):
        """Raw assembled residual r_a,i = sum_q P_ij dN_a,j JxW (no Dirichlet rows)."""
        out = point_outputs(self.u_grads(u), *state, consts, dt)
        P = out["P"].reshape(self.nc, 8, 3, 3)
        re = torch.einsum("cqij,cqaj,cq->cai", P, self.sg, self.jxw).reshape(-1)
        return torch.zeros(self.ndof, dtype=T).index_add(0, self.cell_dofs.reshape(-1), re)


def _oracle_problem(mesh, state, consts, dt):
    """oracle.fe.FEProblem holding the given history and per-point material."""
    Fpi, g, slip, R, mat8 = (t.detach().numpy() for t in state)
    Q = 8 * mesh.nc
    mat = cp_local.Material(consts.P0.numpy(), xm=mat8[:, 0], gss_a=mat8[:, 1], h=mat8[:, 2], t_sat=mat8[:, 3],
                            r=mat8[:, 4], C=cubic_C(*(torch.as_tensor(mat8[:, k]) for k in (5, 6, 7))).numpy(),
                            self_block=consts.self_block, ao=consts.ao)
    prob = fe.FEProblem(mesh.points, mesh.cells, mat, np.tile(np.eye(3), (mesh.nc, 1, 1)), np.ones(mesh.nc), dt=dt,
                        tol=consts.tol, max_sub_step=consts.max_sub_step)
    prob.R = R.reshape(mesh.nc, 8, 3, 3)
    prob.Fp_inv = Fpi.reshape(mesh.nc, 8, 3, 3)
    prob.g = g.reshape(mesh.nc, 8, -1)
    prob.slip = slip.reshape(mesh.nc, 8, -1)
    prob.bc_dofs, prob.bc_vals = mesh.bc_dofs, mesh.bc_vals
    assert Q == prob.nc * 8
    return prob


class _GlobalSolve(torch.autograd.Function):
    """fwd_pred of ad_wrapper (solver.py:856-874): u(params) by the Newton solver, with the
    backward of implicit_vjp (:801-853)."""

    @staticmethod
    def forward(ctx, mesh, consts, dt, solver_kw, u0, *state):
        prob = _oracle_problem(mesh, state, consts, dt)
        u, _ = fe.solve(prob, initial_guess=u0.detach().numpy(), **solver_kw)
        ctx.mesh, ctx.consts, ctx.dt = mesh, consts, dt
        ctx.bc = (mesh.bc_dofs.copy(), mesh.bc_vals.copy())  # the mesh's lists change with the step
        ctx.save_for_backward(torch.as_tensor(u), *state)
        return torch.as_tensor(u)

    @staticmethod
    def backward(ctx, ubar):
        u, *state = ctx.saved_tensors
        mesh, consts, dt = ctx.mesh, ctx.consts, ctx.dt
        prob = _oracle_problem(mesh, state, consts, dt)
        prob.bc_dofs, prob.bc_vals = ctx.bc
        prob.newton_update(u.numpy())
        A = prob.get_A()
        lam = spla.splu(A.T.tocsc()).solve(ubar.numpy())
        lam[ctx.bc[0]] = 0.0  # Dirichlet rows of c = apply_bc(res) are u - ubar: no p dependence
        with torch.enable_grad():
            xs = [t.detach().requires_grad_(True) for t in state]
            r = mesh.residual(u, xs, consts, dt)
            grads = torch.autograd.grad(r, xs, grad_outputs=-torch.as_tensor(lam), allow_unused=True)
        return (None, None, None, None, None) + tuple(grads)


def simulate(mesh, consts, bc_lists, disps, ts, g0, mat8, R, solver_kw=None):
    """Mean-over-cells Cauchy stress after every load step, (steps, 3, 3), differentiable in
    g0 (Q, N), mat8 (Q, 8) and R (Q, 3, 3): the drivers' loop (singlecrystal_copper.py:179-233
    with ad_wrapper, compute_avg_stress, update_int_vars_gp) under reverse-mode AD.
    bc_lists: (node_inds, comp, value or None = the driven displacement)."""
    solver_kw = dict(solver_kw or {})
    Q = 8 * mesh.nc
    Fpi = torch.eye(3, dtype=T).expand(Q, 3, 3).clone()
    g = g0
    slip = torch.zeros(Q, consts.N, dtype=T)
    u = torch.zeros(mesh.ndof, dtype=T)
    out = []
    for i in range(len(ts) - 1):
        dt = float(ts[i + 1] - ts[i])
        mesh.set_dirichlet([b[0] for b in bc_lists], [b[1] for b in bc_lists],
                           [disps[i + 1] if b[2] is None else b[2] for b in bc_lists])
        state = (Fpi, g, slip, R, mat8)
        u = _GlobalSolve.apply(mesh, consts, dt, solver_kw, u.detach(), *state)
        pts = point_outputs(mesh.u_grads(u), *state, consts, dt)
        sig = pts["sigma"].reshape(mesh.nc, 8, 3, 3)
        w = mesh.jxw
        sig_cells = torch.einsum("cqij,cq->cij", sig, w) / w.sum(dim=1)[:, None, None]
        out.append(sig_cells.mean(dim=0))
        Fpi, g, slip = pts["Fp_inv_new"], pts["g_new"], pts["slip_new"]
    return torch.stack(out)
