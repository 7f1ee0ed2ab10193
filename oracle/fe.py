"""Oracle: hex8 finite elements, assembly, Dirichlet BCs, Jacobi-BiCGSTAB, global Newton.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates (FP64, NumPy/SciPy):
  * the upstream jax-fem hex8 element used by the reference (external library,
    deepmodeling/jax-fem, version unpinned in the reference; its API era is fixed by
    the vendored ``crystal_plasticity_OR_design/solver.py`` and the "JAX (version
    0.4.13)" note at ``models_copper.py:206``): trilinear Lagrange shape functions on
    [0,1]^3, 2x2x2 Gauss-Legendre points, physical gradients J^-T dN/dxi,
    JxW = det J * w; total-Lagrangian weak form r_a,i = sum_q P_ij dN_a/dX_j JxW;
    element Jacobian = exact derivative; COO assembly with summed duplicates.
    Reference call sites: models_copper.py:277-278,315; solver.py:279-304,392.
  * ``apply_bc_vec`` / ``assign_bc`` / ``copy_bc`` (solver.py:119-192),
    ``get_A`` row elimination (solver.py:279-304),
  * ``jax_solve``: jax.scipy.sparse.linalg.bicgstab (JAX 0.4.x, the published
    algorithm of jax/_src/scipy/sparse/linalg.py, "preconditioned BiCGSTAB" of
    Wikipedia) with tol=atol=1e-10, maxiter=10000, Jacobi M, and the
    |Ax-b| < 0.1 check (solver.py:19-48),
  * ``linear_incremental_solver``, ``line_search`` and ``solver`` (solver.py:213-437),
  * the load-step driver loop (singlecrystal_copper.py:179-233).

Pins (tests/test_oracle_fe.py, tests/test_oracle_golden.py): sum JxW = volume,
partition of unity of dN, the patch test (u = H X gives u_grad = H), element
stiffness vs central FD of the element residual, elastic stiffness symmetry with
six rigid-body null modes, BiCGSTAB vs scipy spsolve, and the reference-held
single-crystal golden curves end to end.
"""
import numpy as np
import scipy.sparse as sp

from . import cp_local

# ----------------------------------------------------------------------------- hex8
# Local node coordinates on the unit reference cube, VTK / meshio hexahedron order.
HEX8_NODES = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                       [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
_G = 0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)
# Tensor-product Gauss points, first coordinate slowest.
HEX8_QP = np.array([[a, b, c] for a in _G for b in _G for c in _G])
HEX8_QW = np.full(8, 0.125)


def hex8_shape(xi):
    """N_a(xi) (8,) and dN_a/dxi (8,3) at reference point xi in [0,1]^3."""
    xi = np.asarray(xi, dtype=np.float64)
    f = np.where(HEX8_NODES == 1, xi, 1.0 - xi)          # (8,3) 1D factors
    df = np.where(HEX8_NODES == 1, 1.0, -1.0)
    N = f.prod(axis=1)
    dN = np.empty((8, 3))
    dN[:, 0] = df[:, 0] * f[:, 1] * f[:, 2]
    dN[:, 1] = f[:, 0] * df[:, 1] * f[:, 2]
    dN[:, 2] = f[:, 0] * f[:, 1] * df[:, 2]
    return N, dN


def shape_grads(points, cells):
    """Physical shape gradients (cells,8q,8a,3) and JxW (cells,8q)."""
    X = points[cells]  # (C,8,3)
    dNr = np.stack([hex8_shape(x)[1] for x in HEX8_QP])  # (8q,8a,3)
    J = np.einsum("cai,qaj->cqij", X, dNr)               # dx_i/dxi_j
    detJ = np.linalg.det(J)
    Jinv = np.linalg.inv(J)
    sg = np.einsum("qaj,cqji->cqai", dNr, Jinv)          # dN/dX_i = dN/dxi_j dxi_j/dX_i
    return sg, detJ * HEX8_QW[None, :]


def box_mesh(Nx, Ny, Nz, Lx, Ly, Lz):
    """Structured hex8 box; node id = i + (Nx+1)(j + (Ny+1)k); cells in VTK order."""
    xs = np.linspace(0.0, Lx, Nx + 1)
    ys = np.linspace(0.0, Ly, Ny + 1)
    zs = np.linspace(0.0, Lz, Nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)

    def nid(i, j, k):
        return i + (Nx + 1) * (j + (Ny + 1) * k)

    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    cells = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                      nid(i, j + 1, k + 1)], axis=1).astype(np.int64)
    return points, cells


def u_grads(sol, cells, sg):
    """u_grad[c,q] = sum_a u_a (x) dN_a; models_copper.py:277-278."""
    return np.einsum("cai,cqaj->cqij", sol[cells], sg)


# ----------------------------------------------------------------------------- problem
class FEProblem:
    """Mesh + material + history, the oracle counterpart of the reference Problem."""

    def __init__(self, points, cells, mat, R_cell, g0, dt=1.0, tol=1e-8, max_sub_step=5):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells, dtype=np.int64)
        self.nc = len(self.cells)
        self.nn = len(self.points)
        self.ndof = 3 * self.nn
        self.sg, self.JxW = shape_grads(self.points, self.cells)
        self.mat = mat
        Q = self.nc * 8
        self.R = np.repeat(np.asarray(R_cell, dtype=np.float64).reshape(self.nc, 1, 3, 3), 8, axis=1)
        self.Fp_inv = np.broadcast_to(np.eye(3), (self.nc, 8, 3, 3)).copy()
        g0 = np.asarray(g0, dtype=np.float64)
        if g0.ndim == 1:  # one initial resistance per cell
            g0 = g0[:, None, None]
        self.g = np.broadcast_to(g0, (self.nc, 8, mat.N)).copy()
        self.slip = np.zeros((self.nc, 8, mat.N))
        self.dt = dt
        self.tol = tol
        self.max_sub_step = max_sub_step
        self.bc_dofs = np.zeros(0, dtype=np.int64)
        self.bc_vals = np.zeros(0)
        # COO pattern: models jax-fem's I, J (per-cell dof outer products)
        cd = (3 * self.cells[:, :, None] + np.arange(3)).reshape(self.nc, 24)
        self.I = np.repeat(cd, 24, axis=1).ravel()
        self.J = np.tile(cd, (1, 24)).ravel()
        self.cell_dofs = cd
        self.n_newton = 0
        self.krylov_iters = []

    # --- Dirichlet: node_inds_list / vec_inds_list / vals_list semantics
    def set_dirichlet(self, node_lists, comps, vals):
        """Resolve the BC lists to unique dofs; a later list wins on a shared dof
        (the .set/.add order of apply_bc_vec / assign_bc, solver.py:119-153)."""
        d = {}
        for nodes, c, v in zip(node_lists, comps, vals):
            for n in np.asarray(nodes, dtype=np.int64):
                d[3 * int(n) + int(c)] = float(v)
        keys = np.array(sorted(d), dtype=np.int64)
        self.bc_dofs = keys
        self.bc_vals = np.array([d[k] for k in keys])

    def _flat(self, a):
        return a.reshape((self.nc * 8,) + a.shape[2:])

    def local(self, sol, want_tangent=True, Fp_inv=None, g=None, slip=None):
        ug = u_grads(sol.reshape(-1, 3), self.cells, self.sg)
        Fp_inv = self.Fp_inv if Fp_inv is None else Fp_inv
        g = self.g if g is None else g
        slip = self.slip if slip is None else slip
        return cp_local.local_update(self._flat(ug), self._flat(Fp_inv), self._flat(g),
                                     self._flat(slip), self._flat(self.R), self.mat, self.dt,
                                     tol=self.tol, max_sub_step=self.max_sub_step,
                                     want_tangent=want_tangent), ug

    def newton_update(self, dofs):
        """Residual (no BC) and element Jacobians V; jax-fem problem.newton_update."""
        out, _ = self.local(dofs, want_tangent=True)
        P = out["P"].reshape(self.nc, 8, 3, 3)
        A = out["A"].reshape(self.nc, 8, 3, 3, 3, 3)
        sg, w = self.sg, self.JxW
        re = np.einsum("cqij,cqaj,cq->cai", P, sg, w).reshape(self.nc, 24)
        Ke = np.einsum("cqaj,cqijkl,cqbl,cq->caibk", sg, A, sg, w).reshape(self.nc, 24, 24)
        res = np.zeros(self.ndof)
        np.add.at(res, self.cell_dofs.ravel(), re.ravel())
        self.V = Ke.ravel()
        self.n_newton += 1
        return res

    def compute_residual(self, dofs):
        out, _ = self.local(dofs, want_tangent=False)
        P = out["P"].reshape(self.nc, 8, 3, 3)
        re = np.einsum("cqij,cqaj,cq->cai", P, self.sg, self.JxW).reshape(self.nc, 24)
        res = np.zeros(self.ndof)
        np.add.at(res, self.cell_dofs.ravel(), re.ravel())
        return res

    def apply_bc_vec(self, res, dofs):
        """res[bc] = u[bc] - ubar; solver.py:119-133."""
        res = res.copy()
        res[self.bc_dofs] = dofs[self.bc_dofs] - self.bc_vals
        return res

    def get_A(self):
        """CSR with summed duplicates, BC rows zeroed with unit diagonal; solver.py:279-304."""
        A = sp.csr_matrix((self.V, (self.I, self.J)), shape=(self.ndof, self.ndof))
        A.sum_duplicates()
        mask = np.zeros(self.ndof, dtype=bool)
        mask[self.bc_dofs] = True
        rows = np.repeat(np.arange(self.ndof), np.diff(A.indptr))
        kill = mask[rows]
        A.data[kill] = 0.0
        A.data[kill & (A.indices == rows)] = 1.0
        return A

    # --- post-processing / history, models_copper.py:273-319
    def cell_stress(self, dofs):
        out, _ = self.local(dofs, want_tangent=False)
        sig = out["sigma"].reshape(self.nc, 8, 3, 3)
        return np.einsum("cqij,cq->cij", sig, self.JxW) / self.JxW.sum(axis=1)[:, None, None]

    def update_history(self, dofs):
        out, _ = self.local(dofs, want_tangent=False)
        self.Fp_inv = out["Fp_inv_new"].reshape(self.nc, 8, 3, 3)
        self.g = out["g_new"].reshape(self.nc, 8, -1)
        self.slip = out["slip_new"].reshape(self.nc, 8, -1)
        return out


# ----------------------------------------------------------------------------- Krylov
def bicgstab_jax(A, b, x0, M=None, tol=1e-10, atol=1e-10, maxiter=10000):
    """jax.scipy.sparse.linalg.bicgstab semantics (JAX 0.4.x), called at solver.py:34-40.

    Stops when |r|^2 <= max(tol^2 |b|^2, atol^2) on the recurrence residual;
    breakdown (rho=0 / alpha=0 / omega=0) ends the loop with a negative counter.
    Returns (x, iterations, breakdown_code).
    """
    Mf = (lambda v: v) if M is None else M
    bs = b @ b
    atol2 = max(tol * tol * bs, atol * atol)
    x = x0.copy()
    r = b - A @ x
    rhat = r.copy()
    p = r.copy()
    q = r.copy()
    alpha = omega = rho = 1.0
    k = 0
    while (r @ r) > atol2 and k < maxiter and k >= 0:
        rho_ = rhat @ r
        beta = rho_ / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = Mf(p)
        q = A @ phat
        alpha_ = rho_ / (rhat @ q)
        s = r - alpha_ * q
        exit_early = (s @ s) < atol2
        shat = Mf(s)
        t = A @ shat
        omega_ = (t @ s) / (t @ t)
        if exit_early:
            x = x + alpha_ * phat
            r = s
        else:
            x = x + (alpha_ * phat + omega_ * shat)
            r = s - omega_ * t
        kn = -11 if (omega_ == 0 or alpha_ == 0) else k + 1
        kn = -10 if rho_ == 0 else kn
        k, rho, alpha, omega = kn, rho_, alpha_, omega_
    return x, k


def jax_solve(A, b, x0, precond=True, tol=1e-10, atol=1e-10, maxiter=10000):
    """solver.py:19-48: Jacobi from the diagonal, BiCGSTAB, |Ax-b| < 0.1 assertion."""
    inv_d = 1.0 / A.diagonal()
    M = (lambda v: v * inv_d) if precond else None
    x, k = bicgstab_jax(A, b, x0, M=M, tol=tol, atol=atol, maxiter=maxiter)
    err = np.linalg.norm(A @ x - b)
    assert err < 0.1, f"JAX linear solver failed to converge with err = {err}"
    return x, k


def line_search(prob, dofs, inc):
    """solver.py:240-276 (three halvings of the full step on the BC residual norm)."""
    def res_norm(alpha):
        d = dofs + alpha * inc
        return np.linalg.norm(prob.apply_bc_vec(prob.compute_residual(d), d))

    alpha = 1.0
    rn = res_norm(alpha)
    for _ in range(3):
        alpha *= 0.5
        rh = res_norm(alpha)
        if rh > rn:
            alpha *= 2.0
            break
        rn = rh
    return dofs + alpha * inc


def solve(prob, initial_guess=None, tol=1e-6, rel_tol=1e-8, line_search_flag=False,
          krylov=None, max_newton=200):
    """Row-elimination Newton; solver.py:310-437 with linear_incremental_solver :213-237.
    ``max_newton`` only guards the oracle (the reference loop has no cap) and raises."""
    dofs = np.zeros(prob.ndof) if initial_guess is None else np.asarray(initial_guess, float).ravel().copy()
    krylov = krylov or {}

    def helper(d):
        res = prob.apply_bc_vec(prob.newton_update(d), d)
        return res, prob.get_A()

    res, A = helper(dofs)
    res_val = np.linalg.norm(res)
    res0 = res_val
    rel = res_val / res0
    its = 0
    while rel > rel_tol and res_val > tol:
        if its >= max_newton:
            raise RuntimeError(f"oracle Newton did not converge in {max_newton} iterations (res {res_val})")
        b = -res
        x0 = np.zeros(prob.ndof)
        x0[prob.bc_dofs] = prob.bc_vals - dofs[prob.bc_dofs]
        inc, k = jax_solve(A, b, x0, **krylov)
        prob.krylov_iters.append(k)
        dofs = line_search(prob, dofs, inc) if line_search_flag else dofs + inc
        res, A = helper(dofs)
        res_val = np.linalg.norm(res)
        rel = res_val / res0
        its += 1
    assert np.all(np.isfinite(res_val)), "res_val contains NaN, stop the program!"
    assert np.all(np.isfinite(dofs)), "dofs contains NaN, stop the program!"
    return dofs, its


# ----------------------------------------------------------------------------- driver
def faces(points, L, atol=1e-5):
    x, y, z = points[:, 0], points[:, 1], points[:, 2]
    Lx, Ly, Lz = L
    c = np.isclose
    return dict(left=np.nonzero(c(x, 0.0, atol=atol))[0], right=np.nonzero(c(x, Lx, atol=atol))[0],
                front=np.nonzero(c(y, 0.0, atol=atol))[0], back=np.nonzero(c(y, Ly, atol=atol))[0],
                bottom=np.nonzero(c(z, 0.0, atol=atol))[0], top=np.nonzero(c(z, Lz, atol=atol))[0],
                corner=np.nonzero(c(x, 0.0, atol=atol) & c(y, 0.0, atol=atol) & c(z, 0.0, atol=atol))[0])


def run_load_steps(prob, bc_lists, disps, ts, solver_kw=None, record=None):
    """The driver loop of singlecrystal_copper.py:179-233 (set dt, BCs, solve,
    cell-averaged Cauchy stress, history update).

    ``bc_lists`` = list of (node_inds, comp, value_or_None); None = the driven
    value disps[i+1].  Returns per-step mean-over-cells sigma (steps,3,3).
    """
    solver_kw = solver_kw or {}
    sol = np.zeros(prob.ndof)
    out = []
    for i in range(len(ts) - 1):
        prob.dt = ts[i + 1] - ts[i]
        prob.set_dirichlet([b[0] for b in bc_lists], [b[1] for b in bc_lists],
                           [disps[i + 1] if b[2] is None else b[2] for b in bc_lists])
        sol, _ = solve(prob, initial_guess=sol, **solver_kw)
        sig = prob.cell_stress(sol)
        out.append(sig.mean(axis=0))
        if record is not None:
            record.append(dict(sol=sol.copy(), sig_cells=sig))
        prob.update_history(sol)
    return np.array(out)
