"""Hand-written adjoint of the crystal-plasticity path behind torch.autograd.

Two layers:

* Drop-in pieces of the reference's AD path: ``implicit_vjp`` and ``ad_wrapper``
  (crystal_plasticity_OR_design/solver.py:801-874; driver copies e.g.
  calibration_case1_singleCrystalCopper_GB.py:73-94), plus autograd Functions behind
  ``CrystalPlasticity.compute_avg_stress`` / ``update_int_vars_gp`` (what JAX derives through
  the custom_jvp local solver, models_copper.py:251-259).  With them the reference's
  calibration loop (fwd_pred -> compute_avg_stress -> update_int_vars_gp -> loss) runs under
  torch autograd unchanged in structure.
* ``simulate``: the whole load history as one autograd Function with history checkpoints
  every k steps; the backward pass recomputes the history inside each segment from the stored
  converged displacements (point updates only, no Newton solves) and sweeps the steps in
  reverse: per step two point VJPs, one re-assembly, one transposed Krylov solve.

All reverse-mode arithmetic runs in libcpfem_hip.so (cpfem_local_vjp, cpfem_ugrad_transpose,
cpfem_grad_field, cpfem_transpose, cpfem_solve).
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from .models import soa_empty, soa_ptr
from .problem import device, ptr, stream_handle, to_device
from .solver import krylov_solve, linear_solver, solver as _solver


def _soa(t, nc, comps):
    """SoA [prod(comps)][Q] view/copy of a (cells, 8, *comps) tensor (None -> None)."""
    if t is None:
        return None
    p, k = soa_ptr(t, nc, comps)
    return k.reshape(int(np.prod(comps)) if comps else 1, nc * 8)


def _from_soa(a, nc, comps):
    """(cells, 8, *comps) view of an SoA [prod(comps)][Q] tensor."""
    ncomp = len(comps)
    return a.reshape(tuple(comps) + (nc, 8)).permute(ncomp, ncomp + 1, *range(ncomp))


def local_vjp(problem, dofs, params, Pbar=None, sbar_cell=None, Fbar=None, gbar=None, slipbar=None,
              mask=N.VJP_H | N.VJP_FPI | N.VJP_G | N.VJP_MAT | N.VJP_R, params_struct=None):
    """Point VJP through the converged local solve.  Cotangents are SoA device tensors
    (Pbar, Fbar [9][Q]; gbar, slipbar [N][Q]; sbar_cell [9][n_cells]).  Returns a dict with
    Hbar, Fpibar [9][Q], gbar_old [N][Q], pbar [8][Q] (xm, gss_a, h, t_sat, r, C11, C12, C44),
    Rbar [9][Q] for the groups in mask."""
    ctx = problem.ctx
    Q, Ns = ctx.n_qp, problem.num_slip_sys
    if params_struct is None:
        st, pr, keep = problem._marshal(params)
    else:
        st, pr, keep = params_struct
    z = lambda r: torch.zeros(r, Q, dtype=torch.float64, device=device())
    out = dict(Hbar=z(9) if mask & N.VJP_H else None, Fpibar=z(9) if mask & N.VJP_FPI else None,
               gbar_old=z(Ns) if mask & N.VJP_G else None, pbar=z(8) if mask & N.VJP_MAT else None,
               Rbar=z(9) if mask & N.VJP_R else None)
    v = N.CpfemVjp(ptr(Pbar), ptr(sbar_cell), ptr(Fbar), ptr(gbar), ptr(slipbar), ptr(out["Hbar"]),
                   ptr(out["Fpibar"]), ptr(out["gbar_old"]), ptr(out["pbar"]), ptr(out["Rbar"]), mask, 0)
    N.check(ctx._lib.cpfem_local_vjp(ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st), ctypes.byref(pr),
                                     ctypes.byref(v)))
    return out


def ugrad_transpose(problem, Hbar):
    out = torch.empty(problem.ctx.n_dofs, dtype=torch.float64, device=device())
    N.check(problem.ctx._lib.cpfem_ugrad_transpose(problem.ctx.h, stream_handle(), ptr(Hbar), ptr(out)))
    return out


def grad_field(problem, lam, scale=-1.0):
    out = torch.empty(9, problem.ctx.n_qp, dtype=torch.float64, device=device())
    N.check(problem.ctx._lib.cpfem_grad_field(problem.ctx.h, stream_handle(), ptr(lam), float(scale), ptr(out)))
    return out


def transpose(problem, mat):
    matT = torch.empty_like(mat)
    N.check(problem.ctx._lib.cpfem_transpose(problem.ctx.h, stream_handle(), ptr(mat), ptr(matT)))
    return matT


def adjoint_solve(problem, dofs, v, params=None, solver_options=None):
    """lambda with A^T lambda = v, A the row-eliminated Jacobian at (dofs, params)."""
    _, mat = problem.assemble(dofs, params=params)
    matT = transpose(problem, mat)
    x0 = torch.zeros_like(v)
    # the same dispatch as the forward solve (solver.py:844 passes adjoint_solver_options to
    # linear_solver), on a copy: linear_solver adds 'jax_solver' when no solver is named
    return linear_solver(problem, matT, v.contiguous(), x0, dict(solver_options or {}))


def _param_cotangents(problem, params, vj, nc):
    """Cotangent list shaped like the reference's internal_vars from point-VJP outputs."""
    Ns = problem.num_slip_sys
    res = [_from_soa(vj["Fpibar"], nc, (3, 3)), _from_soa(vj["gbar_old"], nc, (Ns,)),
           torch.zeros(nc, 8, Ns, dtype=torch.float64, device=device()), _from_soa(vj["Rbar"], nc, (3, 3))]
    pb = vj["pbar"]
    if len(params) >= 9:  # 9-array order gss_a, h, t_sat, xm, r
        for idx in (N.GSS_A, N.H, N.TSAT, N.XM, N.R):
            res.append(pb[idx].reshape(nc, 8))
    if len(params) >= 10:
        Cb = torch.zeros(nc, 8, 3, 3, 3, 3, dtype=torch.float64, device=device())
        Cb[..., 0, 0, 0, 0] = pb[N.C11].reshape(nc, 8)
        Cb[..., 0, 0, 1, 1] = pb[N.C12].reshape(nc, 8)
        Cb[..., 1, 2, 1, 2] = pb[N.C44].reshape(nc, 8)
        res.append(Cb)
    return res


def implicit_vjp(problem, sol_list, params, v_list, adjoint_solver_options=None):
    """solver.py:801-853: A^T lambda = v at the converged solution, then
    -(dc/dparams)^T lambda, returned shaped like params."""
    problem.set_params(params)
    dofs = problem.flatten(sol_list)
    v = problem.flatten(v_list)
    lam = adjoint_solve(problem, dofs, v, params, adjoint_solver_options)
    Pbar = grad_field(problem, lam, -1.0)
    vj = local_vjp(problem, dofs, params, Pbar=Pbar, mask=N.VJP_FPI | N.VJP_G | N.VJP_MAT | N.VJP_R)
    return _param_cotangents(problem, params, vj, problem.fes[0].num_cells)


# ----------------------------------------------------------------------------- autograd Functions
class _SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, solver_options, adjoint_solver_options, *params):
        problem.set_params(list(params))
        sol = _solver(problem, solver_options)[0]
        ctx.problem, ctx.adj = problem, adjoint_solver_options
        ctx.params = params
        ctx.save_for_backward(sol)
        return sol

    @staticmethod
    def backward(ctx, v):
        (sol,) = ctx.saved_tensors
        cot = implicit_vjp(ctx.problem, [sol], list(ctx.params), [v], ctx.adj)
        return (None, None, None) + tuple(cot[: len(ctx.params)])


def ad_wrapper(problem, solver_options=None, adjoint_solver_options=None):
    """fwd_pred(params) -> [sol] differentiable with respect to params (solver.py:856-874)."""
    solver_options = dict(solver_options or {"jax_solver": {}})
    adj = dict(adjoint_solver_options or {})

    def fwd_pred(params):
        opts = dict(solver_options)
        if getattr(problem, "initial_guess", None) is not None:
            opts["initial_guess"] = problem.initial_guess
        sol = _SolveFn.apply(problem, opts, adj, *params)
        problem.set_initial_guess([sol.detach()])
        return [sol]

    return fwd_pred


class _AvgStressFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, sol, *params):
        ctx.problem = problem
        ctx.save_for_backward(sol, *params)
        return problem._compute_avg_stress_raw(sol, list(params))

    @staticmethod
    def backward(ctx, sbar):
        problem = ctx.problem
        sol, *params = ctx.saved_tensors
        nc = problem.fes[0].num_cells
        dofs = problem.flatten(sol)
        sbc = sbar.reshape(nc, 9).t().contiguous()
        vj = local_vjp(problem, dofs, params, sbar_cell=sbc)
        ubar = ugrad_transpose(problem, vj["Hbar"]).reshape(sol.shape)
        cot = _param_cotangents(problem, params, vj, nc)
        return (None, ubar) + tuple(cot[: len(params)])


class _UpdateFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, sol, *params):
        ctx.problem = problem
        ctx.save_for_backward(sol, *params)
        new = problem._update_raw(sol, list(params))
        return tuple(new[:3])

    @staticmethod
    def backward(ctx, Fb, gb, slb):
        problem = ctx.problem
        sol, *params = ctx.saved_tensors
        nc, Ns = problem.fes[0].num_cells, problem.num_slip_sys
        dofs = problem.flatten(sol)
        Fs = _soa(Fb, nc, (3, 3)) if Fb is not None else None
        gs = _soa(gb, nc, (Ns,)) if gb is not None else None
        ss = _soa(slb, nc, (Ns,)) if slb is not None else None
        vj = local_vjp(problem, dofs, params, Fbar=Fs, gbar=gs, slipbar=ss)
        ubar = ugrad_transpose(problem, vj["Hbar"]).reshape(sol.shape)
        cot = _param_cotangents(problem, params, vj, nc)
        if slb is not None:
            cot[2] = slb
        return (None, ubar) + tuple(cot[: len(params)])


def differentiable(params, sol=None):
    return torch.is_grad_enabled() and any(isinstance(p, torch.Tensor) and p.requires_grad
                                           for p in list(params) + ([sol] if sol is not None else []))


# ----------------------------------------------------------------------------- whole history
class LoadHistory:
    """Loading of a simulation: displacement / time histories and the Dirichlet-info builder
    (bc_info(disp) -> dirichlet_bc_info)."""

    def __init__(self, disps, ts, bc_info):
        self.disps = np.asarray(disps, dtype=np.float64)
        self.ts = np.asarray(ts, dtype=np.float64)
        self.bc_info = bc_info

    @property
    def n_steps(self):
        return len(self.ts) - 1

    def apply(self, problem, i):
        """Set dt and the Dirichlet values of step i (0-based, i.e. the (i+1)-th load step)."""
        problem.dt = float(self.ts[i + 1] - self.ts[i])
        problem.fes[0].update_Dirichlet_boundary_conditions(self.bc_info(self.disps[i + 1]))


def _mat_struct(problem, Fpi, g, slip, R, mat8):
    """CpfemState/CpfemParams with all eight material parameters per point (SoA [8][Q])."""
    st = N.CpfemState(ptr(Fpi), ptr(g), ptr(slip), ptr(R))
    pr = N.CpfemParams()
    Q = problem.ctx.n_qp
    for k in range(N.NPARAM):
        pr.ptr[k] = mat8.data_ptr() + 8 * k * Q
        pr.stride[k] = 1
    m = problem.material
    pr.gamma0 = m.ao
    pr.dt = float(problem.dt)
    pr.tol = m.tol
    pr.max_sub_step = m.max_sub_step
    pr.max_iter = m.max_iter
    pr.self_block = 1 if m.self_block else 0
    return st, pr, (Fpi, g, slip, R, mat8)


class _HistoryFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, load, k, solver_options, adj_options, g0, mat8, R):
        nc, Ns = problem.fes[0].num_cells, problem.num_slip_sys
        Q = 8 * nc
        g0s = _soa(g0, nc, (Ns,)).contiguous()
        ms = mat8.reshape(Q, 8).t().contiguous()
        Rs = _soa(R, nc, (3, 3)).contiguous()
        Fpi = soa_empty(nc, (3, 3))
        Fpi.copy_(torch.eye(3, dtype=torch.float64, device=device()).expand(nc, 8, 3, 3))
        Fpi = _soa(Fpi, nc, (3, 3)).contiguous()
        g = g0s.clone()
        slip = torch.zeros_like(g)
        checkpoints = {0: (Fpi.clone(), g.clone(), slip.clone())}
        us, sig = [], []
        dofs = torch.zeros(problem.ctx.n_dofs, dtype=torch.float64, device=device())
        for i in range(load.n_steps):
            load.apply(problem, i)
            problem._override = _mat_struct(problem, Fpi, g, slip, Rs, ms)
            try:
                opts = {"jax_solver": {}, "initial_guess": dofs.reshape(-1, 3)}
                opts.update(solver_options)
                dofs = _solver(problem, opts)[0].reshape(-1).clone()
                st, pr, _ = problem._override
                s = torch.empty(nc, 3, 3, dtype=torch.float64, device=device())
                N.check(problem.ctx._lib.cpfem_cell_stress(problem.ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st),
                                                           ctypes.byref(pr), ptr(s)))
                sig.append(s.mean(dim=0))
                Fn, gn, sn = torch.empty_like(Fpi), torch.empty_like(g), torch.empty_like(slip)
                out = N.CpfemStateOut(ptr(Fn), ptr(gn), ptr(sn))
                N.check(problem.ctx._lib.cpfem_update_state(problem.ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st),
                                                            ctypes.byref(pr), ctypes.byref(out)))
            finally:
                problem._override = None
            Fpi, g, slip = Fn, gn, sn
            us.append(dofs)
            if (i + 1) % k == 0 and i + 1 < load.n_steps:
                checkpoints[i + 1] = (Fpi.clone(), g.clone(), slip.clone())
        ctx.problem, ctx.load, ctx.k, ctx.adj = problem, load, k, adj_options
        ctx.us, ctx.checkpoints = us, checkpoints
        ctx.save_for_backward(ms, Rs)
        ctx.shape = (nc, Ns)
        return torch.stack(sig)

    @staticmethod
    def backward(ctx, sbar):
        problem, load, k = ctx.problem, ctx.load, ctx.k
        ms, Rs = ctx.saved_tensors
        nc, Ns = ctx.shape
        Q = 8 * nc
        # Input directions of the point VJPs: only those some requested gradient depends on.
        # Fp_inv_old and g_old carry the adjoint to the previous load step (and g_old at step 0
        # to g0); the material and rotation groups feed mat8 and R.  Each direction is one
        # dual-number pass through the point's residual, so an orientation gradient over one
        # load step evaluates 6 + 18 and 6 + 9 passes per point instead of 6 + 59 and 6 + 50.
        need_g0, need_mat, need_R = ctx.needs_input_grad[5], ctx.needs_input_grad[6], ctx.needs_input_grad[7]
        own = (N.VJP_MAT if need_mat else 0) | (N.VJP_R if need_R else 0)
        Fbar = torch.zeros(9, Q, dtype=torch.float64, device=device())
        gbar = torch.zeros(Ns, Q, dtype=torch.float64, device=device())
        pbar_tot = torch.zeros(8, Q, dtype=torch.float64, device=device())
        Rbar_tot = torch.zeros(9, Q, dtype=torch.float64, device=device())
        n = load.n_steps
        starts = sorted(ctx.checkpoints)
        for seg_start in reversed(starts):
            seg_end = min(seg_start + k, n)
            # recompute the history states h_{seg_start} .. h_{seg_end-1} from the checkpoint
            states = [ctx.checkpoints[seg_start]]
            for i in range(seg_start, seg_end - 1):
                load.apply(problem, i)
                Fpi, g, slip = states[-1]
                st, pr, _ = _mat_struct(problem, Fpi, g, slip, Rs, ms)
                Fn, gn, sn = torch.empty_like(Fpi), torch.empty_like(g), torch.empty_like(slip)
                out = N.CpfemStateOut(ptr(Fn), ptr(gn), ptr(sn))
                N.check(problem.ctx._lib.cpfem_update_state(problem.ctx.h, stream_handle(), ptr(ctx.us[i]),
                                                            ctypes.byref(st), ctypes.byref(pr), ctypes.byref(out)))
                states.append((Fn, gn, sn))
            for i in range(seg_end - 1, seg_start - 1, -1):
                load.apply(problem, i)
                Fpi, g, slip = states[i - seg_start]
                ps = _mat_struct(problem, Fpi, g, slip, Rs, ms)
                dofs = ctx.us[i]
                sbc = (sbar[i].reshape(9, 1) / nc).expand(9, nc).contiguous()
                mask = own | (N.VJP_FPI | N.VJP_G if i > 0 else (N.VJP_G if need_g0 else 0))
                v1 = local_vjp(problem, dofs, None, sbar_cell=sbc, Fbar=Fbar, gbar=gbar, mask=N.VJP_H | mask,
                               params_struct=ps)
                ubar = ugrad_transpose(problem, v1["Hbar"])
                problem._override = ps
                try:
                    lam = adjoint_solve(problem, dofs, ubar, None, ctx.adj)
                finally:
                    problem._override = None
                Pbar = grad_field(problem, lam, -1.0)
                vs = [v1]
                if mask:
                    vs.append(local_vjp(problem, dofs, None, Pbar=Pbar, mask=mask, params_struct=ps))
                if mask & N.VJP_FPI:
                    Fbar = sum(v["Fpibar"] for v in vs)
                if mask & N.VJP_G:
                    gbar = sum(v["gbar_old"] for v in vs)
                if need_mat:
                    pbar_tot += sum(v["pbar"] for v in vs)
                if need_R:
                    Rbar_tot += sum(v["Rbar"] for v in vs)
        g0bar = _from_soa(gbar, nc, (Ns,)) if need_g0 else None
        mbar = pbar_tot.t().reshape(nc, 8, 8) if need_mat else None
        Rbar = _from_soa(Rbar_tot, nc, (3, 3)) if need_R else None
        return None, None, None, None, None, g0bar, mbar, Rbar


def simulate(problem, load, g0, mat8, R, checkpoint_every=10, solver_options=None, adjoint_solver_options=None):
    """Mean-over-cells Cauchy stress after every load step, (n_steps, 3, 3), differentiable
    with respect to the initial slip resistance g0 (cells, 8, N), the per-point material
    parameters mat8 (cells, 8, 8: xm, gss_a, h, t_sat, r, C11, C12, C44) and the per-point
    rotations R (cells, 8, 3, 3).  History is checkpointed every `checkpoint_every` steps."""
    return _HistoryFn.apply(problem, load, int(checkpoint_every), dict(solver_options or {}),
                            dict(adjoint_solver_options or {}), g0, mat8, R)
