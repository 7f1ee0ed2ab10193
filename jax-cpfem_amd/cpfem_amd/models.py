"""CrystalPlasticity(Problem): the model-class surface of the reference, on the HIP library.

Surface kept from singlecrystal_copper/models_copper.py:51-319 (and the 9/10-array variants
calibration/case1.py, polycrystal_DPsteel/models_DPsteel_inhomo.py):
``custom_init(quat, cell_ori_inds)``, ``get_tensor_map()``, ``get_maps()``,
``update_int_vars_gp(sol, params)``, ``set_params(params)``, ``set_initial_guess``,
``inspect_interval_vars(params)``, ``compute_avg_stress(sol, params)``, attributes
``internal_vars``, ``dt``, ``fes[0]``, ``Schmid_tensors``, ``q``, ``C``.

``internal_vars`` is a list of torch float64 device tensors shaped like the reference's
arrays ((cells, 8, ...)), whose storage is structure-of-arrays ([component][cells*8]) so the
kernels read them coalesced without copies.  Element-wise arithmetic on them (the
calibration drivers' ``params[k] = coeff * params[k]``) keeps that storage order.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from . import materials as M
from .problem import Problem, device, ptr, stream_handle, to_device
from .rotations import get_rot_mat


def soa_empty(n_cells, comps):
    """(n_cells, 8, *comps) float64 device tensor with storage [prod(comps)][n_cells*8]."""
    nc = len(comps)
    base = torch.empty(tuple(comps) + (n_cells, 8), dtype=torch.float64, device=device())
    return base.permute(nc, nc + 1, *range(nc))


def soa_ptr(t, n_cells, comps):
    """Pointer to the SoA storage of t, converting (one copy) if its strides differ.
    Returns (pointer, keepalive)."""
    comps = tuple(comps)
    t = to_device(t)
    if tuple(t.shape) != (n_cells, 8) + comps:
        t = t.reshape((n_cells, 8) + comps)
    nc = len(comps)
    if nc == 0:
        t = t.contiguous()
        return ptr(t), t
    perm = tuple(range(2, 2 + nc)) + (0, 1)
    tp = t.permute(perm)
    if not tp.is_contiguous():
        tp = tp.contiguous()
    return ptr(tp), tp


class CrystalPlasticity(Problem):
    """Kalidindi crystal plasticity on hex8 meshes (HIP kernels for every point and element)."""

    def __init__(self, mesh, vec=3, dim=3, ele_type="HEX8", gauss_order=None, dirichlet_bc_info=None,
                 location_fns=None, additional_info=(), material=None):
        self.material = material if material is not None else M.COPPER
        self.solver_info = []
        super().__init__(mesh, vec, dim, ele_type, gauss_order, dirichlet_bc_info, location_fns, additional_info)

    # ------------------------------------------------------------------ init
    def custom_init(self, quat, cell_ori_inds, phase_inds=None, phase_materials=None):
        """Slip system, latent matrix, stiffness and the initial internal variables
        (models_copper.py:52-133).  Orientation indices are clamped to the quaternion table,
        as JAX gathers clamp out-of-range indices (polycrystal_DPsteel_inhomo.py:73,83).
        With ``phase_inds`` (per cell) and ``phase_materials`` the 10-array layout of
        models_DPsteel_inhomo.py:54-229 is built (per-point constants and stiffness)."""
        mat = self.material
        fe = self.fes[0]
        table = M.load_slip_table(mat.slip)
        self.num_slip_sys = len(table)
        self.slip_directions, self.slip_normals = M.slip_vectors(table)
        self.Schmid_tensors = np.einsum("ai,aj->aij", self.slip_directions, self.slip_normals)
        r = mat.r
        self.q = r * np.ones((self.num_slip_sys, self.num_slip_sys))
        if mat.self_block:
            for i in range(self.num_slip_sys):
                for j in range(3):
                    self.q[i, i // 3 * 3 + j] = 1.0
        self.C = M.cubic_stiffness(mat.C11, mat.C12, mat.C44)
        self._attach(self.slip_directions, self.slip_normals)
        nc, Ns = fe.num_cells, self.num_slip_sys
        quat = np.asarray(quat, dtype=np.float64).reshape(-1, 4)
        inds = np.clip(np.asarray(cell_ori_inds, dtype=np.int64).reshape(-1), 0, len(quat) - 1)
        rot = get_rot_mat(quat)[inds]  # (cells, 3, 3)
        Fp_inv = soa_empty(nc, (3, 3))
        Fp_inv.copy_(torch.eye(3, dtype=torch.float64, device=device()).expand(nc, 8, 3, 3))
        R = soa_empty(nc, (3, 3))
        R.copy_(torch.as_tensor(rot, device=device()).unsqueeze(1).expand(nc, 8, 3, 3))
        if phase_inds is not None:
            pm = [phase_materials[int(k)] for k in range(len(phase_materials))]
            ph = np.asarray(phase_inds, dtype=np.int64).reshape(-1)
            per_cell = lambda f: torch.as_tensor(np.array([getattr(pm[k], f) for k in ph], dtype=np.float64),
                                                 device=device()).unsqueeze(1).expand(nc, 8).contiguous()
            g0 = per_cell("g0")
        else:
            g0 = torch.full((nc, 8), float(mat.g0), dtype=torch.float64, device=device())
        g = soa_empty(nc, (Ns,))
        g.copy_(g0.unsqueeze(2).expand(nc, 8, Ns))
        slip = soa_empty(nc, (Ns,))
        slip.zero_()
        iv = [Fp_inv, g, slip, R]
        layout = "10" if phase_inds is not None else mat.layout
        if layout in ("9", "10"):
            for f in ("gss_a", "h", "t_sat", "xm", "r"):
                if phase_inds is not None:
                    iv.append(per_cell(f))
                else:
                    iv.append(torch.full((nc, 8), float(getattr(mat, f)), dtype=torch.float64, device=device()))
        if layout == "10":
            Cs = torch.as_tensor(np.stack([M.cubic_stiffness(pm[k].C11, pm[k].C12, pm[k].C44) for k in ph]),
                                 device=device())
            iv.append(Cs.unsqueeze(1).expand(nc, 8, 3, 3, 3, 3).contiguous())
        self.layout = layout
        self.internal_vars = iv
        self._mat_scalars = torch.tensor([mat.xm, mat.gss_a, mat.h, mat.t_sat, mat.r, mat.C11, mat.C12, mat.C44],
                                         dtype=torch.float64, device=device())

    # ------------------------------------------------------------------ state marshalling
    def _marshal(self, params, n_cells=None, n_qp=None):
        """CpfemState + CpfemParams for a parameter list (no copies when already SoA).
        While ``_override`` is set (whole-history adjoint driver) it is used instead."""
        if getattr(self, "_override", None) is not None and n_cells is None:
            return self._override
        nc = self.fes[0].num_cells if n_cells is None else n_cells
        Ns = self.num_slip_sys
        keep = []

        def sp(t, comps):
            p, k = soa_ptr(t, nc, comps)
            keep.append(k)
            return p

        st = N.CpfemState(sp(params[0], (3, 3)), sp(params[1], (Ns,)), sp(params[2], (Ns,)), sp(params[3], (3, 3)))
        pr = N.CpfemParams()
        base = self._mat_scalars.data_ptr()
        for k in range(N.NPARAM):
            pr.ptr[k] = base + 8 * k
            pr.stride[k] = 0
        if len(params) >= 9:
            # 9-array order: gss_a, h, t_sat, xm, r  ->  ABI order XM, GSS_A, H, TSAT, R
            for abi, idx in ((N.GSS_A, 4), (N.H, 5), (N.TSAT, 6), (N.XM, 7), (N.R, 8)):
                pr.ptr[abi] = sp(params[idx], ()).value
                pr.stride[abi] = 1
        if len(params) >= 10:
            C = to_device(params[9]).reshape(nc * 8, 3, 3, 3, 3)
            for abi, ix in ((N.C11, (0, 0, 0, 0)), (N.C12, (0, 0, 1, 1)), (N.C44, (1, 2, 1, 2))):
                v = C[(slice(None),) + ix].contiguous()
                keep.append(v)
                pr.ptr[abi] = v.data_ptr()
                pr.stride[abi] = 1
        mat = self.material
        pr.gamma0 = mat.ao
        pr.dt = float(self.dt)
        pr.tol = mat.tol
        pr.max_sub_step = mat.max_sub_step
        pr.max_iter = mat.max_iter
        pr.self_block = 1 if mat.self_block else 0
        return st, pr, keep

    # ------------------------------------------------------------------ reference surface
    def get_tensor_map(self):
        tensor_map, _ = self.get_maps()
        return tensor_map

    def get_maps(self):
        """(tensor_map, update_int_vars_map) over batches of points (models_copper.py:139-271).

        tensor_map(u_grad, *iv) -> P and update_int_vars_map(u_grad, *iv) -> new iv take
        u_grad of shape (..., 3, 3) and internal variables with the same leading shape."""
        def _run(u_grad, iv, want_P, want_update):
            ug = to_device(u_grad)
            lead = ug.shape[:-2]
            Q = int(np.prod(lead)) if len(lead) else 1
            if Q % 8:
                raise ValueError("point batches must hold a multiple of 8 points")
            ncb = Q // 8
            ivr = [to_device(v).reshape((Q,) + tuple(to_device(v).shape[len(lead):])) for v in iv]
            ivr = [v.reshape((ncb, 8) + tuple(v.shape[1:])) for v in ivr]
            st, pr, keep = self._marshal(ivr, n_cells=ncb)
            H = ug.reshape(Q, 9).t().contiguous()
            P = torch.empty(9, Q, dtype=torch.float64, device=device()) if want_P else None
            out = None
            if want_update:
                Fn, gn, sn = (torch.empty(9, Q, dtype=torch.float64, device=device()),
                              torch.empty(self.num_slip_sys, Q, dtype=torch.float64, device=device()),
                              torch.empty(self.num_slip_sys, Q, dtype=torch.float64, device=device()))
                out = N.CpfemStateOut(ptr(Fn), ptr(gn), ptr(sn))
            N.check(self.ctx._lib.cpfem_local(self.ctx.h, stream_handle(), Q, ptr(H), None, ctypes.byref(st),
                                              ctypes.byref(pr), ptr(P), None, None,
                                              ctypes.byref(out) if out is not None else None, None))
            if want_P:
                return P.t().reshape(tuple(lead) + (3, 3))
            Ns = self.num_slip_sys
            res = [Fn.t().reshape(tuple(lead) + (3, 3)), gn.t().reshape(tuple(lead) + (Ns,)),
                   sn.t().reshape(tuple(lead) + (Ns,))] + [to_device(v) for v in iv[3:]]
            return res

        def tensor_map(u_grad, *iv):
            return _run(u_grad, iv, True, False)

        def update_int_vars_map(u_grad, *iv):
            return _run(u_grad, iv, False, True)

        return tensor_map, update_int_vars_map

    def set_params(self, params):
        self.internal_vars = params

    def set_initial_guess(self, initial_sol):
        self.initial_guess = initial_sol

    def update_int_vars_gp(self, sol, params):
        """New history at the converged solution (models_copper.py:273-282); differentiable
        under torch autograd when sol or params require grad."""
        from . import adjoint
        if adjoint.differentiable(params, sol):
            new = adjoint._UpdateFn.apply(self, sol, *params)
            return list(new) + list(params[3:])
        return self._update_raw(sol, params)

    def _update_raw(self, sol, params):
        nc, Ns = self.fes[0].num_cells, self.num_slip_sys
        dofs = self.flatten(sol)
        st, pr, keep = self._marshal(params)
        Fn, gn, sn = soa_empty(nc, (3, 3)), soa_empty(nc, (Ns,)), soa_empty(nc, (Ns,))
        out = N.CpfemStateOut(Fn.data_ptr(), gn.data_ptr(), sn.data_ptr())
        N.check(self.ctx._lib.cpfem_update_state(self.ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st),
                                                 ctypes.byref(pr), ctypes.byref(out)))
        return [Fn, gn, sn] + [p if isinstance(p, torch.Tensor) else to_device(p) for p in params[3:]]

    def compute_avg_stress(self, sol, params):
        """JxW-weighted cell average of the Cauchy stress, (cells, 3, 3) (models_copper.py:297-319);
        differentiable under torch autograd when sol or params require grad."""
        from . import adjoint
        if adjoint.differentiable(params, sol):
            return adjoint._AvgStressFn.apply(self, sol, *params)
        return self._compute_avg_stress_raw(sol, params)

    def _compute_avg_stress_raw(self, sol, params):
        dofs = self.flatten(sol)
        st, pr, keep = self._marshal(params)
        sig = torch.empty(self.fes[0].num_cells, 3, 3, dtype=torch.float64, device=device())
        N.check(self.ctx._lib.cpfem_cell_stress(self.ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st),
                                                ctypes.byref(pr), ptr(sig)))
        return sig

    def inspect_interval_vars(self, params):
        """Fp, g and slip of the first point (models_copper.py:287-295)."""
        Fp = np.linalg.inv(to_device(params[0])[0, 0].cpu().numpy())
        g0 = float(to_device(params[1])[0, 0, 0])
        print(f"Fp = \n{Fp}")
        print(f"slip_resistance index 0 = {g0}, max slip_resistance = {float(to_device(params[1]).max())}")
        return Fp[2, 2], g0, float(to_device(params[2])[0, 0, 0])

    # ------------------------------------------------------------------ assembly (solver side)
    def assemble(self, dofs, want_matrix=True, apply_bc=True, params=None, iters=None):
        """Residual (and block-ELL Jacobian values) at dofs; newton_update + apply_bc_vec + get_A."""
        params = self.internal_vars if params is None else params
        st, pr, keep = self._marshal(params)
        dofs = dofs.detach()
        res = torch.empty(self.ctx.n_dofs, dtype=torch.float64, device=device())
        mat = None
        if want_matrix:
            if getattr(self, "_mat_buf", None) is None:
                self._mat_buf = torch.empty(self.ctx.mat_len, dtype=torch.float64, device=device())
            mat = self._mat_buf
        N.check(self.ctx._lib.cpfem_assemble(self.ctx.h, stream_handle(), ptr(dofs), ctypes.byref(st),
                                             ctypes.byref(pr), ptr(res), ptr(mat), ptr(iters),
                                             N.APPLY_BC if apply_bc else 0))
        return res, mat

    def newton_update(self, sol_list):
        """Raw residual list; the Jacobian is kept as ``self.A`` (jax-fem newton_update)."""
        dofs = self.flatten(sol_list)
        res, mat = self.assemble(dofs, want_matrix=True, apply_bc=False)
        from .problem import BlockEllMatrix
        self.A = BlockEllMatrix(self.ctx, mat.clone())
        return self.unflatten_fn_sol_list(res)

    def compute_residual(self, sol_list):
        dofs = self.flatten(sol_list)
        res, _ = self.assemble(dofs, want_matrix=False, apply_bc=False)
        return self.unflatten_fn_sol_list(res)


def local_full(problem, u_grad, params):
    """Every output of the point kernel for a batch (P, dP/du_grad, new history, iterations)
    from explicit u_grad (Q, 3, 3) and per-point params (each (Q, ...)); for tests and
    diagnostics.  Q must be a multiple of 8."""
    ug = to_device(u_grad).reshape(-1, 3, 3)
    Q = ug.shape[0]
    ncb = Q // 8
    ivr = [to_device(v).reshape((ncb, 8) + tuple(to_device(v).shape[1:])) for v in params]
    st, pr, keep = problem._marshal(ivr, n_cells=ncb)
    Ns = problem.num_slip_sys
    H = ug.reshape(Q, 9).t().contiguous()
    P = torch.empty(9, Q, dtype=torch.float64, device=device())
    A = torch.empty(81, Q, dtype=torch.float64, device=device())
    sig = torch.empty(9, Q, dtype=torch.float64, device=device())
    Fn = torch.empty(9, Q, dtype=torch.float64, device=device())
    gn = torch.empty(Ns, Q, dtype=torch.float64, device=device())
    sn = torch.empty(Ns, Q, dtype=torch.float64, device=device())
    it = torch.empty(Q, dtype=torch.int32, device=device())
    out = N.CpfemStateOut(ptr(Fn), ptr(gn), ptr(sn))
    N.check(problem.ctx._lib.cpfem_local(problem.ctx.h, stream_handle(), Q, ptr(H), None, ctypes.byref(st),
                                         ctypes.byref(pr), ptr(P), ptr(A), ptr(sig), ctypes.byref(out), ptr(it)))
    torch.cuda.synchronize()
    return dict(P=P.t().reshape(Q, 3, 3).cpu().numpy(), A=A.t().reshape(Q, 9, 9).cpu().numpy(),
                sigma=sig.t().reshape(Q, 3, 3).cpu().numpy(), Fp_inv_new=Fn.t().reshape(Q, 3, 3).cpu().numpy(),
                g_new=gn.t().cpu().numpy(), slip_new=sn.t().cpu().numpy(), iters=it.cpu().numpy())
