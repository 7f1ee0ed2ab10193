"""Problem / FiniteElement surface of the FE layer the reference builds on, backed by the HIP library.

Mirrors the parts of jax-fem's ``Problem`` and ``FiniteElement`` that the reference model and
solver use (models_copper.py:51-319, solver.py:119-437): ``fes[0].{cells, points,
num_quads, num_total_nodes, vec, dim, shape_grads, JxW, node_inds_list, vec_inds_list,
vals_list, update_Dirichlet_boundary_conditions}``, ``num_total_dofs_all_vars``,
``offset``, ``unflatten_fn_sol_list``, ``newton_update`` and ``compute_residual``.
Arrays that live on the device are torch float64 tensors on the current HIP device; the
DLPack bridge (``torch.from_dlpack``) accepts JAX/CuPy arrays, NumPy arrays are copied.
"""
import ctypes

import numpy as np
import torch

from . import _native as N


def device():
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_device(x, dtype=torch.float64):
    """Device tensor view of x (torch / DLPack producer / NumPy)."""
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, "__dlpack__") and not isinstance(x, np.ndarray):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(np.asarray(x))
    return t.to(device=device(), dtype=dtype)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class NativeContext:
    """Owns one cpfem_ctx (mesh geometry, block-ELL pattern, work buffers)."""

    def __init__(self, points, cells, directions, normals):
        L = N.lib()
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        d = np.ascontiguousarray(directions, dtype=np.float64)
        n = np.ascontiguousarray(normals, dtype=np.float64)
        mesh = N.CpfemMesh(len(self.points), len(self.cells), self.points.ctypes.data, self.cells.ctypes.data)
        slip = N.CpfemSlip(len(d), d.ctypes.data, n.ctypes.data)
        h = ctypes.c_void_p()
        N.check(L.cpfem_create(ctypes.byref(mesh), ctypes.byref(slip), ctypes.byref(h)))
        self.h = h
        self._lib = L
        info = (ctypes.c_int64 * 8)()
        N.check(L.cpfem_info(h, info))
        (self.n_nodes, self.n_cells, self.n_qp, self.n_dofs, self.n_slots, self.uniform, self.mat_len,
         self.n_slip) = [int(v) for v in info]

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._lib.cpfem_destroy(h)
            self.h = None

    def geometry(self):
        sg = np.empty((self.n_cells, 8, 8, 3))
        jxw = np.empty((self.n_cells, 8))
        N.check(self._lib.cpfem_geometry(self.h, sg.ctypes.data, jxw.ctypes.data))
        return sg, jxw

    def set_point_park(self, lanes):
        """Parking of the assembly point kernel's slow lanes (0 = none; cpfem_set_point_park):
        changes the time only, never a result."""
        N.check(self._lib.cpfem_set_point_park(self.h, int(lanes)))

    def set_spmv_form(self, form):
        """Forces the BiCGSTAB SpMV form (-1: by system size; cpfem_set_spmv_form): changes the
        time, and the rounding of the dot products, never the algorithm."""
        N.check(self._lib.cpfem_set_spmv_form(self.h, int(form)))

    def set_box_columns(self, enable):
        """Computed node columns of the one-row-per-thread SpMV on a box-numbered mesh on (the
        default) or off (cpfem_set_box_columns): changes the time, never a result bit."""
        N.check(self._lib.cpfem_set_box_columns(self.h, 1 if enable else 0))

    def krylov_config(self):
        """(SpMV form, SpMV blocks, vector-kernel blocks) of this context's BiCGSTAB."""
        out = (ctypes.c_int32 * 3)()
        N.check(self._lib.cpfem_krylov_config(self.h, out))
        return tuple(int(v) for v in out)

    def point_counts(self, reset=True):
        """(local Newton iterations, residual evaluations, parked points) of the assembly's
        point kernel since the last reset."""
        out = (ctypes.c_int64 * 3)()
        N.check(self._lib.cpfem_point_counts(self.h, stream_handle(), 1 if reset else 0, out))
        return tuple(int(v) for v in out)

    def pattern(self):
        nslot = np.empty(self.n_nodes, dtype=np.int32)
        col = np.empty((self.n_slots, self.n_nodes), dtype=np.int32)
        N.check(self._lib.cpfem_pattern(self.h, nslot.ctypes.data, col.ctypes.data))
        return nslot, col

    def box_numbering(self):
        """(nx, ny, nz) nodes if the mesh is a box numbered x-fastest whose nodes off the faces
        have the 27 stencil columns (cpfem_box_numbering), else (0, 0, 0)."""
        nslot, col = self.pattern()
        dims = np.zeros(3, dtype=np.int64)
        N.check(self._lib.cpfem_box_numbering(self.n_nodes, col.ctypes.data, nslot.ctypes.data, dims.ctypes.data))
        return tuple(int(d) for d in dims)


class BlockEllMatrix:
    """Global Jacobian in block-ELL form (3x3 node blocks), values on the device.

    ``to_scipy()`` exports the CSR of scipy/PETSc's AIJ with sorted column indices.
    """

    def __init__(self, ctx, values):
        self.ctx = ctx
        self.values = values
        self.shape = (ctx.n_dofs, ctx.n_dofs)

    def matvec(self, x):
        x = to_device(x).contiguous()
        y = torch.empty_like(x)
        N.check(self.ctx._lib.cpfem_spmv(self.ctx.h, stream_handle(), ptr(self.values), ptr(x), ptr(y)))
        return y

    __matmul__ = matvec

    def to_scipy(self):
        import scipy.sparse as sp
        nslot, _ = self.ctx.pattern()
        nnz = int(3 * np.repeat(nslot, 3).sum())
        indptr = np.empty(self.ctx.n_dofs + 1, dtype=np.int64)
        indices = np.empty(nnz, dtype=np.int32)
        data = np.empty(nnz)
        N.check(self.ctx._lib.cpfem_to_csr(self.ctx.h, stream_handle(), ptr(self.values), indptr.ctypes.data,
                                           indices.ctypes.data, data.ctypes.data))
        return sp.csr_matrix((data, indices, indptr), shape=self.shape)


def _eval_location(fn, points):
    """A reference location / value function at every mesh point, with the per-point
    semantics of jax-fem's ``vmap(fn)(points)``.  The function is first called once on the
    coordinate rows (points.T): componentwise functions return one value per point there.
    Anything else (an exception, or a 0-d result: a constant, or a function that reduces over
    its point such as ``np.all(np.isclose(p, c))``) is evaluated point by point."""
    try:
        v = np.asarray(fn(points.T))
        if v.shape == (len(points),):
            return v
    except Exception:
        pass
    return np.array([np.asarray(fn(p)).item() for p in points])


class FiniteElement:
    """fes[0] of the reference Problem: HEX8 mesh, 2x2x2 quadrature, Dirichlet lists."""

    def __init__(self, mesh, vec, dim, ele_type, dirichlet_bc_info):
        if ele_type.upper() != "HEX8":
            raise NotImplementedError("only HEX8 elements are supported")
        if vec != 3 or dim != 3:
            raise NotImplementedError("crystal plasticity needs vec = dim = 3")
        self.points = np.asarray(mesh.points, dtype=np.float64)
        self.cells = np.asarray(mesh.cells, dtype=np.int32)
        self.vec = vec
        self.dim = dim
        self.ele_type = "HEX8"
        self.num_cells = len(self.cells)
        self.num_quads = 8
        self.num_nodes = 8
        self.num_total_nodes = len(self.points)
        self.num_total_dofs = self.vec * self.num_total_nodes
        self._geom = None
        self.ctx = None
        self.node_inds_list, self.vec_inds_list, self.vals_list = [], [], []
        self.dirichlet_bc_info = dirichlet_bc_info
        if dirichlet_bc_info is not None:
            self.update_Dirichlet_boundary_conditions(dirichlet_bc_info)

    def update_Dirichlet_boundary_conditions(self, dirichlet_bc_info):
        """location_fns, vecs, value_fns -> node/vec/value lists (jax-fem semantics,
        refreshed every load step at singlecrystal_copper.py:186)."""
        self.dirichlet_bc_info = dirichlet_bc_info
        location_fns, vecs, value_fns = dirichlet_bc_info
        self.node_inds_list, self.vec_inds_list, self.vals_list = [], [], []
        for loc, v, val in zip(location_fns, vecs, value_fns):
            nodes = np.argwhere(_eval_location(loc, self.points).astype(bool)).reshape(-1)
            self.node_inds_list.append(nodes)
            self.vec_inds_list.append(np.full(len(nodes), int(v), dtype=np.int64))
            self.vals_list.append(_eval_location(val, self.points[nodes]).astype(np.float64).reshape(-1)
                                  if len(nodes) else np.zeros(0))
        if self.ctx is not None:
            self.push_dirichlet()

    def dirichlet_dofs(self):
        """Unique dofs and values; a later list overrides an earlier one on a shared dof
        (the .set/.add order of apply_bc_vec and assign_bc, solver.py:119-153)."""
        dof_parts = [3 * n.astype(np.int64) + c for n, c in zip(self.node_inds_list, self.vec_inds_list)]
        if not dof_parts:
            return np.zeros(0, dtype=np.int64), np.zeros(0)
        dofs = np.concatenate(dof_parts)
        vals = np.concatenate(self.vals_list)
        # keep the last occurrence of every dof
        rev = dofs[::-1]
        u, first_rev = np.unique(rev, return_index=True)
        return u.astype(np.int64), vals[::-1][first_rev]

    def push_dirichlet(self):
        dofs, vals = self.dirichlet_dofs()
        dofs = np.ascontiguousarray(dofs)
        vals = np.ascontiguousarray(vals)
        N.check(self.ctx._lib.cpfem_set_dirichlet(self.ctx.h, stream_handle(), len(dofs), dofs.ctypes.data,
                                                  vals.ctypes.data))

    def _geometry(self):
        if self._geom is None:
            self._geom = self.ctx.geometry()
        return self._geom

    @property
    def shape_grads(self):
        """(cells, 8 quads, 8 nodes, 3) physical shape-function gradients."""
        return self._geometry()[0]

    @property
    def JxW(self):
        """(cells, 8 quads) Jacobian determinant times Gauss weight."""
        return self._geometry()[1]


class Problem:
    """Single-variable FE problem; subclasses implement ``custom_init`` and the state layout."""

    def __init__(self, mesh, vec=3, dim=3, ele_type="HEX8", gauss_order=None, dirichlet_bc_info=None,
                 location_fns=None, additional_info=()):
        self.mesh = mesh
        self.vec = vec
        self.dim = dim
        self.fes = [FiniteElement(mesh, vec, dim, ele_type, dirichlet_bc_info)]
        self.num_vars = 1
        self.offset = [0]
        self.num_total_dofs_all_vars = self.fes[0].num_total_dofs
        self.num_cells = self.fes[0].num_cells
        self.dt = 1.0
        self.custom_init(*additional_info)

    # --- hooks
    def custom_init(self, *args):
        pass

    def _attach(self, directions, normals):
        fe = self.fes[0]
        if fe.ctx is None:
            fe.ctx = NativeContext(fe.points, fe.cells, directions, normals)
            fe.push_dirichlet()
        self.ctx = fe.ctx

    # --- dof helpers
    def unflatten_fn_sol_list(self, dofs):
        return [dofs.reshape(self.fes[0].num_total_nodes, self.vec)]

    @staticmethod
    def flatten(sol_list):
        if isinstance(sol_list, (list, tuple)):
            sol_list = sol_list[0]
        return to_device(sol_list).detach().reshape(-1).contiguous()
