"""The five benchmark workloads as concrete synthetic inputs (units MPa, mm, s).

Each builder returns a Workload: mesh, material, orientations (scalar-first quaternions) and
per-cell orientation indices, the symmetry boundary conditions of
polycrystal_DPsteel_inhomo.py:180-182 (x=0: u_x=0, y=0: u_y=0, z=0: u_z=0, z=L: u_z=eps L),
and the displacement / time histories (strain rate 1e-3 /s, so dt = d_eps / 1e-3).
Randomness: numpy default_rng(seed) (PCG64); Voronoi seeds are drawn first, then the grain
quaternions (Shoemake), then any phase permutation.
"""
import dataclasses

import numpy as np

from . import materials as M
from .generate_mesh import Mesh, box_mesh, voronoi_grains
from .rotations import uniform_quaternions


@dataclasses.dataclass
class Workload:
    name: str
    mesh: Mesh
    material: M.Material
    quat: np.ndarray
    cell_ori_inds: np.ndarray
    disps: np.ndarray
    ts: np.ndarray
    L: float = 1.0
    phase_inds: np.ndarray = None
    phase_materials: tuple = None
    precision: str = "fp64"
    solver_options: dict = dataclasses.field(default_factory=dict)  # the reference driver's options

    @property
    def n_qp(self):
        return 8 * len(self.mesh.cells)

    def dirichlet_bc_info(self, disp):
        L = self.L
        left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
        front = lambda p: np.isclose(p[1], 0.0, atol=1e-5)
        bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
        top = lambda p: np.isclose(p[2], L, atol=1e-5)
        # constant values written componentwise (one value per point for a coordinate-row
        # array as for a single point), so that the 64^3 boundary lists need no per-point calls
        zero = lambda p: np.zeros_like(p[0], dtype=np.float64)
        const = lambda d: (lambda p: np.full_like(p[0], d, dtype=np.float64))
        return [[left, front, bottom, top], [0, 1, 2, 2], [zero, zero, zero, const(disp)]]


def _mesh(n, L=1.0):
    m = box_mesh(n, n, n, L, L, L)
    return Mesh(m.points, m.cells_dict["hexahedron"])


def _grains(mesh, n_grains, rng, L=1.0):
    cent = mesh.points[mesh.cells].mean(axis=1)
    ids, _ = voronoi_grains(cent, n_grains, (L, L, L), rng)
    quat = uniform_quaternions(n_grains, rng)
    return quat, ids


def _history(total_strain, steps, L=1.0, rate=1e-3):
    disps = np.linspace(0.0, total_strain * L, steps + 1)
    ts = np.linspace(0.0, total_strain / rate, steps + 1)
    return disps, ts


def config1(seed=0, steps=10):
    """Single-crystal FCC copper, 8^3 hex8 (4,096 points), 1% tension in `steps` steps.

    With these constants (m = 0.02, g0 = 16 MPa) the reference algorithm's local Newton from
    S = 0 overflows within the first load step at 10 steps (1e-3 strain each), in the oracle as
    on the device; 40 steps (2.5e-4 each) run cleanly."""
    rng = np.random.default_rng(seed)
    mesh = _mesh(8)
    quat = uniform_quaternions(1, rng)
    mat = M.Material(slip="fcc12", g0=16.0, h=180.0, t_sat=148.0, gss_a=2.25, xm=0.02, r=1.4,
                     C11=168.4e3, C12=121.4e3, C44=75.4e3)
    d, t = _history(0.01, steps)
    return Workload(f"C1 Cu single crystal 8^3 FP64, {steps} steps", mesh, mat, quat, np.zeros(len(mesh.cells), np.int64), d, t)


def config2(seed=0, n=32):
    """Polycrystal FCC 304-steel-like, 32^3 hex8 (262,144 points), 100 Voronoi grains,
    2% tension in 40 steps."""
    rng = np.random.default_rng(seed)
    mesh = _mesh(n)
    quat, ids = _grains(mesh, 100, rng)
    mat = M.Material(slip="fcc12", g0=88.0, h=500.0, t_sat=300.0, gss_a=2.0, xm=0.02, r=1.4,
                     C11=204.6e3, C12=137.7e3, C44=126.2e3)
    d, t = _history(0.02, 40)
    return Workload(f"C2 FCC polycrystal {n}^3, 100 grains, FP64", mesh, mat, quat, ids, d, t)


def config3(seed=0, n=40):
    """Polycrystal BCC tantalum-like, 40^3 (512,000 points), 64 grains, BCC24, 3% in 30 steps;
    mixed precision against plain FP64."""
    rng = np.random.default_rng(seed)
    mesh = _mesh(n)
    quat, ids = _grains(mesh, 64, rng)
    mat = M.Material(slip="bcc24", g0=150.0, h=800.0, t_sat=400.0, gss_a=2.0, xm=0.02, r=1.4,
                     C11=266.0e3, C12=161.0e3, C44=82.5e3)
    d, t = _history(0.03, 30)
    return Workload(f"C3 BCC polycrystal {n}^3, 64 grains", mesh, mat, quat, ids, d, t, precision="mixed")


def config4(seed=0, n=24):
    """Adjoint calibration workload: 24^3, 64 grains, FCC with 304-like stiffness; hidden
    parameters (90, 450, 280, 2.2, 0.02), 5% in 50 steps."""
    rng = np.random.default_rng(seed)
    mesh = _mesh(n)
    quat, ids = _grains(mesh, 64, rng)
    mat = M.Material(slip="fcc12", g0=90.0, h=450.0, t_sat=280.0, gss_a=2.2, xm=0.02, r=1.4,
                     C11=204.6e3, C12=137.7e3, C44=126.2e3, layout="9")
    d, t = _history(0.05, 50)
    # solver options of the steel calibration driver (calibration_case4_UQ_polyCrystalSteel_1D_GB.py:77)
    return Workload(f"C4 calibration {n}^3, 64 grains", mesh, mat, quat, ids, d, t,
                    solver_options={"tol": 1e-7, "line_search_flag": True})


def config5(seed=0, n=64):
    """Dual-phase BCC24 orientation-design workload: 64^3 (2,097,152 points), 300 grains,
    20% of grains hard (g0 x4, h0 x2), 2% in 20 steps."""
    rng = np.random.default_rng(seed)
    mesh = _mesh(n)
    quat, ids = _grains(mesh, 300, rng)
    soft = M.Material(slip="bcc24", g0=100.0, h=400.0, t_sat=250.0, gss_a=2.0, xm=0.02, r=1.4,
                      C11=231.0e3, C12=135.0e3, C44=116.0e3, layout="10")
    hard = soft.replace(g0=400.0, h=800.0)
    hard_grains = rng.permutation(300)[:60]
    grain_phase = np.zeros(300, dtype=np.int64)
    grain_phase[hard_grains] = 1
    d, t = _history(0.02, 20)
    # solver options of the dual-phase driver (polycrystal_DPsteel_inhomo.py:227)
    return Workload(f"C5 dual-phase BCC {n}^3, 300 grains", mesh, soft, quat, ids, d, t,
                    phase_inds=grain_phase[ids], phase_materials=(soft, hard), precision="mixed",
                    solver_options={"line_search_flag": True})


CONFIGS = {1: config1, 2: config2, 3: config3, 4: config4, 5: config5}


# ----------------------------------------------------------------------------- history inputs
# Inputs of cpfem_amd.simulate (the differentiable whole-history form) for a workload:
# per-point initial slip resistance g0 (cells, 8, N), the eight per-point material
# parameters mat8 (cells, 8, 8: xm, gss_a, h, t_sat, r, C11, C12, C44) and rotations.

def _per_point_materials(w):
    nc = len(w.mesh.cells)
    if w.phase_inds is None:
        return [w.material] * 1, np.zeros(nc, dtype=np.int64)
    return list(w.phase_materials), np.asarray(w.phase_inds, dtype=np.int64)


def history_inputs(w, prob=None, device=None):
    """(g0, mat8) float64 tensors of a workload, per phase where it has phases
    (models_DPsteel_inhomo.py:55-186), on `device` (default: the current HIP device)."""
    import torch
    nc = len(w.mesh.cells)
    Ns = len(M.load_slip_table(w.material.slip))
    mats, ph = _per_point_materials(w)
    table = np.array([[m.xm, m.gss_a, m.h, m.t_sat, m.r, m.C11, m.C12, m.C44] for m in mats], dtype=np.float64)
    g0s = np.array([m.g0 for m in mats], dtype=np.float64)
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    mat8 = torch.as_tensor(table[ph], device=dev)[:, None, :].expand(nc, 8, 8).contiguous()
    g0 = torch.as_tensor(g0s[ph], device=dev)[:, None, None].expand(nc, 8, Ns).contiguous()
    return g0, mat8


def history_rotations(w):
    """(cells, 8, 3, 3) device tensor of the workload's crystal rotations (JAX-style clamped
    orientation indices, as custom_init)."""
    import torch
    from .rotations import get_rot_mat
    nc = len(w.mesh.cells)
    inds = np.clip(np.asarray(w.cell_ori_inds, dtype=np.int64), 0, len(w.quat) - 1)
    R = torch.as_tensor(get_rot_mat(w.quat)[inds], device=torch.device("cuda", torch.cuda.current_device()))
    return R[:, None].expand(nc, 8, 3, 3)


def calibration_inputs(w, prob, theta):
    """(g0, mat8) of the calibration workload as functions of the five scale factors theta of
    (g0, h0, g_sat, a, m) (differentiable in theta, on theta's device)."""
    import torch
    m = w.material
    nc = len(w.mesh.cells)
    Ns = len(M.load_slip_table(m.slip))
    base = torch.tensor([m.xm, m.gss_a, m.h, m.t_sat, m.r, m.C11, m.C12, m.C44], dtype=torch.float64,
                        device=theta.device)
    one = torch.ones((), dtype=torch.float64, device=theta.device)
    sc = torch.stack([theta[4], theta[3], theta[1], theta[2], one, one, one, one])
    return (m.g0 * theta[0]).expand(nc, 8, Ns), (base * sc).expand(nc, 8, 8)


def calibration_query(seed=0):
    """Query scale factors of the calibration workload: hidden x (1 + U(-0.2, 0.2)), seeded."""
    return 1.0 + np.random.default_rng(seed).uniform(-0.2, 0.2, size=5)


def build_problem(w):
    """CrystalPlasticity for a workload (BCs at the first displacement)."""
    from .models import CrystalPlasticity
    info = w.dirichlet_bc_info(w.disps[0])
    extra = (w.phase_inds, w.phase_materials) if w.phase_inds is not None else ()
    prob = CrystalPlasticity(w.mesh, vec=3, dim=3, ele_type="HEX8", dirichlet_bc_info=info,
                             additional_info=(w.quat, w.cell_ori_inds) + extra, material=w.material)
    prob._bc_info = info
    return prob


def run_steps(prob, w, params, sol, i0, i1, solver_options=None, on_step=None):
    """The reference driver loop (singlecrystal_copper.py:179-233) over steps [i0, i1).
    Returns (params, sol, sigma_bar list (3x3 numpy per step))."""
    import torch
    from .solver import solver
    opts = dict(w.solver_options)
    opts.update(solver_options or {})
    out = []
    info = prob._bc_info
    for i in range(i0, i1):
        prob.dt = w.ts[i + 1] - w.ts[i]
        info[-1][-1] = w.dirichlet_bc_info(w.disps[i + 1])[-1][-1]
        prob.fes[0].update_Dirichlet_boundary_conditions(info)
        prob.set_params(params)
        o = {"initial_guess": sol}
        o.update(opts)
        if not (o.keys() & {"jax_solver", "umfpack_solver", "petsc_solver", "custom_solver"}):
            o["jax_solver"] = {}
        sol = solver(prob, o)
        sig = prob.compute_avg_stress(sol[0], params)
        out.append(sig.mean(dim=0).cpu().numpy())
        params = prob.update_int_vars_gp(sol[0], params)
        if on_step is not None:
            on_step(i, prob, out[-1])
    torch.cuda.synchronize()
    return params, sol, out
