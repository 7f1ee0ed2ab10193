"""Row-elimination Newton solver with the reference's ``solver(problem, solver_options)`` call.

Follows crystal_plasticity_OR_design/solver.py:310-437 (and linear_incremental_solver
:213-237, linear_solver :91-113, line_search :240-276).  The residual, the Jacobian with the
Dirichlet row elimination and the Krylov solve all run in libcpfem_hip.so; this module only
sequences the calls and reads back norms.

solver_options keys (as in the reference): 'jax_solver': {'precond': bool}, 'umfpack_solver',
'petsc_solver', 'custom_solver', 'initial_guess', 'tol' (1e-6), 'rel_tol' (1e-8),
'line_search_flag' (False).  Added: 'precision': 'fp64' (default) | 'mixed' (FP32 Krylov matrix
with FP64 refinement), 'krylov_tol'/'krylov_atol' (1e-10), 'krylov_maxiter' (10000),
'max_newton' (200; the reference loop has no cap and spins forever on a stagnating residual).
'umfpack_solver' is served by the device banded direct solver (cpfem_solve_direct: unpivoted
band LU under a bandwidth-reducing node order plus FP64 iterative refinement, 'refine' steps,
default 2); 'petsc_solver' (an iterative Krylov solve in the reference, bcgsl + ilu by default)
by the device Krylov solver at the same 1e-10 tolerance.
"""
import ctypes
import logging
import math
import time

import torch

from . import _native as N
from .problem import BlockEllMatrix, device, ptr, stream_handle

logger = logging.getLogger("cpfem_amd")


def norm(problem, v, status_out=None, allow_nonfinite=False):
    """|v|_2 of a residual, with the point-status check of the kernels queued before it.

    A point that hit the local iteration cap raises (CpfemError, code ELOCAL): the
    reference's local Newton has no cap and would never return there.  A non-finite residual
    or point state raises (code ENONFINITE) unless ``allow_nonfinite``, in which case NaN is
    returned, as np.linalg.norm does in the reference's Newton loop and line search
    (solver.py:247-249,408-416)."""
    out = ctypes.c_double()
    st = ctypes.c_int32()
    rc = problem.ctx._lib.cpfem_norm(problem.ctx.h, stream_handle(), ptr(v), ctypes.byref(out), ctypes.byref(st))
    if status_out is not None:
        status_out.append(int(st.value))
    if rc == N.ENONFINITE and allow_nonfinite:
        logger.debug("non-finite residual: %s", problem.ctx._lib.cpfem_last_error().decode(errors="replace"))
        return math.nan
    N.check(rc)
    return out.value


def _precision(solver_options):
    p = solver_options.get("precision", "fp64")
    if p not in ("fp64", "mixed"):
        raise ValueError(f"unknown precision {p!r}")
    return N.MIXED if p == "mixed" else N.FP64


def krylov_solve(problem, mat, b, x0, solver_options=None, precond=True):
    """jax_solve (solver.py:19-48) on the device: Jacobi-BiCGSTAB, tol = atol = 1e-10,
    maxiter 10000, then the |Ax - b| < 0.1 assertion."""
    solver_options = solver_options or {}
    x = x0.clone()
    info = N.CpfemSolveInfo()
    rc = problem.ctx._lib.cpfem_solve(problem.ctx.h, stream_handle(), ptr(mat), ptr(b), ptr(x),
                                      float(solver_options.get("krylov_tol", 1e-10)),
                                      float(solver_options.get("krylov_atol", 1e-10)),
                                      int(solver_options.get("krylov_maxiter", 10000)), 1 if precond else 0,
                                      _precision(solver_options), ctypes.byref(info))
    stats = dict(iterations=info.iterations, outer=info.outer, residual=info.residual, b_norm=info.b_norm)
    problem.solver_info.append(stats)
    if rc == N.EKRYLOV:
        raise AssertionError(f"JAX linear solver failed to converge with err = {info.residual}")
    N.check(rc)
    return x


def direct_solve(problem, mat, b, solver_options=None):
    """umfpack_solve (solver.py:50-60) on the device: banded LU + FP64 iterative refinement
    (no |Ax - b| assertion, as in the reference)."""
    solver_options = solver_options or {}
    x = torch.empty_like(b)
    info = N.CpfemSolveInfo()
    refine = int(solver_options.get("umfpack_solver", {}).get("refine", 2)) if isinstance(
        solver_options.get("umfpack_solver"), dict) else 2
    rc = problem.ctx._lib.cpfem_solve_direct(problem.ctx.h, stream_handle(), ptr(mat), ptr(b), ptr(x), refine,
                                             ctypes.byref(info))
    problem.solver_info.append(dict(iterations=info.iterations, outer=info.outer, residual=info.residual,
                                    b_norm=info.b_norm, direct=True))
    N.check(rc)
    return x


def linear_solver(problem, mat, b, x0, solver_options):
    """solver.py:91-113 dispatch."""
    if not (solver_options.keys() & {"jax_solver", "umfpack_solver", "petsc_solver", "custom_solver"}):
        solver_options["jax_solver"] = {}
    if "jax_solver" in solver_options:
        precond = solver_options["jax_solver"].get("precond", True)
        return krylov_solve(problem, mat, b, x0, solver_options, precond)
    if "umfpack_solver" in solver_options:
        return direct_solve(problem, mat, b, solver_options)
    if "petsc_solver" in solver_options:
        return krylov_solve(problem, mat, b, x0, solver_options, True)
    if "custom_solver" in solver_options:
        return solver_options["custom_solver"](BlockEllMatrix(problem.ctx, mat), b, x0, solver_options)
    raise NotImplementedError("Unknown linear solver.")


def line_search(problem, dofs, inc):
    """solver.py:240-276: halve the step while the BC residual norm keeps falling (3 tries)."""
    def res_norm(alpha):
        d = dofs + alpha * inc
        res, _ = problem.assemble(d, want_matrix=False, apply_bc=True)
        return norm(problem, res, allow_nonfinite=True)

    alpha = 1.0
    rn = res_norm(alpha)
    for i in range(3):
        alpha *= 0.5
        rh = res_norm(alpha)
        logger.debug(f"i = {i}, res_norm = {rn}, res_norm_half = {rh}")
        if rh > rn:
            alpha *= 2.0
            break
        rn = rh
    return dofs + alpha * inc


def bc_x0(problem, dofs):
    x0 = torch.empty_like(dofs)
    N.check(problem.ctx._lib.cpfem_bc_x0(problem.ctx.h, stream_handle(), ptr(dofs), ptr(x0)))
    return x0


def linear_incremental_solver(problem, res, mat, dofs, solver_options):
    """solver.py:213-237: b = -res, x0 = ubar - u on Dirichlet dofs, then dofs += inc."""
    b = -res
    x0 = bc_x0(problem, dofs)
    inc = linear_solver(problem, mat, b, x0, solver_options)
    if solver_options.get("line_search_flag", False):
        return line_search(problem, dofs, inc)
    return dofs + inc


def solver(problem, solver_options={}):
    """Newton with row elimination; returns [sol (nodes, 3)] (solver.py:310-437)."""
    start = time.time()
    if solver_options.get("initial_guess", None) is not None:
        dofs = problem.flatten(solver_options["initial_guess"]).clone()
    else:
        dofs = torch.zeros(problem.num_total_dofs_all_vars, dtype=torch.float64, device=device())
    rel_tol = solver_options.get("rel_tol", 1e-8)
    tol = solver_options.get("tol", 1e-6)
    status = []
    res, mat = problem.assemble(dofs)
    res_val = norm(problem, res, status, allow_nonfinite=True)
    res_val_initial = res_val
    rel_res_val = res_val / res_val_initial if res_val_initial != 0 else math.nan
    logger.debug(f"Before, l_2 res = {res_val}, relative l_2 res = {rel_res_val}")
    newton = 0
    max_newton = int(solver_options.get("max_newton", 200))
    while (rel_res_val > rel_tol) and (res_val > tol):
        if newton >= max_newton:
            raise RuntimeError(f"Newton did not converge in {max_newton} iterations (l_2 res = {res_val})")
        dofs = linear_incremental_solver(problem, res, mat, dofs, solver_options)
        res, mat = problem.assemble(dofs)
        res_val = norm(problem, res, status, allow_nonfinite=True)
        rel_res_val = res_val / res_val_initial
        newton += 1
        logger.debug(f"l_2 res = {res_val}, relative l_2 res = {rel_res_val}")
    if not math.isfinite(res_val):
        flags = 0
        for s in status:
            flags |= s
        logger.error("non-finite residual after %d Newton iterations; point flags: %s", newton,
                     ", ".join(n for b, n in ((N.PT_CAPPED, "local Newton hit its iteration cap"),
                                              (N.PT_NONFINITE, "non-finite local state")) if flags & b) or "none")
    # solver.py:416-417: a NaN residual ends the loop above and fails here
    assert math.isfinite(res_val), "res_val contains NaN, stop the program!"
    assert bool(torch.isfinite(dofs).all()), "dofs contains NaN, stop the program!"
    flags = 0
    for s in status:
        flags |= s
    problem.last_solve = dict(newton=newton, point_status=flags, seconds=time.time() - start)
    logger.info(f"Solve took {time.time() - start} [s]")
    return problem.unflatten_fn_sol_list(dofs)
