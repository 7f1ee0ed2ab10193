"""Failure handling of the drop-in solver, each condition forced from the host (every kernel
still exits normally):

* a local iteration cap that plastic points cannot meet: the residual norm returns
  CPFEM_ELOCAL and the solver raises (the reference's uncapped local loop would not return);
* a NaN in the history: the residual is NaN, the Newton loop ends and the reference's
  assertion text is raised (crystal_plasticity_OR_design/solver.py:416), with the point's
  non-finite flag in the log;
* a Krylov iteration limit the solve cannot meet: the |Ax - b| < 0.1 check of jax_solve
  fails with the reference's AssertionError text (solver.py:45).
"""
import logging

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(material, disp=2e-3):
    import cpfem_amd as C
    m = C.box_mesh(2, 2, 2, 1.0, 1.0, 1.0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    front = lambda p: np.isclose(p[1], 0.0, atol=1e-5)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    top = lambda p: np.isclose(p[2], 1.0, atol=1e-5)
    zero = lambda p: 0.0
    info = [[left, front, bottom, top], [0, 1, 2, 2], [zero, zero, zero, lambda p: disp]]
    rng = np.random.default_rng(0)
    quat = rng.normal(size=(3, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    prob = C.CrystalPlasticity(mesh, dirichlet_bc_info=info, additional_info=(quat, rng.integers(0, 3, size=8)),
                               material=material)
    prob.dt = 1.0
    return prob


def test_local_iteration_cap_raises_elocal():
    import cpfem_amd as C
    from cpfem_amd import _native as N
    prob = _problem(C.STEEL304.replace(max_iter=1))
    with pytest.raises(N.CpfemError) as e:
        C.solver(prob, {"jax_solver": {}})
    assert e.value.code == N.ELOCAL
    assert "iteration cap" in str(e.value)
    # the same step with the default cap converges
    prob = _problem(C.STEEL304)
    C.solver(prob, {"jax_solver": {}})
    assert prob.last_solve["point_status"] == 0


def test_nan_history_raises_the_reference_assertion(caplog):
    import cpfem_amd as C
    from cpfem_amd import _native as N
    prob = _problem(C.STEEL304)
    params = [p.clone() for p in prob.internal_vars]
    params[1][3, 5, 7] = float("nan")  # one slip resistance of one point
    prob.set_params(params)
    with caplog.at_level(logging.ERROR, logger="cpfem_amd"):
        with pytest.raises(AssertionError, match="res_val contains NaN, stop the program!"):
            C.solver(prob, {"jax_solver": {}})
    assert "non-finite local state" in caplog.text
    # the status word was read and cleared: a clean problem on the same context solves
    prob.set_params([p.clone() for p in prob.internal_vars])
    del N


def test_krylov_iteration_limit_raises_the_reference_assertion():
    import cpfem_amd as C
    prob = _problem(C.STEEL304, disp=1.0)  # |b| >> 0.1 in the first Newton system
    with pytest.raises(AssertionError, match="JAX linear solver failed to converge with err = "):
        C.solver(prob, {"jax_solver": {}, "krylov_maxiter": 1})
    assert prob.solver_info[-1]["iterations"] == 1
