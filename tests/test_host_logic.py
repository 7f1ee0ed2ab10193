"""Host-side logic of the drop-in package that runs without a GPU: Dirichlet lists (jax-fem
semantics), location-function evaluation, the synthetic workloads, slip tables, meshes."""
import os

import numpy as np
import pytest

pytest.importorskip("torch")


def _fe(n=2):
    from cpfem_amd.generate_mesh import Mesh, box_mesh
    from cpfem_amd.problem import FiniteElement
    m = box_mesh(n, n, n, 1.0, 1.0, 1.0)
    return FiniteElement(Mesh(m.points, m.cells_dict["hexahedron"]), 3, 3, "HEX8", None)


def test_dirichlet_lists_last_wins_and_values():
    """solver.py:119-153 applies the lists in order (.set), so a dof named twice takes the
    value of the later list; node lists come from the location functions."""
    fe = _fe(2)
    bottom = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    corner = lambda p: np.isclose(p[0], 0.0, atol=1e-5) & np.isclose(p[1], 0.0, atol=1e-5) & np.isclose(p[2], 0.0, atol=1e-5)
    fe.update_Dirichlet_boundary_conditions([[bottom, corner], [2, 2], [lambda p: 0.0, lambda p: 7.0]])
    assert len(fe.node_inds_list[0]) == 9 and len(fe.node_inds_list[1]) == 1
    dofs, vals = fe.dirichlet_dofs()
    assert len(dofs) == 9 and np.all(np.diff(dofs) > 0)
    c = int(fe.node_inds_list[1][0])
    assert vals[list(dofs).index(3 * c + 2)] == 7.0
    assert np.count_nonzero(vals) == 1


def test_location_functions_vectorised_and_scalar_fallback():
    """Vectorised NumPy location/value functions and plain-Python scalar ones give the same lists."""
    fe = _fe(3)
    top_vec = lambda p: np.isclose(p[2], 1.0, atol=1e-5)

    def top_scalar(p):  # not vectorisable: Python `if` on an array raises
        if abs(p[2] - 1.0) < 1e-5:
            return True
        return False

    def val_scalar(p):
        if p[0] > 0.5:
            return 2e-3
        return 1e-3

    fe.update_Dirichlet_boundary_conditions([[top_vec, top_scalar], [2, 1], [lambda p: 1e-3, val_scalar]])
    np.testing.assert_array_equal(fe.node_inds_list[0], fe.node_inds_list[1])
    pts = fe.points[fe.node_inds_list[1]]
    np.testing.assert_allclose(fe.vals_list[1], np.where(pts[:, 0] > 0.5, 2e-3, 1e-3))
    np.testing.assert_allclose(fe.vals_list[0], 1e-3)


def test_location_function_that_reduces_over_its_point():
    """jax-fem evaluates location and value functions point by point (vmap).  A function that
    reduces over its point, such as np.all(np.isclose(p, c)), returns one 0-d value for the
    whole coordinate-row array, which must not be copied to every point: here it selects
    exactly the origin, and a value function reducing the same way gives per-point values."""
    fe = _fe(2)
    origin = lambda p: np.all(np.isclose(p, 0.0, atol=1e-5))
    far = lambda p: np.all(np.isclose(p, 1.0, atol=1e-5))
    val = lambda p: float(np.sum(p))  # 0-d for any input
    fe.update_Dirichlet_boundary_conditions([[origin, far, lambda p: np.isclose(p[2], 1.0, atol=1e-5)], [0, 1, 2],
                                             [lambda p: 0.0, lambda p: 3.0, val]])
    np.testing.assert_array_equal(fe.points[fe.node_inds_list[0]], [[0.0, 0.0, 0.0]])
    np.testing.assert_array_equal(fe.points[fe.node_inds_list[1]], [[1.0, 1.0, 1.0]])
    np.testing.assert_allclose(fe.vals_list[1], [3.0])
    top = fe.points[fe.node_inds_list[2]]
    assert len(top) == 9
    np.testing.assert_allclose(fe.vals_list[2], top.sum(axis=1))


def test_workloads_are_deterministic_and_well_formed():
    from cpfem_amd import configs
    a, b = configs.config2(n=4), configs.config2(n=4)
    np.testing.assert_array_equal(a.cell_ori_inds, b.cell_ori_inds)
    np.testing.assert_array_equal(a.quat, b.quat)
    assert a.n_qp == 8 * 64 and a.cell_ori_inds.min() >= 0 and a.cell_ori_inds.max() < 100
    np.testing.assert_allclose(np.linalg.norm(a.quat, axis=1), 1.0, atol=1e-14)
    # 2% in 40 steps at a strain rate of 1e-3 /s
    np.testing.assert_allclose(np.diff(a.disps), 5e-4)
    np.testing.assert_allclose(np.diff(a.ts), 0.5)
    c5 = configs.config5(n=4)
    assert set(np.unique(c5.phase_inds)) <= {0, 1}
    hard, soft = c5.phase_materials[1], c5.phase_materials[0]
    assert hard.g0 == 4 * soft.g0 and hard.h == 2 * soft.h
    c4 = configs.config4(n=4)
    assert c4.solver_options == {"tol": 1e-7, "line_search_flag": True}
    c1 = configs.config1(steps=40)
    assert len(c1.ts) == 41 and c1.ts[-1] == pytest.approx(10.0)


def test_slip_tables_and_box_mesh():
    from cpfem_amd.generate_mesh import box_mesh
    from cpfem_amd.materials import load_slip_table, slip_vectors
    for name, n in (("fcc12", 12), ("bcc12", 12), ("bcc24", 24)):
        d, nrm = slip_vectors(load_slip_table(name))
        assert d.shape == (n, 3)
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-15)
        np.testing.assert_allclose(np.einsum("ij,ij->i", d, nrm), 0.0, atol=1e-15)
    m = box_mesh(2, 3, 4, 1.0, 2.0, 3.0)
    cells = m.cells_dict["hexahedron"]
    assert m.points.shape == (3 * 4 * 5, 3) and cells.shape == (24, 8)
    # node id = i + (Nx+1)(j + (Ny+1)k)
    np.testing.assert_allclose(m.points[1 + 3 * (2 + 4 * 3)], [0.5, 4.0 / 3.0, 2.25])
    # every hexahedron has positive volume in VTK order (corner 0 -> 1 -> 3 -> 4 edge triple)
    p = m.points[cells]
    e1, e2, e3 = p[:, 1] - p[:, 0], p[:, 3] - p[:, 0], p[:, 4] - p[:, 0]
    assert np.all(np.einsum("ij,ij->i", np.cross(e1, e2), e3) > 0)


def test_measured_traffic_table():
    """bench.py reads the committed rocprof traffic table for its roofline."""
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    t = bench.measured_traffic("c2_32", "fp64", ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"])
    assert t is not None and 60e6 < t < 120e6
    assert bench.measured_traffic("c2_32", "fp64", ["cpfem::no_such_kernel"]) is None
    assert bench.measured_traffic("c2_32", "mixed", ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"]) is None
    assert bench.measured_traffic("nowhere", "fp64", ["cpfem::k_kpq_flat1"]) is None
    t64 = bench.measured_traffic("c5_64", "fp64", ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"])
    assert t64 is not None and 600e6 < t64 < 800e6
    # kernel names per SpMV form (the form itself is the library's report, cpfem_krylov_config)
    assert bench.spmv_kernel_names(2, "double") == ["cpfem::k_kpq<double, 2>", "cpfem::k_kst<double, 2>"]
    assert bench.spmv_kernel_names(1, "float") == ["cpfem::k_kpq<float, 1>", "cpfem::k_kst<float, 1>"]
    assert bench.spmv_kernel_names(0, "double")[1] == "cpfem::k_kst_all<double>"
    assert bench.spmv_kernel_names(3, "float") == ["cpfem::k_kpq_flat<float>", "cpfem::k_kst_flat<float>"]
    assert bench.spmv_kernel_names(3, "double") == ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"]
    assert bench.spmv_kernel_names(4, "double") == ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"]
    assert bench.spmv_kernel_names(4, "float") == ["cpfem::k_kpq_flat<float>", "cpfem::k_kst_flat<float>"]
