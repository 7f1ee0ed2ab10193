"""The reference's own 304-steel polycrystal run, replayed through the drop-in surface.

polycrystal_304steel.py:66-265: the Neper mesh domain0_mesh16.msh (16^3 hex8, 8 grains, read
with meshio, grain = gmsh:physical - 1), orientations quat.txt[:8, 1:], grain g -> orientation
g, the 304-steel model (models_304steel.py), BCs [corner(0,0,Lz) x, corner y, bottom z, top z],
1% tension in 50 steps of dt 0.002 s, the jax solver, compute_avg_stress and
update_int_vars_gp per step.  The reference wrote the per-cell Cauchy stress of load steps
1-16 to u_000..u_015.vtu (float32); tests/golden/steel304_vtu.npz holds those values
(tests/golden/make_golden.py).

Checks: the mesh read by cpfem_amd.read_gmsh has the VTU's connectivity and orientation
indices; per-step cell means of sigma_xx, sigma_yy, sigma_zz and von Mises match the
reference's to 2e-6 of |mean sigma_zz| (observed 0.7-4.6e-7); per-cell sigma of steps 1, 5,
10, 16 matches to 1e-3 of the field's largest magnitude (observed 7.5e-4 at step 1, 1.3-1.7e-4
later).  The per-cell bar is set by the reference's own stopping rule, not by the kernels: on
this 0.016 mm cube the nodal forces are ~1e-4, so the absolute Newton tolerance 1e-6
(solver.py:385, 407) is met after a single Newton iteration in step 1, and that iterate is the
BiCGSTAB solution at a 1e-10 relative recurrence residual, which two correct runs that differ
only in rounding reproduce to ~1e-4 per cell; cell means average those differences out.  A
tighter Newton tolerance cannot be used to sharpen the check: the residual floors near 1e-10
on this domain.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup():
    import cpfem_amd as C
    m = C.read_gmsh(os.path.join(GOLDEN, "steel304_mesh16.msh.gz"))
    cell_grain_inds = m.cell_data["gmsh:physical"][0] - 1
    grain_oris_inds = np.arange(8)
    cell_ori_inds = np.take(grain_oris_inds, cell_grain_inds, axis=0)
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    quat = np.loadtxt(os.path.join(GOLDEN, "steel304_quat.txt"))[:8, 1:]
    Lx, Ly, Lz = (np.max(mesh.points[:, k]) for k in range(3))
    disps = np.linspace(0.0, 0.01 * Lx, 51)
    ts = np.linspace(0.0, 0.1, 51)

    def corner(point):
        return np.logical_and(np.logical_and(np.isclose(point[0], 0.0, atol=1e-5), np.isclose(point[1], 0.0, atol=1e-5)),
                              np.isclose(point[2], Lz, atol=1e-5))

    bottom = lambda point: np.isclose(point[2], 0.0, atol=1e-5)
    top = lambda point: np.isclose(point[2], Lz, atol=1e-5)
    zero = lambda point: 0.0
    get_top = lambda d: (lambda point: d)
    info = [[corner, corner, bottom, top], [0, 1, 2, 2], [zero, zero, zero, get_top(disps[0])]]
    prob = C.CrystalPlasticity(mesh, vec=3, dim=3, ele_type="HEX8", dirichlet_bc_info=info,
                               additional_info=(quat, cell_ori_inds), material=C.STEEL304)
    return prob, info, get_top, disps, ts, cell_ori_inds, m


def test_mesh_ingest_matches_reference_output():
    ref = np.load(os.path.join(GOLDEN, "steel304_vtu.npz"))
    prob, _, _, _, _, cell_ori_inds, m = _setup()
    np.testing.assert_array_equal(m.cells_dict["hexahedron"], ref["connectivity"])
    np.testing.assert_array_equal(cell_ori_inds, ref["cell_ori_inds"])


def test_polycrystal_304steel_reproduces_reference_stress():
    import cpfem_amd as C
    ref = np.load(os.path.join(GOLDEN, "steel304_vtu.npz"))
    prob, info, get_top, disps, ts, _, _ = _setup()
    sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
    params = prob.internal_vars
    means, full = [], {}
    for i in range(16):
        prob.dt = ts[i + 1] - ts[i]
        info[-1][-1] = get_top(disps[i + 1])
        prob.fes[0].update_Dirichlet_boundary_conditions(info)
        prob.set_params(params)
        sol = C.solver(prob, solver_options={"jax_solver": {}, "initial_guess": sol})
        s = prob.compute_avg_stress(sol[0], params).cpu().numpy()
        xx, yy, zz = s[:, 0, 0], s[:, 1, 1], s[:, 2, 2]
        vm = np.sqrt(0.5 * ((xx - yy) ** 2 + (yy - zz) ** 2 + (zz - xx) ** 2)
                     + 3.0 * (s[:, 0, 1] ** 2 + s[:, 1, 2] ** 2 + s[:, 0, 2] ** 2))
        means.append([xx.mean(), yy.mean(), zz.mean(), vm.mean()])
        if i in ref["full_steps"]:
            full[i] = np.stack([xx, yy, zz], axis=1)
        params = prob.update_int_vars_gp(sol[0], params)
    means = np.array(means)
    rm = ref["mean_sigma"]
    # sigma_zz and von Mises carry the load; the lateral means are small, compare on the zz scale
    scale = np.abs(rm[:, 2:3])
    assert (np.abs(means - rm) / scale).max() < 2e-6, (means[:, 2], rm[:, 2])
    for k, i in enumerate(ref["full_steps"]):
        rf = ref["full_sigma"][k].astype(np.float64)
        err = np.abs(full[int(i)] - rf).max() / np.abs(rf).max()
        assert err < 1e-3, (int(i), err)
