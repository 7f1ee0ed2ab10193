"""Three of the reference's own example runs, replayed through the drop-in surface and checked
against the VTU output the reference ships with them (tests/golden/make_golden.py extracts it).

steel304  polycrystal_304steel.py:66-265: Neper mesh domain0_mesh16.msh (16^3 hex8, 8 grains,
          grain = gmsh:physical - 1), orientations quat.txt[:8, 1:], grain g -> orientation g,
          the 304-steel model (models_304steel.py), BCs [corner(0,0,Lz) x, corner y, bottom z,
          top z], 1% tension in 50 steps of dt 0.002 s; VTUs of load steps 1-16.
ta        singlecrystal_tantalum.py:66-251: Neper mesh10.msh (10^3 hex8, one grain; the stored
          VTUs have its 1000 cells, while the driver text names mesh16.msh), the identity
          orientation, the tantalum model (models_tantalum.py, BCC12), the same BCs, 1.25%
          compression in 50 steps of dt 0.25 s; VTUs of all 50 steps.
dp        polycrystal_DPsteel_inhomo.py:62-265: Neper n1-id1-mesh10.msh (10^3 hex8, 2 mm),
          quat.txt[:20, 1:] (8 rows: orientation indices 8-19 clamp to row 7, as the JAX
          gather of models_DPsteel_inhomo.py:212 does), the per-cell orientation and phase
          draws recorded in the VTU (the driver draws both without a seed), ferrite/martensite
          constants per phase (models_DPsteel_inhomo.py:58-110, BCC24) and the cubic
          stiffness per cell that the VTU records (its C11/C12/C44 cell data: the stored run
          paired phase 0 with 4.174e5/2.424e5/2.111e5 and phase 1 with 2.314e5/1.347e5/
          1.164e5, the reverse of models_DPsteel_inhomo.py:84-102; with the stiffness of the
          model text, step 1 is already 11% off, with the recorded one every step agrees), BCs
          [left x, front y, bottom z, top z], 1% tension in 50 steps of dt 0.2 s, line search
          on; VTUs of all 50 steps.

All runs use the jax-solver path with the previous solution as initial guess, then
compute_avg_stress and update_int_vars_gp per step, as the drivers do; the four checked steps
go through cpfem_amd.utils.save_sol (float32 VTU) and are read back, as the reference's were.  The VTU values are
float32 (mesh ingest against the VTU connectivity is checked in test_host_gmsh.py).

Checks: per-step cell means of sigma_xx, sigma_yy, sigma_zz and von Mises on the scale of
|mean sigma_zz|; per-cell sigma and nodal displacement of four stored steps on the scale of
the field's largest magnitude.  Tolerances, with the largest errors observed on MI355X, are in
TOL.  The ta and dp runs agree to the float32 storage.

On steel304 the per-cell bar is looser, and it is set by the reference's own stopping rule,
not by the kernels: on this 0.016 mm cube the nodal forces are ~1e-4, so the absolute Newton
tolerance 1e-6 (solver.py:385, 407) stops every load step at a residual of up to ~1e-2 of the
force scale (in step 1 after a single Newton iteration, whose iterate is a BiCGSTAB solution
at a 1e-10 relative recurrence residual).  Two correct runs that differ only in rounding stop
at different points inside that band and carry the difference into the history, so per cell
they agree to that band's width, not better: 2e-2 of the field's largest magnitude.  Observed:
7.5e-4 per cell with one compilation of the kernels, 7.3e-3 with another that fuses
multiply-adds per source expression only (same algorithm, different last bits); cell means
average the differences out (1.3e-6).

With the single-corner x/y pins of the steel304 and ta runs, rotation about z through the
pinned node is a zero-energy mode: the reference's BiCGSTAB and this one each leave a
rounding-seeded amount of it in u (observed up to 9% of max|u| for steel304), which changes no
stress; the displacement check removes its least-squares multiple from the difference first.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# case -> (tolerance on means / |mean szz|, on per-cell sigma / max|sigma|, on nodal u / max|u|);
# largest observed: steel304 1.3e-6, 7.3e-3, 4.0e-4 (per-cell bar: see above); ta 4.5e-8, 2.5e-7, 8.8e-8;
# dp 2.5e-9, 5.1e-8, 4.8e-8 before this side's own float32 VTU rounding (each float32 storage
# rounds a per-cell value by up to 6e-8 of it, so the two stored copies may differ by 1.2e-7 more)
TOL = {
    "steel304": (2e-6, 2e-2, 2e-2),
    "ta": (2e-7, 1e-6, 5e-7),
    "dp": (1e-8, 3e-7, 3e-7),
}


def _on(c, v):
    return lambda point: np.isclose(point[c], v, atol=1e-5)


def _setup(case):
    import cpfem_amd as C
    ref = np.load(os.path.join(GOLDEN, f"{case}_vtu.npz"))
    m = C.read_gmsh(os.path.join(GOLDEN, f"{case}_mesh.msh.gz"))
    mesh = C.Mesh(m.points, m.cells_dict["hexahedron"])
    Lx, Ly, Lz = (np.max(mesh.points[:, k]) for k in range(3))
    phase_inds = None
    opts = {}
    if case == "steel304":
        cell_ori_inds = np.take(np.arange(8), m.cell_data["gmsh:physical"][0] - 1, axis=0)
        quat = np.loadtxt(os.path.join(GOLDEN, "steel304_quat.txt"))[:8, 1:]
        disps, ts, steps, mat = np.linspace(0.0, 0.01 * Lx, 51), np.linspace(0.0, 0.1, 51), 16, C.STEEL304
    elif case == "ta":
        cell_ori_inds = ref["cell_ori_inds"]
        quat = np.array([[1.0, 0.0, 0.0, 0.0]])
        disps, ts, steps, mat = np.linspace(0.0, -0.0125 * Lx, 51), np.linspace(0.0, 12.5, 51), 50, C.TANTALUM
    else:
        cell_ori_inds, phase_inds = ref["cell_ori_inds"], ref["phase_inds"]
        quat = np.loadtxt(os.path.join(GOLDEN, "dp_quat.txt"))[:20, 1:]
        disps, ts, steps, mat = np.linspace(0.0, 0.01 * Lx, 51), np.linspace(0.0, 10.0, 51), 50, C.DP_FERRITE
        opts = {"line_search_flag": True}

    def corner(point):
        return np.logical_and(np.logical_and(_on(0, 0.0)(point), _on(1, 0.0)(point)), _on(2, Lz)(point))

    zero = lambda point: 0.0
    get_top = lambda d: (lambda point: d)
    locs = [corner, corner] if case != "dp" else [_on(0, 0.0), _on(1, 0.0)]
    info = [locs + [_on(2, 0.0), _on(2, Lz)], [0, 1, 2, 2], [zero, zero, zero, get_top(disps[0])]]
    extra = ()
    if phase_inds is not None:
        # the run's stiffness per phase, as its VTU records it per cell
        stiff = [{k: float(np.unique(ref[k][phase_inds == p])[0]) for k in ("C11", "C12", "C44")} for p in (0, 1)]
        extra = (phase_inds, (C.DP_FERRITE.replace(**stiff[0]), C.DP_MARTENSITE.replace(**stiff[1])))
    prob = C.CrystalPlasticity(mesh, vec=3, dim=3, ele_type="HEX8", dirichlet_bc_info=info,
                               additional_info=(quat, cell_ori_inds) + extra, material=mat)
    return prob, info, get_top, disps, ts, steps, opts, ref


@pytest.mark.parametrize("case", ["steel304", "ta", "dp"])
def test_reference_run_reproduces_stored_fields(case, tmp_path):
    import cpfem_amd as C
    from cpfem_amd.utils import cell_stress_fields, save_sol
    from test_host_vtu import _decode
    prob, info, get_top, disps, ts, steps, opts, ref = _setup(case)
    sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
    params = prob.internal_vars
    means, full, usol = [], {}, {}
    for i in range(steps):
        prob.dt = ts[i + 1] - ts[i]
        info[-1][-1] = get_top(disps[i + 1])
        prob.fes[0].update_Dirichlet_boundary_conditions(info)
        prob.set_params(params)
        sol = C.solver(prob, solver_options={"jax_solver": {}, "initial_guess": sol, **opts})
        s = prob.compute_avg_stress(sol[0], params).cpu().numpy()
        xx, yy, zz = s[:, 0, 0], s[:, 1, 1], s[:, 2, 2]
        vm = np.sqrt(0.5 * ((xx - yy) ** 2 + (yy - zz) ** 2 + (zz - xx) ** 2)
                     + 3.0 * (s[:, 0, 1] ** 2 + s[:, 1, 2] ** 2 + s[:, 0, 2] ** 2))
        means.append([xx.mean(), yy.mean(), zz.mean(), vm.mean()])
        if i in ref["full_steps"]:
            # through the drivers' VTU output and back, as the reference's fields were stored
            path = str(tmp_path / f"u_{i:03d}.vtu")
            save_sol(prob.fes[0], sol[0], path, cell_infos=list(cell_stress_fields(s).items()))
            a = _decode(path)
            full[i] = np.stack([a["sigma_xx"], a["sigma_yy"], a["sigma_zz"]], axis=1).astype(np.float64)
            usol[i] = a["sol"].reshape(-1, 3).astype(np.float64)
            np.testing.assert_allclose(a["von_Mises_stress"], vm, rtol=1e-6)
        params = prob.update_int_vars_gp(sol[0], params)
    t_mean, t_cell, t_u = TOL[case]
    pts = prob.fes[0].points
    means = np.array(means)
    rm = ref["mean_sigma"]
    # sigma_zz and von Mises carry the load; the lateral means are small, compare on the zz scale
    e_mean = (np.abs(means - rm) / np.abs(rm[:, 2:3])).max(axis=1)
    e_cell, e_u = [], []
    for k, i in enumerate(ref["full_steps"]):
        rf = ref["full_sigma"][k].astype(np.float64)
        e_cell.append(np.abs(full[int(i)] - rf).max() / np.abs(rf).max())
        ru = ref["full_sol"][k].astype(np.float64)
        d = usol[int(i)] - ru
        if case != "dp":
            # corner x/y pins leave rotation about the z axis through the pinned node free:
            # u = theta (-y, x, 0) costs no energy, so each run's Krylov solves carry their own
            # rounding-seeded amount of it; take out the least-squares theta before comparing
            mode = np.stack([-pts[:, 1], pts[:, 0], np.zeros(len(pts))], axis=1)
            d = d - (d * mode).sum() / (mode * mode).sum() * mode
        e_u.append(np.abs(d).max() / np.abs(ru).max())
    fmt = lambda v: " ".join(f"{x:.2e}" for x in v)
    print(f"{case}: mean err max {e_mean.max():.2e} (step {e_mean.argmax() + 1}), cell {fmt(e_cell)}, u {fmt(e_u)}")
    assert e_mean.max() < t_mean, (case, e_mean)
    assert max(e_cell) < t_cell, (case, e_cell)
    assert max(e_u) < t_u, (case, e_u)
