"""Oracle runs of small instances of the five benchmark workloads, written as fixtures for
the end-to-end GPU tests (tests/test_gpu_workloads.py).

The oracle (oracle/, a restatement of the reference algorithm pinned to the reference-held
curves) runs the same driver loop as the device path: identical meshes, grains,
orientations, materials, boundary conditions and load histories, built by
cpfem_amd.configs with the same seeds.  Output per case: the mean-over-cells Cauchy stress
after every load step, (steps, 9) row-major, in <case>.txt.

Run from the repository root:  python tests/golden/make_oracle_fixtures.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "jax-cpfem_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import cp_local, fe  # noqa: E402
from cpfem_amd import configs  # noqa: E402
from cpfem_amd.materials import load_slip_table  # noqa: E402
from cpfem_amd.rotations import get_rot_mat  # noqa: E402

# name -> (builder, load steps)
CASES = {
    "wl_c1_cu_8": (lambda: configs.config1(steps=40), 10),  # 10 of 40 steps (10 steps of 1e-3 overflow)
    "wl_c2_fcc_4": (lambda: configs.config2(n=4), 6),
    "wl_c3_bcc24_4": (lambda: configs.config3(n=4), 4),
    "wl_c4_calib_4": (lambda: configs.config4(n=4), 4),
    "wl_c5_dual_4": (lambda: configs.config5(n=4), 3),
}


def oracle_problem(w):
    m = w.material
    nc = len(w.mesh.cells)
    idx = np.clip(np.asarray(w.cell_ori_inds, dtype=np.int64), 0, len(w.quat) - 1)  # JAX index clamping
    R = get_rot_mat(w.quat)[idx]
    P0 = cp_local.schmid_tensors(load_slip_table(m.slip))
    if w.phase_inds is None:
        om = cp_local.Material(P0, xm=m.xm, gss_a=m.gss_a, h=m.h, t_sat=m.t_sat, r=m.r, C11=m.C11, C12=m.C12,
                               C44=m.C44, self_block=m.self_block, ao=m.ao)
        g0 = m.g0
    else:  # dual phase: per-point parameters by phase (models_DPsteel_inhomo.py:55-186)
        ph = np.repeat(np.asarray(w.phase_inds, dtype=np.int64), 8)
        mats = w.phase_materials
        pick = lambda f: np.array([getattr(mats[k], f) for k in ph], dtype=np.float64)
        om = cp_local.Material(P0, xm=pick("xm"), gss_a=pick("gss_a"), h=pick("h"), t_sat=pick("t_sat"), r=pick("r"),
                               C11=mats[0].C11, C12=mats[0].C12, C44=mats[0].C44, self_block=m.self_block, ao=m.ao)
        g0 = np.array([mats[k].g0 for k in np.asarray(w.phase_inds)], dtype=np.float64)
    prob = fe.FEProblem(w.mesh.points, w.mesh.cells, om, R, g0, dt=1.0, tol=m.tol, max_sub_step=m.max_sub_step)
    return prob


def run_case(w, steps):
    prob = oracle_problem(w)
    pts = w.mesh.points
    L = w.L
    on = lambda c, v: np.argwhere(np.isclose(pts[:, c], v, atol=1e-5)).ravel()
    bcs = [(on(0, 0.0), 0, 0.0), (on(1, 0.0), 1, 0.0), (on(2, 0.0), 2, 0.0), (on(2, L), 2, None)]
    kw = dict(tol=w.solver_options.get("tol", 1e-6), line_search_flag=w.solver_options.get("line_search_flag", False))
    sig = fe.run_load_steps(prob, bcs, w.disps[:steps + 1], w.ts[:steps + 1], solver_kw=kw)
    return sig.reshape(steps, 9)


def main():
    names = sys.argv[1:] or list(CASES)
    for name in names:
        build, steps = CASES[name]
        t0 = time.time()
        sig = run_case(build(), steps)
        np.savetxt(os.path.join(HERE, name + ".txt"), sig, fmt="%.17g")
        print(f"{name}: {steps} steps, szz = {sig[:, 8]}, {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
