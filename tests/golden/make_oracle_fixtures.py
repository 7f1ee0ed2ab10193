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
    # load step 1 of the dual-phase workload at 24^3 (110,592 points): the largest size the
    # oracle runs in reasonable time (about 35 minutes of NumPy); written only when named
    "wl_c5_dual_24": (lambda: configs.config5(n=24), 1),
}
SLOW = {"wl_c5_dual_24"}


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


# --------------------------------------------------------------------------- gradients
# The reference's AD gradient restated (oracle/ad.py) on small instances of the two gradient
# workloads, with tightened tolerances (local |r| <= GRAD_LOCAL_TOL, global Newton to
# GRAD_SOLVER) so that the converged states, hence the implicit-function gradients, are
# reproducible to ~1e-9 by any correct implementation.  Written: <case>.txt with the query
# point, the loss and the gradient (see the case functions).
GRAD_LOCAL_TOL = 1e-10
GRAD_SOLVER = dict(tol=1e-9, rel_tol=1e-14, krylov=dict(tol=1e-13, atol=1e-20))


def _ad_setup(w, steps):
    import torch
    from oracle import ad
    m = w.material
    mesh = ad.Mesh(w.mesh.points, w.mesh.cells)
    consts = ad.Consts(cp_local.schmid_tensors(load_slip_table(m.slip)), self_block=m.self_block, ao=m.ao,
                       tol=GRAD_LOCAL_TOL, max_sub_step=m.max_sub_step)
    pts = w.mesh.points
    on = lambda c, v: np.argwhere(np.isclose(pts[:, c], v, atol=1e-5)).ravel()
    bcs = [(on(0, 0.0), 0, 0.0), (on(1, 0.0), 1, 0.0), (on(2, 0.0), 2, 0.0), (on(2, w.L), 2, None)]
    Q = w.n_qp
    g0, mat8 = configs.history_inputs(w, device=torch.device("cpu"))
    sim = lambda g0_, mat8_, R_: ad.simulate(mesh, consts, bcs, w.disps[:steps + 1], w.ts[:steps + 1],
                                             g0_.reshape(Q, -1), mat8_.reshape(Q, 8), R_, GRAD_SOLVER)
    return sim, g0, mat8


def grad_calibration(n=4, steps=4):
    """Calibration workload at n^3, `steps` load steps: loss = sum_steps (szz - target)^2 with
    the target from the hidden parameters, gradient w.r.t. the five scale factors of
    (g0, h0, g_sat, a, m) at the seeded query point (configs.calibration_query)."""
    import torch
    w = configs.config4(n=n)
    sim, _, _ = _ad_setup(w, steps)
    R = torch.as_tensor(get_rot_mat(w.quat)[np.clip(w.cell_ori_inds, 0, len(w.quat) - 1)]).repeat_interleave(8, 0)
    f = lambda th: sim(*configs.calibration_inputs(w, None, th), R)[:, 2, 2]
    with torch.no_grad():
        target = f(torch.ones(5, dtype=torch.float64))
    th = torch.as_tensor(configs.calibration_query()).requires_grad_(True)
    loss = ((f(th) - target) ** 2).sum()
    loss.backward()
    return np.concatenate([th.detach().numpy(), [loss.item()], th.grad.numpy(), target.numpy()])


def grad_orientation(n=4, steps=3, seed=1):
    """Dual-phase workload at n^3, `steps` load steps: loss = (final mean szz - 1.01 x its value at
    the seeded angles)^2, gradient w.r.t. all per-grain z-y-x Euler angles (R = R_euler R_quat
    per grain, poly_sensitivity_rot_mesh2.py:41-65,203) at angles U(-30, 30) degrees."""
    import torch
    from oracle import ad
    w = configs.config5(n=n)
    sim, g0, mat8 = _ad_setup(w, steps)
    Rq = torch.as_tensor(get_rot_mat(w.quat))
    gid = torch.as_tensor(np.clip(w.cell_ori_inds, 0, len(w.quat) - 1)).repeat_interleave(8)
    f = lambda e: sim(g0, mat8, (ad.euler_zyx(e) @ Rq)[gid])[-1, 2, 2]
    e0 = torch.as_tensor(np.random.default_rng(seed).uniform(-30, 30, size=(len(w.quat), 3)))
    with torch.no_grad():
        target = 1.01 * f(e0)
    e = e0.clone().requires_grad_(True)
    loss = (f(e) - target) ** 2
    loss.backward()
    return np.concatenate([[target.item(), loss.item()], e0.numpy().ravel(), e.grad.numpy().ravel()])


GRAD_CASES = {"grad_c4_calib_4": grad_calibration, "grad_c5_dual_4": grad_orientation}


def main():
    names = sys.argv[1:] or [c for c in CASES if c not in SLOW] + list(GRAD_CASES)
    for name in names:
        t0 = time.time()
        if name in GRAD_CASES:
            out = GRAD_CASES[name]()
            np.savetxt(os.path.join(HERE, name + ".txt"), out, fmt="%.17g")
            print(f"{name}: {time.time() - t0:.0f} s", flush=True)
            continue
        build, steps = CASES[name]
        sig = run_case(build(), steps)
        np.savetxt(os.path.join(HERE, name + ".txt"), sig, fmt="%.17g")
        print(f"{name}: {steps} steps, szz = {sig[:, 8]}, {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
