#!/usr/bin/env python
"""Benchmark of the crystal-plasticity hot path on one MI355X.

Default workload: the dual-phase BCC24 polycrystal of the orientation-design case at 64^3
hex8 (262,144 cells, 2,097,152 quadrature points, 823,875 dofs), 300 Voronoi grains with
seeded uniform-SO(3) orientations, 20% of the grains hard (g0 x4, h0 x2), FP64.

A step is one load step (load step 1: 1e-3 strain, dt = 1 s) re-run from its snapshot (the
state before it), with all of its work: the global Newton to the reference tolerances with
the dual-phase driver's line search (per-point Kalidindi update + consistent tangent, hex8
assembly into the block-ELL Jacobian, Jacobi-BiCGSTAB solves, line-search residuals), then
the cell-averaged Cauchy stress and the history update, through the drop-in
``solver(problem, solver_options)`` loop.  (The reference algorithm, local Newton from S = 0
with cut-half backtracking, does not get through load step 2 of this workload from 24^3 up:
some points' local Newton neither converges nor stays finite, in the oracle as on the device,
so load step 1 is the load step this workload has at full size.)

value = quadrature points x timed load steps / wall time (whole job, inputs resident in HBM).
``roofline`` refers to the kernel class with the largest total time in the timed region,
timed with HIP events on its launch stream; ``detail`` carries the mixed-precision load step,
the adjoint gradients (this workload's load step w.r.t. its 900 Euler angles, and the
calibration workload's 50-step history w.r.t. five hardening parameters), and the point
kernel's FP64 roofline.  ``cpu_baseline`` times the C++/OpenMP restatement of the same
algorithm on the host's cores on a bounded sample of the same state.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jax-cpfem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0    # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector

METRIC = "CP quad-pt updates/s, load-step time, adjoint-grad time (mixed vs FP64)"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------ byte / flop models
def spmv_bytes(prob, value_bytes=8, computed_columns=False, one_operand=False):
    """Algorithmic HBM bytes of one Krylov SpMV launch (mean of the two kernels): the block
    values (value_bytes each), one int32 column index per 3x3 block (plus the slot counts, read
    unless every row is padded to 27 slots; padding is not counted) and FP64 row vectors, each
    read or written once.  Pair recurrence (k_kpq: 7 vectors, k_kst: 5): the two vectors the
    operand is formed from, and k_kpq reads rhat and the Jacobi vector and writes phat, q and
    M q; k_kst reads r, q and writes t.  One-operand recurrence (k_kpq_flat1: 7, k_kst_flat1:
    6): k_kpq_flat1 gathers phat, reads rhat, the Jacobi vector and rh and writes q and the
    (rh, qh) pair; k_kst_flat1 gathers the pair, reads r, q, rhat and writes t.
    computed_columns: the one-row-per-thread form on a box-numbered mesh computes the
    columns of the nodes off the box's faces, so only the face nodes' indices count."""
    nslot, _ = prob.ctx.pattern()
    nnzb = int(nslot.sum())
    n = prob.ctx.n_dofs
    idx = 4 * nnzb + (0 if prob.ctx.n_slots == 27 else 4 * prob.ctx.n_nodes)
    if computed_columns:
        nx, ny, nz = prob.ctx.box_numbering()
        if nx > 0:
            idx -= 4 * 27 * (nx - 2) * (ny - 2) * (nz - 2)
    vec = (7 + 6) / 2 if one_operand else (7 + 5) / 2
    return round(9 * value_bytes * nnzb + idx + vec * 8 * n), nnzb


def point_flops(n_slip, tangent_points, newton_iters, residual_evals):
    """Algorithmic FP64 flops of the assembly's point kernels (k_element_res, the parked points'
    k_element_qp_park and k_point_tangent) from measured totals: the reference algorithm's per-unit counts
    (FMA = 2 flops; div, abs, sign, pow, sqrt = 1 each), F_res(N) = 41 N + 350 per residual
    evaluation, F_jac(N) = 165 N + 2574 per local Jacobian and F_LU = 648 per 9x9 LU and solve
    (once per iteration), F_tan = 9300 per consistent tangent (assemblies with the Jacobian
    only; the line search's residual-only assemblies compute none).  The element stiffness
    (F_elem = 4752 + 144 per point) is k_element_K's work, not counted here."""
    N = n_slip
    f_res, f_jac, f_lu, f_tan = 41 * N + 350, 165 * N + 2574, 648, 9300
    return residual_evals * f_res + newton_iters * (f_jac + f_lu) + tangent_points * f_tan


def spmv_kernel_names(form, tm):
    """Names of the BiCGSTAB SpMV kernel pair of an SpMV form (as the library reports it,
    cpfem_krylov_config): 4 = one row per thread over XCD-local row ranges, 3 = one row per
    thread, 2 = 128-row tiles, 1 = 64-row tiles, 0 = 64-row tiles with every slot's loads in
    flight."""
    if form in (3, 4) and tm == "double":  # the one-operand recurrence of the FP64 flat forms
        return ["cpfem::k_kpq_flat1", "cpfem::k_kst_flat1"]
    return {4: [f"cpfem::k_kpq_flat<{tm}>", f"cpfem::k_kst_flat<{tm}>"],
            3: [f"cpfem::k_kpq_flat<{tm}>", f"cpfem::k_kst_flat<{tm}>"],
            2: [f"cpfem::k_kpq<{tm}, 2>", f"cpfem::k_kst<{tm}, 2>"],
            1: [f"cpfem::k_kpq<{tm}, 1>", f"cpfem::k_kst<{tm}, 1>"],
            0: [f"cpfem::k_kpq_all<{tm}>", f"cpfem::k_kst_all<{tm}>"]}[form]


def measured_traffic(workload_key, precision, kernels):
    """HBM bytes per dispatch of `kernels` (mean) from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/traffic_<workload>.txt), or None if it has no entry."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        f = line.split()  # workload precision kernel-name (may contain spaces) bytes
        if len(f) >= 4 and not line.startswith("#") and f[0] == workload_key and f[1] == precision:
            vals[" ".join(f[2:-1])] = float(f[-1])
    got = [vals[k] for k in kernels if k in vals]
    return sum(got) / len(got) if len(got) == len(kernels) else None


# ------------------------------------------------------------------------------ library profile
class Profile:
    """Kernel-class timing of libcpfem_hip.so (HIP events on the launch stream around every
    `sample`-th launch of a class) plus the point kernel's work counters."""

    def __init__(self, prob, sample):
        from cpfem_amd import _native as N
        self.N, self.L, self.h, self.sample = N, prob.ctx._lib, prob.ctx.h, sample

    def start(self):
        from cpfem_amd.problem import stream_handle
        self.N.check(self.L.cpfem_profile(self.h, 1, self.sample))
        cnt = (ctypes.c_int64 * 3)()
        self.N.check(self.L.cpfem_point_counts(self.h, stream_handle(), 1, cnt))

    def read(self):
        from cpfem_amd.problem import stream_handle
        n = len(self.N.PROF_CLASSES)
        ms, la, ti = (ctypes.c_double * n)(), (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)()
        self.N.check(self.L.cpfem_profile_read(self.h, n, ms, la, ti))
        self.N.check(self.L.cpfem_profile(self.h, 0, 1))
        cnt = (ctypes.c_int64 * 3)()
        self.N.check(self.L.cpfem_point_counts(self.h, stream_handle(), 1, cnt))
        prof = {c: dict(launches=int(la[k]), timed=int(ti[k]), avg_ms=(ms[k] / ti[k] if ti[k] else None),
                        est_total_ms=(ms[k] / ti[k] * la[k] if ti[k] else 0.0))
                for k, c in enumerate(self.N.PROF_CLASSES)}
        return prof, dict(newton_iters=int(cnt[0]), residual_evals=int(cnt[1]), parked_points=int(cnt[2]))


# ------------------------------------------------------------------------------ load steps
class StepRunner:
    """Re-runs load step `index` (0-based) of a workload from a fixed snapshot."""

    def __init__(self, prob, w, params, sol, index):
        self.prob, self.w, self.index = prob, w, index
        self.params, self.sol = params, sol

    def run(self, precision, on_step=None):
        from cpfem_amd import configs
        return configs.run_steps(self.prob, self.w, self.params, [self.sol[0]], self.index, self.index + 1,
                                 {"precision": precision}, on_step=on_step)


# ------------------------------------------------------------------------------ adjoints
def step_adjoint(prob, w, index, szz_ref, precision):
    """Gradient of (mean szz after load step `index` - target)^2 w.r.t. the per-grain z-y-x
    Euler angles (the orientation-design loss of poly_sensitivity_rot_mesh2.py:179-239, with
    R = R_euler(theta) R_quat per grain as at :203), through the forward load step and its
    hand-written adjoint (cpfem_amd.simulate).  theta = 0 (the workload's own orientations);
    target = 1.05 x the FP64 forward value.  Returns timing and the gradient."""
    import numpy as np
    import torch
    import cpfem_amd as C
    from cpfem_amd import configs
    from cpfem_amd.rotations import alpha_rotate_tensor_rank_2, get_rot_mat, rotate_theta_transfer
    if index != 0:
        raise ValueError("the load-step adjoint starts from the initial state")
    nc = len(w.mesh.cells)
    g0, mat8 = configs.history_inputs(w, prob)
    Rq = torch.as_tensor(get_rot_mat(w.quat), device="cuda")
    gid = torch.as_tensor(np.clip(w.cell_ori_inds, 0, len(w.quat) - 1), device="cuda")
    load = C.LoadHistory(w.disps[:2], w.ts[:2], w.dirichlet_bc_info)
    theta = torch.zeros(len(w.quat), 3, dtype=torch.float64, device="cuda", requires_grad=True)
    target = 1.05 * szz_ref
    opts = dict(w.solver_options, precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    R = alpha_rotate_tensor_rank_2(rotate_theta_transfer(theta), Rq)[gid].unsqueeze(1).expand(nc, 8, 3, 3)
    szz = C.simulate(prob, load, g0, mat8, R, checkpoint_every=5, solver_options=opts,
                     adjoint_solver_options={"precision": precision})[-1, 2, 2]
    loss = (szz - target) ** 2
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_info = len(prob.solver_info)
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    adj_iters = sum(max(d["iterations"], 0) for d in prob.solver_info[n_info:])
    return dict(forward_s=t1 - t0, backward_s=t2 - t1, grad_s=t2 - t0, szz=float(szz.detach()), loss=float(loss.detach()),
                backward_krylov_iters=adj_iters, grad=theta.grad.detach().cpu().numpy())


def history_adjoint(n=16, k=10, fd=True, mixed=True):
    """Gradient of the calibration loss w.r.t. the five hardening scales through the whole
    load history (hand-written adjoint, history checkpointed every k steps).  Workload: the
    calibration case (64 grains, 304-like stiffness, hidden (g0, h0, gs, a, m) =
    (90, 450, 280, 2.2, 0.02), 5% in 50 steps, the steel calibration driver's solver options),
    loss = sum over steps of (szz - target)^2 with the target from the hidden parameters and
    the query = hidden x (1 + U(-0.2, 0.2)) (seeded).  Default mesh 16^3: from 20^3 up the
    forward run of the hidden parameters itself fails (a point's local Newton overflows, in the
    oracle as on the device), so no gradient exists there.  Also reports a central difference
    (relative step 1e-6) of the g0 component and the gradient of the mixed-precision forward and
    adjoint."""
    import numpy as np
    import torch
    import cpfem_amd as C
    from cpfem_amd import configs
    w = configs.config4(n=n)
    prob = configs.build_problem(w)
    load = C.LoadHistory(w.disps, w.ts, w.dirichlet_bc_info)
    R = configs.history_rotations(w)
    opts = dict(w.solver_options)

    def loss_of(theta, target, precision="fp64"):
        g0, m8 = configs.calibration_inputs(w, prob, theta)
        o = dict(opts, precision=precision)
        szz = C.simulate(prob, load, g0, m8, R, checkpoint_every=k, solver_options=o,
                         adjoint_solver_options={"precision": precision})[:, 2, 2]
        return ((szz - target) ** 2).sum() if target is not None else szz

    hidden = torch.ones(5, dtype=torch.float64, device="cuda")
    q0 = torch.as_tensor(configs.calibration_query(), device="cuda")
    with torch.no_grad():
        target = loss_of(hidden, None)
    query = q0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = loss_of(query, target)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out = dict(workload=w.name, load_steps=load.n_steps, checkpoint_every=k, n_qp=w.n_qp,
               forward_s=t1 - t0, backward_s=t2 - t1, grad_s=t2 - t0, loss=float(loss),
               grad=[float(v) for v in query.grad.cpu()])
    if fd:
        h = 1e-6 * float(q0[0])
        with torch.no_grad():
            qp, qm = q0.clone(), q0.clone()
            qp[0] += h
            qm[0] -= h
            fdv = (float(loss_of(qp, target)) - float(loss_of(qm, target))) / (2 * h)
        out["fd_g0"] = fdv
        out["fd_g0_rel_err"] = abs(out["grad"][0] - fdv) / abs(fdv)
    if mixed:
        qx = q0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        lx = loss_of(qx, target, "mixed")
        lx.backward()
        torch.cuda.synchronize()
        gx = qx.grad.cpu().numpy()
        g64 = np.array(out["grad"])
        out["mixed"] = dict(grad_s=time.perf_counter() - t3, loss_rel_err=abs(float(lx) - out["loss"]) / abs(out["loss"]),
                            grad_max_rel_err=float(np.max(np.abs(gx - g64)) / np.max(np.abs(g64))),
                            speedup_over_fp64=out["grad_s"] / (time.perf_counter() - t3))
    return out


# ------------------------------------------------------------------------------ CPU baseline
def cpu_baseline(prob, w, params, sol, index, per_step, krylov_iters=40, adjoint_krylov_iters=None):
    """The C++/OpenMP restatement of the reference algorithm (oracle/cpu) on all host cores,
    on a bounded sample of the timed load step: one full assembly (every point's local Newton
    and tangent, element stiffness, CSR gather), one residual-only assembly (the line search's),
    one point pass without tangent and `krylov_iters` Jacobi-BiCGSTAB iterations on the
    assembled CSR matrix at the converged state, scaled by the GPU run's per-step counts of
    Jacobian assemblies, residual-only assemblies, point passes (cell stress, history update)
    and Krylov iterations.  With adjoint_krylov_iters (the GPU backward's transposed-solve
    iterations of the load-step adjoint) also one per-point VJP pass, and the adjoint gradient
    is estimated as the load step plus its backward: two VJP passes, one Jacobian assembly at
    the converged state and the transposed solve."""
    from oracle.cpu import cpu_ref
    cores = cpu_ref.threads()
    mat = cpu_ref.Model.from_problem(prob, params, w)
    u = sol[0].reshape(-1).detach().cpu().numpy()
    # the timed step's Dirichlet values (the history is the snapshot's)
    prob.dt = float(w.ts[index + 1] - w.ts[index])
    t_asm, t_post, t_it, info = mat.time_sample(u, float(prob.dt), krylov_iters, vjp=adjoint_krylov_iters is not None)
    t_res = info["res_s"]
    step_s = (per_step["assemblies"] * t_asm + per_step["residual_assemblies"] * t_res + 2 * t_post
              + per_step["krylov_iters"] * t_it)
    Q = w.n_qp
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out = dict(value=Q / step_s, unit="quad-pt load-step updates/s", cores=cores, kind="port",
               sample=(f"C++/OpenMP restatement of the reference algorithm (oracle/cpu), {cores} threads on "
                       f"{os.cpu_count()} logical CPUs ({model}): one full assembly of all {Q} points at the timed "
                       f"step's converged state ({t_asm:.2f} s: local Newton + tangent + element stiffness + CSR "
                       f"gather; {info['newton_iters'] / Q:.2f} local iterations per point), one residual-only "
                       f"assembly ({t_res:.2f} s), one point pass without tangent ({t_post:.2f} s) and {krylov_iters} "
                       f"Jacobi-BiCGSTAB iterations on the {info['n_dofs']}-dof CSR matrix ({t_it * 1e3:.1f} "
                       f"ms/iteration), scaled by the GPU run's per-step counts ({per_step['assemblies']:.1f} Jacobian "
                       f"assemblies, {per_step['residual_assemblies']:.1f} residual-only assemblies, 2 point passes, "
                       f"{per_step['krylov_iters']:.0f} Krylov iterations): est. {step_s:.1f} s per load step"),
               load_step_s=step_s)
    if adjoint_krylov_iters is not None:
        t_vjp = info["vjp_s"]
        back = 2 * t_vjp + t_asm + adjoint_krylov_iters * t_it
        out["adjoint_grad_s"] = step_s + back
        out["adjoint_sample"] = (f"load step ({step_s:.1f} s) + backward: 2 VJP passes ({t_vjp:.2f} s each), one "
                                 f"Jacobian assembly, {adjoint_krylov_iters} transposed-solve iterations "
                                 f"= est. {step_s + back:.1f} s")
    return out


def direct_bench(n, reps):
    """Banded direct solve (factor + solve + two refinement solves) on the row-eliminated
    elastic Newton matrix of the n^3 FCC polycrystal, against the Krylov solve of the same
    system; one JSON line."""
    import torch
    from cpfem_amd import configs
    from cpfem_amd.solver import direct_solve, krylov_solve
    torch.cuda.set_device(0)
    w = configs.config2(n=n)
    prob = configs.build_problem(w)
    prob.fes[0].update_Dirichlet_boundary_conditions(w.dirichlet_bc_info(w.disps[1]))
    prob.dt = w.ts[1] - w.ts[0]
    prob.set_params(prob.internal_vars)
    res, mat = prob.assemble(torch.zeros(prob.ctx.n_dofs, dtype=torch.float64, device="cuda"))
    b = torch.randn(prob.ctx.n_dofs, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(0))
    direct_solve(prob, mat, b)  # allocates the band storage
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(reps, 1)):
        x = direct_solve(prob, mat, b)
    torch.cuda.synchronize()
    t_direct = (time.perf_counter() - t0) / max(reps, 1)
    info = prob.solver_info[-1]
    t0 = time.perf_counter()
    xk = krylov_solve(prob, mat, b, torch.zeros_like(b))
    torch.cuda.synchronize()
    t_kry = time.perf_counter() - t0
    print(json.dumps({
        "metric": "direct solve time (banded LU + 2 refinement steps)", "value": round(t_direct * 1e3, 2), "unit": "ms",
        "higher_is_better": False, "n_gpus": 1, "steps": max(reps, 1), "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"FCC polycrystal {n}^3 elastic Newton matrix", "n_dofs": prob.ctx.n_dofs},
        "residual_rel": info["residual"] / info["b_norm"], "krylov_ms": round(t_kry * 1e3, 2),
        "krylov_iterations": prob.solver_info[-1]["iterations"],
        "diff_vs_krylov": float(torch.linalg.norm(x - xk) / torch.linalg.norm(x))}))


# ------------------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--mesh", type=int, default=None, help="override the mesh size n (n^3 cells)")
    ap.add_argument("--load-step", type=int, default=1, help="the load step (1-based) every timed step re-runs")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--profile-sample", type=int, default=64,
                    help="time every n-th launch of each kernel class with HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=2, help="timed load steps in the other precision (0: none)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip both adjoint gradients")
    ap.add_argument("--no-history-adjoint", action="store_true", help="skip the calibration workload's history adjoint")
    ap.add_argument("--park", type=int, default=-1,
                    help="park the point kernel's lanes when at most this many of a wave still iterate "
                         "(0: never; -1: the library's default)")
    ap.add_argument("--spmv-form", type=int, default=-1,
                    help="force the BiCGSTAB SpMV form (0-4; -1: the library's choice by size)")
    ap.add_argument("--no-box-columns", action="store_true",
                    help="load every column in the one-row-per-thread SpMV (default: computed on box meshes)")
    ap.add_argument("--history-adjoint-mesh", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true", help="also run a plain copy kernel (counter scale)")
    ap.add_argument("--direct", action="store_true",
                    help="instead of the load steps: time the banded direct solve ('umfpack_solver') on the "
                         "elastic Newton matrix of the FCC polycrystal at --mesh (default 20)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "cpfem_amd runs on one GPU; runs on several GPUs are not supported", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.direct:
        direct_bench(args.mesh or 20, args.steps)
        return

    import numpy as np
    import torch
    from cpfem_amd import _native as N
    from cpfem_amd import configs

    torch.cuda.set_device(0)
    kw = {} if args.mesh is None else {"n": args.mesh}
    w = configs.CONFIGS[args.config](**kw)
    n_mesh = round(len(w.mesh.cells) ** (1 / 3))
    S = args.load_step - 1
    if not 0 <= S < len(w.ts) - 1:
        raise SystemExit(f"--load-step must be in 1..{len(w.ts) - 1}")
    log(f"[bench] workload {w.name}: {len(w.mesh.cells)} cells, {w.n_qp} points; load step {S + 1}")
    t0 = time.perf_counter()
    prob = configs.build_problem(w)
    if args.park >= 0:
        prob.ctx.set_point_park(args.park)
    if args.spmv_form >= 0:
        prob.ctx.set_spmv_form(args.spmv_form)
    if args.no_box_columns:
        prob.ctx.set_box_columns(False)
    log(f"[bench] context built in {time.perf_counter() - t0:.1f}s")
    params = prob.internal_vars
    sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
    if S > 0:
        params, sol, _ = configs.run_steps(prob, w, params, sol, 0, S, {"precision": args.precision})
    runner = StepRunner(prob, w, [p.clone() for p in params], [sol[0].clone()], S)
    W, K = args.warmup, args.steps

    t0 = time.perf_counter()
    for i in range(W):
        _, _, sig = runner.run(args.precision)
        log(f"[bench] warmup {i + 1}/{W}: szz={sig[0][2, 2]:.10f} newton={prob.last_solve['newton']}")
    log(f"[bench] warmup {W} steps {time.perf_counter() - t0:.2f}s")

    counts = dict(newton=0)

    def on_step(i, p, s):
        counts["newton"] += p.last_solve["newton"]

    prof = Profile(prob, args.profile_sample)
    prob.solver_info.clear()
    torch.cuda.synchronize()
    prof.start()
    t0 = time.perf_counter()
    sigs = []
    for i in range(K):
        final_params, final_sol, sig = runner.run(args.precision, on_step=on_step)
        sigs.append(sig[0])
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    kp, work = prof.read()
    kry = [d["iterations"] for d in prob.solver_info]
    log(f"[bench] timed {K} steps in {T:.3f}s ({T / K:.3f} s/step); newton {counts['newton'] / K:.1f}/step; "
        "profile: " + ", ".join(f"{c}={kp[c]['est_total_ms']:.1f}ms/{kp[c]['launches']}" for c in kp))
    assemblies = counts["newton"] + K  # one Jacobian assembly per Newton iteration plus the initial one
    # every assembly launches the point kernel once; the others are the line search's residuals
    res_assemblies = max(kp["element_qp"]["launches"] - assemblies, 0)
    per_step = dict(assemblies=assemblies / K, residual_assemblies=res_assemblies / K,
                    krylov_iters=sum(max(k, 0) for k in kry) / K)

    Q = w.n_qp
    value = Q * K / T
    vb = 4 if args.precision == "mixed" else 8
    form = prob.ctx.krylov_config()[0]
    sb, nnzb = spmv_bytes(prob, value_bytes=vb, computed_columns=form >= 3 and not args.no_box_columns,
                          one_operand=form >= 3 and vb == 8)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    tm = "float" if vb == 4 else "double"
    names = spmv_kernel_names(form, tm)
    tkey = f"c{args.config}_{n_mesh}"
    point_ms = lambda: (kp["element_qp"]["est_total_ms"] + kp["element_qp_park"]["est_total_ms"]
                        + kp["point_tangent"]["est_total_ms"])
    dom = max(("krylov_spmv", "element_qp", "element_K", "gather_matrix", "krylov_vector", "point_local"),
              key=lambda c: point_ms() if c == "element_qp" else kp[c]["est_total_ms"])
    # Launches that move the matrix: two per BiCGSTAB iteration plus the initial residual of each
    # solve (the launches queued past convergence return at once and move nothing, so the class
    # time, sampled mean x all launches, is divided by the working launches only)
    working = 2 * sum(max(k, 0) for k in kry) + len(kry)
    spmv_roof = None
    if kp["krylov_spmv"]["avg_ms"] and working:
        t_launch = kp["krylov_spmv"]["est_total_ms"] / working
        ach = sb / (t_launch * 1e-3) / 1e9
        traffic = measured_traffic(tkey, args.precision, names)
        spmv_roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(ach / HBM_PEAK_GBS, 4), traffic=(round(traffic) if traffic else None),
                         traffic_source=(f"rocprofv3 FETCH_SIZE+WRITE_SIZE per dispatch, copy-kernel scaled "
                                         f"(profiles/traffic_{tkey}.txt)" if traffic else None),
                         kernel=f"{names[0]} / {names[1]} (BiCGSTAB SpMV with fused p/s updates and dots)",
                         bytes_per_launch=sb, avg_launch_us=round(t_launch * 1e3, 2), working_launches=working,
                         timing="hipExtLaunchKernelGGL start/stop events (kernel-bound) on every "
                                f"{args.profile_sample}th launch of the class, on its launch stream")
    point_roof = None
    qp = kp["element_qp"]
    if qp["avg_ms"] and qp["launches"]:
        tangent_launches = assemblies  # one Jacobian assembly per Newton iteration plus the initial one
        flops = point_flops(prob.num_slip_sys, Q * tangent_launches, work["newton_iters"], work["residual_evals"])
        t_ms = point_ms()
        ach_tf = flops / (t_ms * 1e-3) / 1e12
        point_roof = dict(bound="fp64 valu", achieved=round(ach_tf, 3), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                          frac=round(ach_tf / FP64_PEAK_TFLOPS, 4),
                          kernel=f"cpfem::k_element_res<{prob.num_slip_sys}> + k_point_tangent + k_element_qp_park "
                                 "(local Newton + P; consistent tangent per point)",
                          launches=qp["launches"], tangent_launches=tangent_launches,
                          avg_launch_ms=round(t_ms / qp["launches"], 4), park_lanes=args.park,
                          parked_points_per_launch=round(work["parked_points"] / qp["launches"], 1),
                          flops_per_point_launch=round(flops / (Q * qp["launches"]), 1),
                          newton_iters_per_point_launch=round(work["newton_iters"] / (Q * qp["launches"]), 3),
                          residual_evals_per_point_launch=round(work["residual_evals"] / (Q * qp["launches"]), 3),
                          flop_model="F_res(N)=41N+350 per residual evaluation, F_jac(N)=165N+2574 + F_LU=648 per "
                                     "iteration, F_tan=9300 per tangent (reference algorithm counts)")
    roof = spmv_roof if dom == "krylov_spmv" else (point_roof if dom == "element_qp" else
                                                    dict(bound=None, achieved=None, peak=None, unit=None, frac=None,
                                                         traffic=None, kernel=dom))
    szz = [float(s[2, 2]) for s in sigs]
    detail = dict(
        kernel_profile=kp, dominant=dom, newton_per_step=counts["newton"] / K, krylov_per_solve=kry[: 40],
        per_step=per_step, point_work=work, sigma_bar_zz=szz[0], sigma_bar_zz_spread=max(szz) - min(szz),
        nnz_blocks=nnzb, n_dofs=prob.ctx.n_dofs, krylov_form=form, spmv_bytes_per_launch=sb, load_step_s=T / K,
        spmv_roofline=spmv_roof, point_roofline=point_roof)

    if args.mixed_steps > 0:
        other = "mixed" if args.precision == "fp64" else "fp64"
        runner.run(other)  # warm-up (the FP32 copy of the matrix is allocated on first use)
        prob.solver_info.clear()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.mixed_steps):
            _, _, sig2 = runner.run(other)
        torch.cuda.synchronize()
        T2 = (time.perf_counter() - t1) / args.mixed_steps
        a, b = szz[0], float(sig2[0][2, 2])
        ref, mix = (a, b) if args.precision == "fp64" else (b, a)
        detail["compare"] = dict(other_precision=other, other_load_step_s=T2, load_steps=args.mixed_steps,
                                 speedup_mixed_over_fp64=((T / K) / T2 if other == "mixed" else T2 / (T / K)),
                                 rel_err_sigma_zz=abs(mix - ref) / abs(ref),
                                 other_krylov_per_solve=[d["iterations"] for d in prob.solver_info][: 40])
        log(f"[bench] {other}: {T2:.3f}s/step, speed-up {detail['compare']['speedup_mixed_over_fp64']:.3f}, "
            f"rel err {detail['compare']['rel_err_sigma_zz']:.2e}")

    if args.copy_calibration:
        n = 1 << 26
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        from cpfem_amd.problem import stream_handle
        N.check(prob.ctx._lib.cpfem_profile(prob.ctx.h, 1, 1))
        for _ in range(5):
            N.check(prob.ctx._lib.cpfem_copy(prob.ctx.h, stream_handle(), ctypes.c_void_p(src.data_ptr()),
                                             ctypes.c_void_p(dst.data_ptr()), n))
        cp, _ = prof.read()
        detail["copy_calibration"] = dict(bytes=16 * n, avg_ms=cp["copy"]["avg_ms"],
                                          GBps=16 * n / (cp["copy"]["avg_ms"] * 1e-3) / 1e9)
        del src, dst

    if not args.no_adjoint:
        if args.config == 5 and S == 0:
            log("[bench] load-step adjoint (900 Euler angles) ...")
            # one untimed gradient first: the first backward of a process pays one-time costs
            # (first launches of the VJP and transpose kernels, the adjoint's device buffers),
            # 0.2 s of scatter in an otherwise ~0.5 s backward
            step_adjoint(prob, w, S, szz[0], "fp64")
            a64 = step_adjoint(prob, w, S, szz[0], "fp64")
            amx = step_adjoint(prob, w, S, szz[0], "mixed")
            g64, gmx = a64.pop("grad"), amx.pop("grad")
            a64["mixed"] = dict(grad_s=amx["grad_s"], forward_s=amx["forward_s"], backward_s=amx["backward_s"],
                                speedup_over_fp64=a64["grad_s"] / amx["grad_s"],
                                loss_rel_err=abs(amx["loss"] - a64["loss"]) / abs(a64["loss"]),
                                grad_max_rel_err=float(np.abs(gmx - g64).max() / np.abs(g64).max()))
            a64.update(n_angles=int(g64.size), grad_norm=float(np.linalg.norm(g64)), szz_rel_err_vs_timed_step=abs(
                a64["szz"] - szz[0]) / abs(szz[0]), warmup_gradients=1)
            detail["adjoint_load_step"] = a64
            log(f"[bench] load-step adjoint: {json.dumps(a64)}")
        if not args.no_history_adjoint:
            log("[bench] history adjoint (calibration workload) ...")
            detail["adjoint_history"] = history_adjoint(n=args.history_adjoint_mesh)
            log(f"[bench] history adjoint: {json.dumps(detail['adjoint_history'])}")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (C++/OpenMP restatement) ...")
        adj = detail.get("adjoint_load_step")
        cpu = cpu_baseline(prob, w, runner.params, [final_sol[0]], S, per_step,
                           adjoint_krylov_iters=adj["backward_krylov_iters"] if adj else None)
        detail["speedup_vs_cpu_load_step"] = cpu["load_step_s"] / (T / K)
        if adj:
            detail["speedup_vs_cpu_adjoint_grad"] = dict(fp64=cpu["adjoint_grad_s"] / adj["grad_s"],
                                                         mixed=cpu["adjoint_grad_s"] / adj["mixed"]["grad_s"])
        if "compare" in detail and args.precision == "fp64":
            detail["speedup_vs_cpu_load_step_mixed"] = cpu["load_step_s"] / detail["compare"]["other_load_step_s"]
        log(f"[bench] cpu {cpu['value']:.1f} {cpu['unit']} ({cpu['load_step_s']:.1f} s per load step)")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "quad-pt load-step updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(T / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32-krylov",
        "data": "synthetic (seeded Voronoi polycrystal, uniform SO(3) orientations)",
        "config": {"workload": w.name, "n_qp": Q, "load_step": f"{S + 1} of {len(w.ts) - 1}, re-run from its snapshot",
                   "precision": args.precision, "solver": "row-elimination Newton, Jacobi-BiCGSTAB tol=atol=1e-10",
                   "solver_options": w.solver_options},
        "roofline": roof,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(line, default=float), flush=True)


if __name__ == "__main__":
    main()
