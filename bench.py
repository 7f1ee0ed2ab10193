#!/usr/bin/env python
"""Benchmark of the crystal-plasticity hot path on one MI355X.

A step is one load step of the polycrystal workload (global Newton to the reference
tolerances: per-point Kalidindi update + tangent, hex8 assembly into the block-ELL Jacobian,
Jacobi-BiCGSTAB solves; then the cell-averaged Cauchy stress and the history update), run
through the drop-in ``solver(problem, solver_options)`` loop.  Default workload: 32^3 hex8,
262,144 quadrature points, 100 Voronoi grains, FCC12, 2% tension in 40 steps of 5e-4.

value = quadrature points x timed load steps / wall time (whole job, inputs resident in HBM).
The roofline object refers to the kernel class with the largest total time in the timed
region, timed with HIP events on its launch stream; the CPU baseline is the oracle timed on
a bounded sample of the same state on one host core (see the 'sample' string).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jax-cpfem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0    # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def spmv_bytes(prob, value_bytes=8):
    """Algorithmic HBM bytes of one Krylov SpMV launch (k_kpq / k_kst, equal counts): the
    block values (value_bytes each), one int32 column index per 3x3 block (plus the slot
    counts, read unless every row is padded to 27 slots; padding is not counted), the two FP64 vectors the
    operand is formed from (read once each), and five FP64 row vectors (k_kpq: reads rhat and
    the Jacobi vector, writes phat, q, M q; k_kst: reads r, q, writes s, shat, t)."""
    nslot, _ = prob.ctx.pattern()
    nnzb = int(nslot.sum())
    n = prob.ctx.n_dofs
    idx = 4 * nnzb + (0 if prob.ctx.n_slots == 27 else 4 * prob.ctx.n_nodes)
    return 9 * value_bytes * nnzb + idx + 7 * 8 * n, nnzb


def spmv_kernel_names(n_dofs, cus, tm):
    """Names of the BiCGSTAB SpMV kernel pair the library runs for n_dofs rows on a device
    with `cus` compute units (the library's krylov_spmv_form): 128-row tiles when there are
    two per CU, else 64-row tiles, with every slot's loads in flight when all of them fit
    three blocks per CU."""
    if -(-n_dofs // 128) >= 2 * cus:
        return [f"cpfem::k_kpq<{tm}, 2>", f"cpfem::k_kst<{tm}, 2>"]
    if -(-n_dofs // 64) <= 3 * cus:
        return [f"cpfem::k_kpq_all<{tm}>", f"cpfem::k_kst_all<{tm}>"]
    return [f"cpfem::k_kpq<{tm}, 1>", f"cpfem::k_kst<{tm}, 1>"]


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


def cpu_baseline(prob, w, params, sol, per_step, sample_cells=512, krylov_iters=16):
    """Oracle timed on a bounded sample of the state at the start of the timed region."""
    import numpy as np
    import torch
    from threadpoolctl import threadpool_limits
    from oracle import cp_local, fe

    mat = w.material
    om = cp_local.Material(cp_local.schmid_tensors(__import__("cpfem_amd").materials.load_slip_table(mat.slip)),
                           xm=mat.xm, gss_a=mat.gss_a, h=mat.h, t_sat=mat.t_sat, r=mat.r, C11=mat.C11, C12=mat.C12,
                           C44=mat.C44, self_block=mat.self_block)
    S = sample_cells
    cells = w.mesh.cells[:S]
    u = sol[0].reshape(-1).cpu().numpy()
    sg, jxw = prob.fes[0].shape_grads[:S], prob.fes[0].JxW[:S]
    ug = np.einsum("cai,cqaj->cqij", u.reshape(-1, 3)[cells], sg).reshape(-1, 3, 3)
    take = lambda t, comps: t[:S].reshape((S * 8,) + comps).cpu().numpy()
    Ns = prob.num_slip_sys
    Fp, g, sl, R = take(params[0], (3, 3)), take(params[1], (Ns,)), take(params[2], (Ns,)), take(params[3], (3, 3))
    dt = float(w.ts[1] - w.ts[0])
    with threadpool_limits(1):
        t0 = time.perf_counter()
        out = cp_local.local_update(ug, Fp, g, sl, R, om, dt, max_sub_step=mat.max_sub_step)
        A = out["A"].reshape(S, 8, 3, 3, 3, 3)
        P = out["P"].reshape(S, 8, 3, 3)
        np.einsum("cqaj,cqijkl,cqbl,cq->caibk", sg, A, sg, jxw, optimize=True)
        np.einsum("cqij,cqaj,cq->cai", P, sg, jxw)
        t_asm = (time.perf_counter() - t0) / (S * 8)
        t0 = time.perf_counter()
        cp_local.local_update(ug, Fp, g, sl, R, om, dt, max_sub_step=mat.max_sub_step, want_tangent=False)
        t_post = (time.perf_counter() - t0) / (S * 8)
        prob.set_params(params)  # the snapshot history that belongs to `sol`
        _, mat_vals = prob.assemble(sol[0].reshape(-1))
        Asp = __import__("cpfem_amd").BlockEllMatrix(prob.ctx, mat_vals).to_scipy()
        if not np.isfinite(Asp.data).all():
            raise RuntimeError("CPU baseline: the sampled Jacobian is not finite")
        # seeded right-hand side far from convergence, so that all krylov_iters iterations run
        b = Asp @ np.random.default_rng(0).normal(size=Asp.shape[0])
        inv_d = 1.0 / Asp.diagonal()
        t0 = time.perf_counter()
        _, k_run = fe.bicgstab_jax(Asp, b, np.zeros_like(b), M=lambda v: v * inv_d, maxiter=krylov_iters,
                                   tol=0.0, atol=0.0)
        t_iter = (time.perf_counter() - t0) / max(int(k_run), 1)
        if int(k_run) != krylov_iters:
            raise RuntimeError(f"CPU baseline BiCGSTAB ran {k_run} of {krylov_iters} iterations")
    Q = w.n_qp
    step_s = (per_step["assemblies"] * Q * t_asm + per_step["krylov_iters"] * t_iter + 2 * Q * t_post)
    return dict(value=Q / step_s, unit="quad-pt load-step updates/s", cores=1, kind="port",
                sample=(f"oracle (NumPy/SciPy FP64, 1 thread) on {S * 8} points of the timed state "
                        f"({t_asm * 1e6:.1f} us/point update+tangent+element, {t_post * 1e6:.1f} us/point post) and "
                        f"{krylov_iters} BiCGSTAB iterations on the full {Asp.shape[0]}-dof Jacobian "
                        f"({t_iter * 1e3:.1f} ms/iter), scaled by the GPU run's per-step counts "
                        f"({per_step['assemblies']:.1f} assemblies, {per_step['krylov_iters']:.0f} Krylov iterations); "
                        f"est. {step_s:.1f} s per load step"))


def adjoint_bench(n=16, steps=None, k=10, fd=True, mixed=True):
    """Gradient of a calibration loss w.r.t. five hardening scales through the whole load
    history (hand-written adjoint, history checkpointed every k steps).  Workload: the
    calibration case (64 grains, 304-like stiffness, hidden (g0, h0, gs, a, m) =
    (90, 450, 280, 2.2, 0.02), 5% in 50 steps, the steel calibration driver's solver options),
    loss = sum over steps of (szz - target)^2 with the target from the hidden parameters and
    the query = hidden x (1 + U(-0.2, 0.2)) (seeded).  Default mesh 16^3: at 24^3 the forward
    solve of the hidden parameters itself fails in load step 2 (a point's local Newton
    overflows), so no gradient exists there.  Also reports a central difference (relative step
    1e-6) of the g0 component and the gradient of the mixed-precision forward and adjoint."""
    import numpy as np
    import torch
    import cpfem_amd as C
    from cpfem_amd import configs
    w = configs.config4(n=n)
    if steps is not None:
        w.disps, w.ts = w.disps[:steps + 1], w.ts[:steps + 1]
    prob = configs.build_problem(w)
    nc, Ns, m = len(w.mesh.cells), prob.num_slip_sys, w.material
    load = C.LoadHistory(w.disps, w.ts, w.dirichlet_bc_info)
    R = torch.as_tensor(C.get_rot_mat(w.quat)[w.cell_ori_inds], device="cuda")[:, None].expand(nc, 8, 3, 3)

    def inputs(theta):  # theta: scales of (g0, h0, g_sat, a, m)
        base = torch.tensor([m.xm, m.gss_a, m.h, m.t_sat, m.r, m.C11, m.C12, m.C44], dtype=torch.float64, device="cuda")
        one = torch.ones((), dtype=torch.float64, device="cuda")
        sc = torch.stack([theta[4], theta[3], theta[1], theta[2], one, one, one, one])
        return (m.g0 * theta[0]).expand(nc, 8, Ns), (base * sc).expand(nc, 8, 8)

    opts = dict(w.solver_options)

    def loss_of(theta, target, precision="fp64"):
        g0, m8 = inputs(theta)
        o = dict(opts, precision=precision)
        szz = C.simulate(prob, load, g0, m8, R, checkpoint_every=k, solver_options=o,
                         adjoint_solver_options={"precision": precision})[:, 2, 2]
        return ((szz - target) ** 2).sum() if target is not None else szz

    hidden = torch.ones(5, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(0)
    q0 = torch.as_tensor(1.0 + rng.uniform(-0.2, 0.2, size=5), device="cuda")
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
                            grad_max_rel_err=float(np.max(np.abs(gx - g64)) / np.max(np.abs(g64))))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--mesh", type=int, default=None, help="override the mesh size n (n^3 cells)")
    ap.add_argument("--profile-sample", type=int, default=64,
                    help="time every n-th launch of each kernel class with HIP events (events perturb the run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare-mixed", action="store_true", help="rerun the timed steps in the other precision")
    ap.add_argument("--copy-calibration", action="store_true", help="also run a plain copy kernel (counter scale)")
    ap.add_argument("--adjoint", action="store_true", help="also time the history gradient of the calibration workload")
    ap.add_argument("--adjoint-steps", type=int, default=None, help="load steps of the adjoint workload (default: all 50)")
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
    import cpfem_amd as C
    from cpfem_amd import _native as N
    from cpfem_amd import configs

    torch.cuda.set_device(0)
    kw = {} if args.mesh is None else {"n": args.mesh}
    w = configs.CONFIGS[args.config](**kw)
    log(f"[bench] workload {w.name}: {len(w.mesh.cells)} cells, {w.n_qp} points")
    prob = configs.build_problem(w)
    L, h = prob.ctx._lib, prob.ctx.h
    params = prob.internal_vars
    sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
    opts = {"precision": args.precision}
    W, K = args.warmup, args.steps
    if W + K > len(w.ts) - 1:
        raise SystemExit(f"warmup + steps exceeds the {len(w.ts) - 1} load steps of the workload")
    t0 = time.perf_counter()
    params, sol, _ = configs.run_steps(prob, w, params, sol, 0, W, opts,
                                       on_step=lambda i, p, s: log(f"[bench] warmup step {i + 1}: szz={s[2, 2]:.6f}"))
    log(f"[bench] warmup {W} steps {time.perf_counter() - t0:.2f}s")
    snap = ([p.clone() for p in params], [sol[0].clone()])

    counts = dict(newton=0, krylov=0, solves=0)

    def on_step(i, p, s):
        counts["newton"] += p.last_solve["newton"]
        log(f"[bench] step {i + 1}: szz={s[2, 2]:.6f} newton={p.last_solve['newton']} "
            f"krylov={[d['iterations'] for d in p.solver_info[-p.last_solve['newton']:]] if p.last_solve['newton'] else []}")

    prob.solver_info.clear()
    N.check(L.cpfem_profile(h, 1, args.profile_sample))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    params, sol, sig = configs.run_steps(prob, w, params, sol, W, W + K, opts, on_step=on_step)
    torch.cuda.synchronize()
    T = time.perf_counter() - t0
    ms = (ctypes.c_double * 8)()
    la = (ctypes.c_int64 * 8)()
    ti = (ctypes.c_int64 * 8)()
    N.check(L.cpfem_profile_read(h, 8, ms, la, ti))
    N.check(L.cpfem_profile(h, 0, 1))
    kry = [d["iterations"] for d in prob.solver_info]
    assemblies = counts["newton"] + K  # one assembly per Newton iteration plus the initial one
    per_step = dict(assemblies=assemblies / K, krylov_iters=sum(max(k, 0) for k in kry) / K)
    classes = N.PROF_CLASSES
    prof = {c: dict(launches=int(la[k]), timed=int(ti[k]), avg_ms=(ms[k] / ti[k] if ti[k] else None),
                    est_total_ms=(ms[k] / ti[k] * la[k] if ti[k] else 0.0)) for k, c in enumerate(classes)}
    dom = max(prof, key=lambda c: prof[c]["est_total_ms"])
    log(f"[bench] timed {K} steps in {T:.3f}s; profile: " +
        ", ".join(f"{c}={prof[c]['est_total_ms']:.1f}ms/{prof[c]['launches']}" for c in classes))

    Q = w.n_qp
    value = Q * K / T
    vb = 4 if args.precision == "mixed" else 8
    sb, nnzb = spmv_bytes(prob, value_bytes=vb)
    roof = None
    if dom == "krylov_spmv" and prof[dom]["avg_ms"]:
        # Launches that move the matrix: two per BiCGSTAB iteration plus the initial residual of
        # each solve; the launches queued past convergence return at once and move nothing, so
        # the class time (sampled mean x all launches) is divided by the working launches only.
        working = 2 * sum(max(k, 0) for k in kry) + len(kry)
        t_launch = prof[dom]["est_total_ms"] / working
        ach = sb / (t_launch * 1e-3) / 1e9
        tkey = f"poly{args.mesh or 32}" if args.config == 2 else None
        tm = "float" if vb == 4 else "double"
        names = spmv_kernel_names(prob.ctx.n_dofs, torch.cuda.get_device_properties(0).multi_processor_count, tm)
        traffic = measured_traffic(tkey, args.precision, names) if tkey else None
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                    traffic=(round(traffic) if traffic else None),
                    traffic_source=("rocprofv3 FETCH_SIZE+WRITE_SIZE per dispatch, copy-kernel scaled (profiles/)"
                                    if traffic else None),
                    kernel="k_kpq/k_kst (BiCGSTAB SpMV with fused p/s updates and dots)",
                    bytes_per_launch=sb, avg_launch_us=round(t_launch * 1e3, 2), working_launches=working,
                    timing="hipExtLaunchKernelGGL start/stop events (kernel-bound) on every "
                           f"{args.profile_sample}th launch of the class, on its launch stream")
    elif prof[dom]["avg_ms"]:
        # another class dominates (small meshes: the FP64-VALU point kernels); no byte model here
        roof = dict(bound=None, achieved=None, peak=None, unit=None, frac=None, traffic=None,
                    kernel=dom, avg_launch_us=round(prof[dom]["avg_ms"] * 1e3, 2))

    detail = dict(
        kernel_profile=prof, dominant=dom, newton_per_step=counts["newton"] / K, krylov_per_solve=kry,
        point_evals_per_s_in_update_kernel=(Q * prof["element_qp"]["launches"] / (prof["element_qp"]["est_total_ms"] * 1e-3)
                                            if prof["element_qp"]["est_total_ms"] else None),
        sigma_bar_zz=[float(s[2, 2]) for s in sig], nnz_blocks=nnzb, n_dofs=prob.ctx.n_dofs,
        spmv_bytes_per_launch=sb, load_step_s=T / K)

    if args.compare_mixed:
        other = "mixed" if args.precision == "fp64" else "fp64"
        p2, s2 = snap
        prob.solver_info.clear()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, _, sig2 = configs.run_steps(prob, w, p2, s2, W, W + K, {"precision": other})
        torch.cuda.synchronize()
        T2 = time.perf_counter() - t1
        ref = np.array([s[2, 2] for s in (sig if args.precision == "fp64" else sig2)])
        mix = np.array([s[2, 2] for s in (sig2 if args.precision == "fp64" else sig)])
        detail["compare"] = dict(other_precision=other, other_load_step_s=T2 / K,
                                 speedup_mixed_over_fp64=(T / T2 if other == "mixed" else T2 / T),
                                 max_rel_err_sigma_zz=float(np.max(np.abs(mix - ref) / np.abs(ref))),
                                 other_krylov=[d["iterations"] for d in prob.solver_info])
        log(f"[bench] {other}: {T2 / K:.3f}s/step, rel err {detail['compare']['max_rel_err_sigma_zz']:.2e}")

    if args.copy_calibration:
        n = 1 << 26
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        N.check(L.cpfem_profile(h, 1, 1))
        for _ in range(5):
            N.check(L.cpfem_copy(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), ctypes.c_void_p(src.data_ptr()),
                                 ctypes.c_void_p(dst.data_ptr()), n))
        N.check(L.cpfem_profile_read(h, 8, ms, la, ti))
        N.check(L.cpfem_profile(h, 0, 1))
        cms = ms[7] / ti[7]
        detail["copy_calibration"] = dict(bytes=16 * n, avg_ms=cms, GBps=16 * n / (cms * 1e-3) / 1e9)

    if args.adjoint:
        log("[bench] adjoint: calibration-workload history gradient ...")
        detail["adjoint"] = adjoint_bench(steps=args.adjoint_steps)
        log(f"[bench] adjoint: {json.dumps(detail['adjoint'])}")
        log(f"[bench] adjoint: forward {detail['adjoint']['forward_s']:.2f}s, backward {detail['adjoint']['backward_s']:.2f}s")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle sample) ...")
        cpu = cpu_baseline(prob, w, snap[0], snap[1], per_step)
        log(f"[bench] cpu {cpu['value']:.1f} {cpu['unit']}")

    line = {
        "metric": "CP quad-pt updates/s, load-step time, adjoint-grad time (mixed vs FP64)",
        "value": round(value, 1),
        "unit": "quad-pt load-step updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(T / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32-krylov",
        "data": "synthetic (seeded Voronoi polycrystal, uniform SO(3) orientations)",
        "config": {"workload": w.name, "n_qp": Q, "load_steps": f"{W + 1}..{W + K} of {len(w.ts) - 1}",
                   "precision": args.precision, "solver": "row-elimination Newton, Jacobi-BiCGSTAB tol=atol=1e-10"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
