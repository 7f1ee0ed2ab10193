"""The forms of the BiCGSTAB SpMV kernels.  Four are tested at a size that selects them on a
256-CU device: 64-row tiles with every slot's loads in flight (16^3, 14,739 dofs: 231 tiles),
64-row tiles one slot at a time (26^3, 59,049 dofs: 923 tiles, above three per CU), 128-row
tiles (28^3, 73,167 dofs: 572 tiles, two per CU), one row per thread over XCD-local row
ranges (32^3, 107,811 dofs; picked on 27-slot meshes from 100,000 rows); the fifth, one row
per thread in launch order, is picked from 700,000 rows (the 64^3 benchmark system, covered
by test_gpu_fullsize.py).  The two one-row-per-thread forms are also forced here on small
systems.  The oracle is too slow at these sizes, so each
case checks what fixes the answer independently of the device kernels, on the first Newton
system of the 304-like polycrystal (elastic tangent, Dirichlet values in x0), with the true
residual computed by scipy from the exported CSR.  The recurrence stops at
|r| <= 1e-10 |b|; the true residual then sits at that level plus the recurrence's rounding
drift, which grows with the initial residual r0 = b - A x0 (here far larger than b: b is
nonzero only in the Dirichlet rows, x0 carries their values into the coupled rows), about
k eps |r0| after k iterations.  Bar: |b - A x| <= 1e-10 |b| + 1e-9 |r0| in FP64 and in the
mixed form (whose outer loop stops on the FP64 residual); a wrong slot or coefficient leaves
a residual of the order of |r0|.  The mixed solution lands within 1e-6 of the FP64 one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _form(prob):
    """The SpMV form the library runs on this context (cpfem_krylov_config): 2 = 128-row tiles,
    1 = 64-row tiles, 0 = 64-row tiles with all slot loads in flight."""
    import ctypes
    from cpfem_amd import _native as N
    out = (ctypes.c_int32 * 3)()
    N.check(prob.ctx._lib.cpfem_krylov_config(prob.ctx.h, out))
    return int(out[0])


def _first_system(n):
    from cpfem_amd import configs
    from cpfem_amd.solver import bc_x0
    w = configs.config2(n=n)
    prob = configs.build_problem(w)
    prob.fes[0].update_Dirichlet_boundary_conditions(w.dirichlet_bc_info(w.disps[1]))
    prob.dt = w.ts[1] - w.ts[0]
    prob.set_params(prob.internal_vars)
    dofs = torch.zeros(prob.ctx.n_dofs, dtype=torch.float64, device="cuda")
    res, mat = prob.assemble(dofs)
    return prob, mat, -res, bc_x0(prob, dofs)


@pytest.mark.parametrize("n,form", [(16, 0), (26, 1), (28, 2), (32, 4)])
def test_spmv_form_solves_first_newton_system(n, form):
    import cpfem_amd as C
    from cpfem_amd.solver import krylov_solve
    prob, mat, b, x0 = _first_system(n)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    picked = _form(prob)
    print(f"n={n}: {prob.ctx.n_dofs} dofs, {cus} CUs, SpMV form {picked}", flush=True)
    if cus == 256:
        assert picked == form
    x = krylov_solve(prob, mat, b, x0)
    it = prob.solver_info[-1]["iterations"]
    A = C.BlockEllMatrix(prob.ctx, mat).to_scipy()
    xh, bh, x0h = x.cpu().numpy(), b.cpu().numpy(), x0.cpu().numpy()
    nb, nr0 = np.linalg.norm(bh), np.linalg.norm(bh - A @ x0h)
    bar = 1e-10 * nb + 1e-9 * nr0
    res = np.linalg.norm(A @ xh - bh)
    print(f"  |r0|/|b| = {nr0 / nb:.2e}; FP64: {it} iterations, |Ax - b|/|b| = {res / nb:.2e}, "
          f"|Ax - b|/|r0| = {res / nr0:.2e}", flush=True)
    assert 0 < it < 10000
    assert res <= bar, (res, bar)
    xm = krylov_solve(prob, mat, b, x0, dict(precision="mixed")).cpu().numpy()
    info = prob.solver_info[-1]
    resm = np.linalg.norm(A @ xm - bh)
    err = np.linalg.norm(xm - xh) / np.linalg.norm(xh)
    print(f"  mixed: {info['iterations']} iterations in {info['outer']} passes, |Ax - b|/|r0| = {resm / nr0:.2e}, "
          f"|x - x64|/|x64| = {err:.2e}", flush=True)
    assert resm <= bar, (resm, bar)
    assert err < 1e-6, err


def test_all_slots_form_reproduces_one_slot_form_bitwise():
    """At 16^3 the all-slots form (0) and the one-slot-at-a-time form (1) run the same grid of
    64-row tiles with the same products in the same order: forced on the same system, they give
    bitwise-equal solutions and iteration counts (FP64 and mixed); the 128-row form (2) solves
    the same system to the same bar.  Forced on a larger system (28^3), form 0's blocks walk
    runs of tiles and meet the bar too."""
    import cpfem_amd as C
    from cpfem_amd import _native as N
    from cpfem_amd.solver import krylov_solve
    prob, mat, b, x0 = _first_system(16)
    L, h = prob.ctx._lib, prob.ctx.h
    out = {}
    for form in (0, 1, 2):
        N.check(L.cpfem_set_spmv_form(h, form))
        assert _form(prob) == form
        for prec in ("fp64", "mixed"):
            x = krylov_solve(prob, mat, b, x0, dict(precision=prec))
            out[form, prec] = (x.cpu().numpy(), prob.solver_info[-1]["iterations"])
    N.check(L.cpfem_set_spmv_form(h, -1))
    for prec in ("fp64", "mixed"):
        np.testing.assert_array_equal(out[0, prec][0], out[1, prec][0])
        assert out[0, prec][1] == out[1, prec][1]
        x2, x1 = out[2, prec][0], out[1, prec][0]
        assert np.linalg.norm(x2 - x1) <= 1e-6 * np.linalg.norm(x1)
    # forced where the 64-row tiles exceed three per CU (28^3: 1,144 tiles), the all-slots
    # form's blocks walk runs of tiles; the solution meets the same true-residual bar as the
    # size-picked form and lands within 1e-6 of it
    big, bmat, bb, bx0 = _first_system(28)
    xa = krylov_solve(big, bmat, bb, bx0).cpu().numpy()
    N.check(big.ctx._lib.cpfem_set_spmv_form(big.ctx.h, 0))
    assert _form(big) == 0
    x0f = krylov_solve(big, bmat, bb, bx0).cpu().numpy()
    A = C.BlockEllMatrix(big.ctx, bmat).to_scipy()
    bh, x0h = bb.cpu().numpy(), bx0.cpu().numpy()
    bar = 1e-10 * np.linalg.norm(bh) + 1e-9 * np.linalg.norm(bh - A @ x0h)
    assert np.linalg.norm(A @ x0f - bh) <= bar
    assert np.linalg.norm(x0f - xa) <= 1e-6 * np.linalg.norm(xa)


@pytest.mark.parametrize("n", [4, 16, 28])
@pytest.mark.parametrize("flat", [3, 4])
def test_flat_form_solves_first_newton_system(n, flat):
    """The one-row-per-thread forms (3: row blocks in launch order, 4: XCD-local row ranges),
    forced: one partial sum per 256-row block, added by the consuming kernel, so their dot
    products run in a different order from the tile forms'.
    FP64 and mixed solutions meet the same true-residual bar as the size-picked form and land
    within 1e-6 of it.  Iteration counts are not compared: BiCGSTAB's count moves with the
    summation order by up to ~10% on these systems in FP64 (28^3: 353 against 319) and more in
    the mixed form (16^3: 223 against 192); the oracle comparison of the count against its own
    rounding spread is in test_gpu_parity.py."""
    import cpfem_amd as C
    from cpfem_amd import _native as N
    from cpfem_amd.solver import krylov_solve
    prob, mat, b, x0 = _first_system(n)
    A = C.BlockEllMatrix(prob.ctx, mat).to_scipy()
    bh, x0h = b.cpu().numpy(), x0.cpu().numpy()
    bar = 1e-10 * np.linalg.norm(bh) + 1e-9 * np.linalg.norm(bh - A @ x0h)
    for prec in ("fp64", "mixed"):
        N.check(prob.ctx._lib.cpfem_set_spmv_form(prob.ctx.h, -1))
        xa = krylov_solve(prob, mat, b, x0, dict(precision=prec)).cpu().numpy()
        ka = prob.solver_info[-1]["iterations"]
        N.check(prob.ctx._lib.cpfem_set_spmv_form(prob.ctx.h, flat))
        assert _form(prob) == flat
        xf = krylov_solve(prob, mat, b, x0, dict(precision=prec)).cpu().numpy()
        kf = prob.solver_info[-1]["iterations"]
        res = np.linalg.norm(A @ xf - bh)
        err = np.linalg.norm(xf - xa) / np.linalg.norm(xa)
        print(f"n={n} {prec}: picked form {ka} iterations, flat form {flat} {kf}, |Ax - b| {res:.2e} (bar {bar:.2e}), "
              f"|x - x_picked|/|x| {err:.2e}", flush=True)
        assert res <= bar, (prec, res, bar)
        assert err < 1e-6, (prec, err)
    N.check(prob.ctx._lib.cpfem_set_spmv_form(prob.ctx.h, -1))


@pytest.mark.parametrize("n", [4, 16, 28])
def test_flat_form_box_columns_bitwise(n):
    """On a box mesh's pattern the one-row-per-thread form computes the columns of the nodes
    off the box's faces (cpfem_box_numbering finds the numbering) instead of loading them; the
    columns are the same, so with computed columns on and off the solutions and iteration
    counts are bitwise equal, FP64 and mixed."""
    import ctypes
    from cpfem_amd import _native as N
    from cpfem_amd.solver import krylov_solve
    prob, mat, b, x0 = _first_system(n)
    L, h = prob.ctx._lib, prob.ctx.h
    nn = prob.ctx.n_dofs // 3
    nslot = np.empty(nn, dtype=np.int32)
    col = np.empty((27, nn), dtype=np.int32)
    N.check(L.cpfem_pattern(h, nslot.ctypes.data, col.ctypes.data))
    dims = np.zeros(3, dtype=np.int64)
    N.check(L.cpfem_box_numbering(nn, col.ctypes.data, nslot.ctypes.data, dims.ctypes.data))
    assert list(dims) == [n + 1] * 3
    N.check(L.cpfem_set_spmv_form(h, 3))
    out = {}
    for on in (1, 0):
        N.check(L.cpfem_set_box_columns(h, on))
        for prec in ("fp64", "mixed"):
            x = krylov_solve(prob, mat, b, x0, dict(precision=prec))
            out[on, prec] = (x.cpu().numpy(), prob.solver_info[-1]["iterations"])
    N.check(L.cpfem_set_box_columns(h, 1))
    N.check(L.cpfem_set_spmv_form(h, -1))
    for prec in ("fp64", "mixed"):
        print(f"n={n} {prec}: {out[1, prec][1]} iterations with computed columns, {out[0, prec][1]} with loaded",
              flush=True)
        np.testing.assert_array_equal(out[1, prec][0], out[0, prec][0])
        assert out[1, prec][1] == out[0, prec][1]
