"""cpfem_solve_noalloc (the FP64 Jacobi-BiCGSTAB with its workspace in the library's fixed
device array) against cpfem_solve on the same systems, through the C ABI: two calls in a row,
the second on a smaller system (the array keeps the first call's contents), each giving the
same iteration count and bitwise the same solution as cpfem_solve (the same kernels, grids and
summation order); systems beyond the array's limit and the mixed form are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _newton_system(n, seed):
    """The row-eliminated Newton matrix of load step 1 of the n^3 FCC polycrystal workload at
    u = 0 and a seeded right-hand side."""
    from cpfem_amd import configs
    w = configs.config2(n=n)
    prob = configs.build_problem(w)
    prob.fes[0].update_Dirichlet_boundary_conditions(w.dirichlet_bc_info(w.disps[1]))
    prob.dt = w.ts[1] - w.ts[0]
    prob.set_params(prob.internal_vars)
    _, mat = prob.assemble(torch.zeros(prob.ctx.n_dofs, dtype=torch.float64, device="cuda"))
    b = torch.as_tensor(np.random.default_rng(seed).normal(size=prob.ctx.n_dofs), device="cuda")
    return prob, mat, b


def _solve(prob, mat, b, entry):
    from cpfem_amd import _native as N
    from cpfem_amd.problem import ptr, stream_handle
    L = prob.ctx._lib
    x = torch.zeros_like(b)
    info = N.CpfemSolveInfo()
    if entry == "noalloc":
        rc = L.cpfem_solve_noalloc(prob.ctx.h, stream_handle(), ptr(mat), ptr(b), ptr(x), 1e-10, 1e-10, 10000, 1,
                                   ctypes.byref(info))
    else:
        rc = L.cpfem_solve(prob.ctx.h, stream_handle(), ptr(mat), ptr(b), ptr(x), 1e-10, 1e-10, 10000, 1, N.FP64,
                           ctypes.byref(info))
    N.check(rc)
    torch.cuda.synchronize()
    return x.cpu().numpy(), info


def test_noalloc_solve_matches_solve_on_two_calls():
    from cpfem_amd.problem import BlockEllMatrix
    for n, seed in ((12, 1), (6, 2)):  # the second call on a smaller system
        prob, mat, b = _newton_system(n, seed)
        print(f"[no_alloc] {n}^3: {prob.ctx.n_dofs} dofs", flush=True)
        x_na, inf_na = _solve(prob, mat, b, "noalloc")
        x_ref, inf_ref = _solve(prob, mat, b, "solve")
        assert inf_na.iterations == inf_ref.iterations > 10
        np.testing.assert_array_equal(x_na, x_ref)
        assert inf_na.residual == inf_ref.residual
        # and it solves the system: |A x - b| <= 1e-9 |b| on the exported CSR matrix
        A = BlockEllMatrix(prob.ctx, mat).to_scipy()
        bn = b.cpu().numpy()
        assert np.linalg.norm(A @ x_na - bn) <= 1e-9 * np.linalg.norm(bn)


def test_noalloc_refuses_systems_beyond_its_limit():
    from cpfem_amd import _native as N
    from cpfem_amd import configs
    from cpfem_amd.problem import ptr, stream_handle
    # 40^3 cells: 68,921 nodes = 206,763 dofs > CPFEM_NOALLOC_MAX_DOFS
    prob = configs.build_problem(configs.config2(n=40))
    assert prob.ctx.n_dofs > N.NOALLOC_MAX_DOFS
    b = torch.ones(prob.ctx.n_dofs, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    info = N.CpfemSolveInfo()
    rc = prob.ctx._lib.cpfem_solve_noalloc(prob.ctx.h, stream_handle(), ptr(b), ptr(b), ptr(x), 1e-10, 1e-10, 10, 1,
                                           ctypes.byref(info))
    assert rc == N.EBADARG
    assert "fixed workspace" in prob.ctx._lib.cpfem_last_error().decode()
