"""Run-to-run reproducibility.  No sum in the library depends on timing: every dot product is
written as one partial per block and added by the consuming kernel in index order, assembly
gathers element blocks in a fixed order (no atomics in any sum), and the point kernels are
one lane per point.  Repeating a computation on the same inputs therefore gives bitwise-equal
results, in every SpMV form and in both precisions: checked on one Krylov solve per form
(28^3 first Newton system) and on two load steps of the dual-phase workload (16^3, line
search on), each run twice on fresh contexts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("form", [-1, 0, 1, 3, 4])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_krylov_solve_repeats_bitwise(form, precision):
    from test_gpu_spmv_forms import _first_system
    from cpfem_amd import _native as N
    from cpfem_amd.solver import krylov_solve
    prob, mat, b, x0 = _first_system(28)
    N.check(prob.ctx._lib.cpfem_set_spmv_form(prob.ctx.h, form))
    runs = []
    for _ in range(2):
        x = krylov_solve(prob, mat, b, x0, dict(precision=precision)).cpu().numpy()
        runs.append((x, prob.solver_info[-1]["iterations"]))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_load_steps_repeat_bitwise(precision):
    from cpfem_amd import configs
    outs = []
    for _ in range(2):
        w = configs.config5(n=16)
        prob = configs.build_problem(w)
        sol = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
        params, sol, sig = configs.run_steps(prob, w, prob.internal_vars, sol, 0, 2, {"precision": precision})
        outs.append((np.array(sig), sol[0].cpu().numpy(), [p.cpu().numpy() for p in params[:2]]))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    for a, b in zip(outs[0][2], outs[1][2]):
        np.testing.assert_array_equal(a, b)
