"""Parking of the assembly point kernel's slow lanes changes the time only: on the converged
state of the dual-phase workload's first load step (8^3, 4,096 points, BCC24, a wide spread of
local Newton iteration counts), the residual, the block-ELL Jacobian and every point's
iteration count are bitwise the same with parking off and at several thresholds, with and
without the tangent, and points were in fact parked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_parking_is_bitwise_neutral():
    from cpfem_amd import configs
    w = configs.config5(n=8)
    prob = configs.build_problem(w)
    p0 = [p.clone() for p in prob.internal_vars]
    sol0 = [torch.zeros(prob.fes[0].num_total_nodes, 3, dtype=torch.float64, device="cuda")]
    _, sol, _ = configs.run_steps(prob, w, p0, sol0, 0, 1, {"precision": "fp64"})
    prob.set_params(p0)
    dofs = sol[0].reshape(-1).clone()
    # a state away from convergence too: halfway between the start and the solution
    half = 0.5 * dofs
    out = {}
    for lanes in (0, 4, 16, 63):
        prob.ctx.set_point_park(lanes)
        for name, d in (("converged", dofs), ("half", half)):
            prob.ctx.point_counts(reset=True)
            it = torch.empty(w.n_qp, dtype=torch.int32, device="cuda")
            res, mat = prob.assemble(d, iters=it)
            res2, _ = prob.assemble(d, want_matrix=False)
            counts = prob.ctx.point_counts(reset=True)
            out[lanes, name] = (res.cpu().numpy(), mat.cpu().numpy().copy(), it.cpu().numpy(), res2.cpu().numpy(),
                                counts)
    prob.ctx.set_point_park(0)
    for name in ("converged", "half"):
        r0, m0, i0, s0, c0 = out[0, name]
        assert c0[2] == 0 and i0.max() > i0.min()
        for lanes in (4, 16, 63):
            r, m, i, s, c = out[lanes, name]
            np.testing.assert_array_equal(r, r0)
            np.testing.assert_array_equal(m, m0)
            np.testing.assert_array_equal(i, i0)
            np.testing.assert_array_equal(s, s0)
            assert c[0] == c0[0]  # the same total of Newton iterations
        assert out[63, name][4][2] > 0  # some points were parked
