"""The calibration loop (cpfem_amd.optimize.minimize_lbfgsb) on CPU objectives: the
reference drivers' logging, stopping rules and first-iteration behaviour of L-BFGS-B with
bounds (calibration_case1_singleCrystalCopper_GB.py:285-409)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _quad(target, scale=1.0):
    t = torch.as_tensor(target, dtype=torch.float64)
    return lambda x: scale * torch.sum((x - t) ** 2)


def test_converges_and_logs(tmp_path):
    from cpfem_amd.optimize import minimize_lbfgsb
    out = minimize_lbfgsb(_quad([1.0, 2.0, 0.5]), [1.4, 1.4, 1.4], ([0.4] * 3, [3.0] * 3), maxiter=50,
                          obj_stop=-1.0, log_dir=str(tmp_path), x_device="cpu")
    np.testing.assert_allclose(out.x, [1.0, 2.0, 0.5], atol=1e-6)
    alpha = np.loadtxt(tmp_path / "alpha.txt", ndmin=2)
    obj = np.loadtxt(tmp_path / "obj_val.txt", ndmin=1)
    assert len(alpha) == len(obj) == out.nfev == len(out.history)
    np.testing.assert_array_equal(alpha[0], [1.4, 1.4, 1.4])
    assert obj[0] == pytest.approx(0.16 + 0.36 + 0.81)
    assert np.all(np.diff(np.loadtxt(tmp_path / "iteration.txt", ndmin=1)) >= 0)


def test_first_step_to_the_box_corner_and_maxiter():
    """With every variable bounded, L-BFGS-B's first trial point is the projected steepest-
    descent (Cauchy) point at unit step: for a large gradient, the corner picked by its signs.
    The reference's stored alpha.txt shows the same first move (1.4 -> 0.4 / 3.0 per sign)."""
    from cpfem_amd.optimize import minimize_lbfgsb
    f = _quad([0.9, 2.5, 0.5, 0.6, 1.0, 0.7], scale=1e5)
    out = minimize_lbfgsb(f, [1.4] * 6, ([0.4] * 6, [3.0] * 6), maxiter=1, obj_stop=-1.0, x_device="cpu")
    np.testing.assert_array_equal(out.history[1][0], [0.4, 3.0, 0.4, 0.4, 0.4, 0.4])
    assert out.reason == "maxiter" and out.nit == 1


def test_obj_stop_ends_the_run():
    from cpfem_amd.optimize import minimize_lbfgsb
    out = minimize_lbfgsb(_quad([1.0, 1.0]), [1.4, 1.4], ([0.4] * 2, [3.0] * 2), maxiter=50, obj_stop=1.0,
                          x_device="cpu")
    assert out.reason == "obj_stop" and out.nfev == 1 and out.fun < 1.0
