"""Gradient-based calibration / design loop on top of the adjoint: scipy's L-BFGS-B driving a
torch objective whose gradient comes from the hand-written reverse mode.

The reference's calibration drivers (calibration_case1_singleCrystalCopper_GB.py:285-409,
calibration_case2_singleCrystalTa_GB.py, calibration_case4_UQ_polyCrystalSteel_*_GB.py) wrap
``jax.value_and_grad(simulation)`` for ``scipy.optimize.minimize(..., method='L-BFGS-B',
jac=..., bounds=..., callback=...)`` with this behaviour, which ``minimize_lbfgsb`` keeps:

* one forward + backward per evaluation; the gradient is cached for the jac call of the
  same point (objective_wrapper.dJ, :307-308, :349);
* before each evaluation the point is appended to alpha.txt (space-separated, :293-299);
  after it the objective to obj_val.txt and the elapsed seconds to iteration.txt (:313-326);
* the run ends as soon as an objective falls below ``obj_stop`` (1.0 in the drivers, :330-334,
  there by SystemExit) or when the callback has counted ``maxiter`` iterations (10, :357-372,
  there by StopIteration);
* the drivers reset the problem before every evaluation (zero initial guess, custom_init,
  :302-304); ``simulation`` is expected to do the same, as in the drivers.

Instead of ending the process, the loop returns an ``OptimizeOutcome`` naming why it stopped.
"""
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from scipy.optimize import Bounds, minimize

from .problem import device


class _Stop(Exception):
    pass


@dataclass
class OptimizeOutcome:
    x: np.ndarray                 # last evaluated point
    fun: float                    # objective there
    jac: np.ndarray               # gradient there
    nit: int                      # completed L-BFGS-B iterations
    nfev: int                     # objective evaluations
    reason: str                   # "obj_stop", "maxiter", or scipy's message
    history: list = field(default_factory=list)  # (x, f) per evaluation
    scipy_result: object = None


def _append(path, text):
    with open(path, "a+") as f:
        f.write(text + "\n")


def minimize_lbfgsb(simulation, x0, bounds, maxiter=10, obj_stop=1.0, log_dir=None, scipy_options=None,
                    x_device=None):
    """Minimise ``simulation(x)`` (x: float64 torch tensor with requires_grad, on the current
    GPU unless ``x_device`` says otherwise; returns a scalar tensor) with L-BFGS-B, gradient by
    ``backward()``.

    bounds: scipy ``Bounds`` or (lower, upper) sequences.  log_dir: where alpha.txt,
    obj_val.txt and iteration.txt are written (truncated at the start), or None."""
    x0 = np.asarray(x0, dtype=np.float64)
    if not isinstance(bounds, Bounds):
        bounds = Bounds(np.asarray(bounds[0], dtype=np.float64), np.asarray(bounds[1], dtype=np.float64))
    files = None
    if log_dir is not None:
        os.makedirs(log_dir, exist_ok=True)
        files = {k: os.path.join(log_dir, k + ".txt") for k in ("alpha", "obj_val", "iteration")}
        for p in files.values():
            open(p, "w").close()
    t0 = time.time()
    state = dict(x=None, f=None, g=None, nit=0, nfev=0, reason=None, history=[])

    def value_and_grad(x):
        x = np.array(x, dtype=np.float64)
        if files:
            _append(files["alpha"], " ".join(map(str, x)))
        xt = torch.tensor(x, dtype=torch.float64, device=x_device or device(), requires_grad=True)
        obj = simulation(xt)
        obj.backward()
        f = float(obj.detach().cpu())
        g = xt.grad.detach().cpu().numpy().astype(np.float64)
        state.update(x=x, f=f, g=g, nfev=state["nfev"] + 1)
        state["history"].append((x.copy(), f))
        if files:
            _append(files["obj_val"], repr(f))
            _append(files["iteration"], repr(time.time() - t0))
        if f < obj_stop:
            state["reason"] = "obj_stop"
            raise _Stop
        return f, np.asfortranarray(g)

    def callback(xk):
        state["nit"] += 1
        if state["nit"] == maxiter:
            state["reason"] = "maxiter"
            raise _Stop

    res = None
    try:
        res = minimize(value_and_grad, x0, method="L-BFGS-B", jac=True, bounds=bounds, callback=callback,
                       options=dict(scipy_options or {}))
        state["reason"] = str(res.message)
    except (_Stop, StopIteration):
        pass
    return OptimizeOutcome(x=state["x"], fun=state["f"], jac=state["g"], nit=state["nit"], nfev=state["nfev"],
                           reason=state["reason"] or "stopped", history=state["history"], scipy_result=res)
