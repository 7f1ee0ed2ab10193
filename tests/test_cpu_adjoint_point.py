"""The per-point reverse-mode product of the adjoint kernel (point_vjp in cp_dual.hpp, built
for the CPU with g++) against central finite differences of the oracle's local update, for
random output cotangents: xbar_k = d/dx_k [Pbar.P + sbar.sigma + Fbar.Fp_inv_new + gbar.g_new
+ slipbar.slip_new] for every input group (u_grad, Fp_inv_old, g_old, the eight material
parameters, R)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from helpers import FCC_304, random_states, slip_table

SRC = os.path.join(ROOT, "tests", "cpu_point", "point_check.cpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("pcv") / "libpoint_check.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out, SRC], check=True)
    return ctypes.CDLL(out)


def _oracle_f(x, table, N, dt, obar):
    from oracle import cp_local
    H, Fp, g, slip, R, mat = x
    Q = H.shape[0]
    Cs = np.stack([cp_local.cubic_stiffness(*m[5:8]) for m in mat])
    om = cp_local.Material(cp_local.schmid_tensors(slip_table(table)), xm=mat[:, 0], gss_a=mat[:, 1], h=mat[:, 2],
                           t_sat=mat[:, 3], r=mat[:, 4], C=Cs, self_block=True)
    o = cp_local.local_update(H, Fp, g, slip, R, om, dt, tol=1e-13, want_tangent=False, max_iter=400)
    Pb, sb, Fb, gb, slb = obar
    return (np.einsum("qij,qij->q", Pb, o["P"]) + np.einsum("qij,qij->q", sb, o["sigma"])
            + np.einsum("qij,qij->q", Fb, o["Fp_inv_new"]) + np.einsum("qa,qa->q", gb, o["g_new"])
            + np.einsum("qa,qa->q", slb, o["slip_new"]))


@pytest.mark.parametrize("table,N", [("slip_fcc12.txt", 12), ("slip_bcc24.txt", 24)])
def test_point_vjp_matches_fd(lib, table, N):
    from oracle import cp_local
    rng = np.random.default_rng(21)
    Q = 6
    H, Fp, g, slip, q = random_states(Q, N, rng, table=table, cap=1.3)
    R = cp_local.rot_mat_from_quat(q)
    kw = dict(FCC_304, xm=0.05)
    mat = np.tile([kw["xm"], kw["gss_a"], kw["h"], kw["t_sat"], kw["r"], kw["C11"], kw["C12"], kw["C44"]], (Q, 1))
    mat = mat * (1 + 0.05 * rng.random(mat.shape))
    obar = (rng.normal(size=(Q, 3, 3)), rng.normal(size=(Q, 3, 3)), 50 * rng.normal(size=(Q, 3, 3)),
            rng.normal(size=(Q, N)), 100 * rng.normal(size=(Q, N)))
    tab = slip_table(table)
    d = tab[:, 3:6] / np.linalg.norm(tab[:, 3:6], axis=1)[:, None]
    n = tab[:, 0:3] / np.linalg.norm(tab[:, 0:3], axis=1)[:, None]
    dt = 0.5
    outs = dict(H=np.zeros((Q, 9)), F=np.zeros((Q, 9)), g=np.zeros((Q, N)), p=np.zeros((Q, 8)), R=np.zeros((Q, 9)))
    c = lambda a: np.ascontiguousarray(a, dtype=float).ctypes.data_as(ctypes.c_void_p)
    keep = [np.ascontiguousarray(a, dtype=float) for a in (H.reshape(Q, 9), Fp.reshape(Q, 9), g, R.reshape(Q, 9), mat, d, n,
                                                           obar[0].reshape(Q, 9), obar[1].reshape(Q, 9),
                                                           obar[2].reshape(Q, 9), obar[3], obar[4])]
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.cpu_point_vjp(N, ctypes.c_int64(Q), *[p(a) for a in keep[:7]], ctypes.c_double(1e-3), ctypes.c_double(dt),
                           ctypes.c_double(1e-13), 5, 400, 1, *[p(a) for a in keep[7:]], p(outs["H"]), p(outs["F"]),
                           p(outs["g"]), p(outs["p"]), p(outs["R"]))
    assert rc == 0
    base = [H, Fp, g, slip, R, mat]
    groups = [("H", 0, 9, 1e-3), ("F", 1, 9, 1.0), ("g", 2, N, 100.0), ("p", 5, 8, None), ("R", 4, 9, 1.0)]
    for name, idx, width, scale in groups:
        fd = np.zeros((Q, width))
        for k in range(width):
            h = np.zeros((Q, width))
            if name == "p":  # the 1/m = 20 power law amplifies FD round-off: use a larger step
                h[:, k] = 1e-5 * mat[:, k]
            else:
                h[:, k] = 1e-6 * scale
            xp = [b.copy() for b in base]
            xm = [b.copy() for b in base]
            xp[idx] = (base[idx].reshape(Q, width) + h).reshape(base[idx].shape)
            xm[idx] = (base[idx].reshape(Q, width) - h).reshape(base[idx].shape)
            fd[:, k] = (_oracle_f(xp, table, N, dt, obar) - _oracle_f(xm, table, N, dt, obar)) / (2 * h[:, k])
        got = outs[name]
        # error per input component, relative to that component's largest sensitivity in the
        # batch (nearly elastic points have FD-noise-level sensitivities to the flow rule)
        err = np.abs(got - fd) / np.maximum(np.abs(fd).max(axis=0), 1e-30)
        assert err.max() < (5e-5 if name == "p" else 2e-5), (name, err.max(axis=0))
