"""Oracle vs the reference-held golden curves (single-element Cu tension, Ta compression,
calibration objective scalar).  These pin the local update (a6-a12), the FE kernel, the
row-elimination BCs, the Krylov solve and the global Newton (a13-a18) end to end."""
import os

import numpy as np
import pytest

from oracle import cp_local, fe
from conftest import GOLDEN


def _single_element(table, g0, mat_kw, disps, ts, line_search=False):
    P0 = cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, table)))
    mat = cp_local.Material(P0, self_block=False, **mat_kw)
    pts, cells = fe.box_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    R = cp_local.rot_mat_from_quat(np.array([[1.0, 0, 0, 0]]))
    prob = fe.FEProblem(pts, cells, mat, R, g0)
    F = fe.faces(pts, (1.0, 1.0, 1.0))
    # calibration_case1_singleCrystalCopper_GB.py:147-149 (corner x/y, bottom z, top z)
    bcs = [(F["corner"], 0, 0.0), (F["corner"], 1, 0.0), (F["bottom"], 2, 0.0), (F["top"], 2, None)]
    rec = []
    fe.run_load_steps(prob, bcs, disps, ts, solver_kw=dict(tol=1e-7, line_search_flag=line_search),
                      record=rec)
    return np.array([r["sig_cells"][0, 2, 2] for r in rec])


CU = dict(xm=0.1, gss_a=2.5, h=541.5, t_sat=109.8, r=1.0, C11=1.684e5, C12=1.214e5, C44=0.754e5)


def test_cu_tension_curve():
    # calibration/case1.py:57-64,100-102; driver :107,120-121 (20 steps of 1.25e-3, dt 0.125 s)
    ref = np.loadtxt(os.path.join(GOLDEN, "cu_tension_szz.txt"))[:20]
    szz = _single_element("slip_fcc12.txt", 60.8, CU, np.linspace(0, 0.025, 21), np.linspace(0, 2.5, 21))
    np.testing.assert_allclose(szz, ref, rtol=2e-8)


def test_cu_objective_scalar():
    # all six calibration scales = 1.4 (alpha.txt:1); q = r everywhere (case1.py:175 no-op)
    ref = np.loadtxt(os.path.join(GOLDEN, "cu_tension_szz.txt"))[:20]
    a = 1.4
    kw = dict(CU, xm=0.1 * a, gss_a=2.5 * a, h=541.5 * a, t_sat=109.8 * a, r=1.0 * a)
    szz = _single_element("slip_fcc12.txt", 60.8 * a, kw, np.linspace(0, 0.025, 21), np.linspace(0, 2.5, 21))
    idx = np.arange(0, 20, 4)
    obj = 1e6 * np.sum((szz[idx] - ref[idx]) ** 2) / np.sum(ref[idx] ** 2)
    expect = float(np.loadtxt(os.path.join(GOLDEN, "cu_objective.txt")))
    assert abs(obj - expect) / expect < 1e-8


def test_cu_objective_detects_q_quirk():
    """Planted mistake: the 'correct' block-of-3 q (self=1) must NOT reproduce the stored
    objective, proving that the fixture pins the q no-op quirk."""
    ref = np.loadtxt(os.path.join(GOLDEN, "cu_tension_szz.txt"))[:20]
    a = 1.4
    P0 = cp_local.schmid_tensors(np.loadtxt(os.path.join(GOLDEN, "slip_fcc12.txt")))
    mat = cp_local.Material(P0, xm=0.1 * a, gss_a=2.5 * a, h=541.5 * a, t_sat=109.8 * a, r=a,
                            C11=1.684e5, C12=1.214e5, C44=0.754e5, self_block=True)
    pts, cells = fe.box_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    prob = fe.FEProblem(pts, cells, mat, np.eye(3)[None], 60.8 * a)
    F = fe.faces(pts, (1.0, 1.0, 1.0))
    bcs = [(F["corner"], 0, 0.0), (F["corner"], 1, 0.0), (F["bottom"], 2, 0.0), (F["top"], 2, None)]
    rec = []
    fe.run_load_steps(prob, bcs, np.linspace(0, 0.025, 21), np.linspace(0, 2.5, 21),
                      solver_kw=dict(tol=1e-7), record=rec)
    szz = np.array([r["sig_cells"][0, 2, 2] for r in rec])
    idx = np.arange(0, 20, 4)
    obj = 1e6 * np.sum((szz[idx] - ref[idx]) ** 2) / np.sum(ref[idx] ** 2)
    expect = float(np.loadtxt(os.path.join(GOLDEN, "cu_objective.txt")))
    assert abs(obj - expect) / expect > 1e-4


def test_ta_compression_curve():
    # calibration/case2.py (BCC12, g0 67.4641, h 1959.132, a 200, gsat 7295.1754, m 1/45.2726);
    # driver: 40 steps of -2.5e-3, dt 0.25 s, line_search_flag=True
    ref = np.loadtxt(os.path.join(GOLDEN, "ta_compression_szz.txt"))
    kw = dict(xm=1.0 / 45.2726, gss_a=200.0, h=1959.1320, t_sat=7295.1754, r=1.0,
              C11=2.670e5, C12=1.610e5, C44=0.825e5)
    szz = _single_element("slip_bcc12.txt", 67.4641, kw, np.linspace(0, -0.1, 41), np.linspace(0, 10.0, 41),
                          line_search=True)
    np.testing.assert_allclose(szz, ref, rtol=1e-9)
