"""Material presets of the reference model files and the per-point parameter layouts.

Units as in the reference: MPa, mm, s.  Each preset names the reference file whose
constants it carries.
"""
import dataclasses
import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
SLIP_TABLES = {"fcc12": "slip_fcc12.txt", "bcc12": "slip_bcc12.txt", "bcc24": "slip_bcc24.txt"}


def load_slip_table(name_or_path):
    """(N,6) table, columns 0:3 plane normal, 3:6 slip direction (input_slip_sys*.txt)."""
    path = os.path.join(DATA, SLIP_TABLES[name_or_path]) if name_or_path in SLIP_TABLES else name_or_path
    return np.loadtxt(path)


def slip_vectors(table):
    """Normalised directions and normals, models_copper.py:63-66."""
    t = np.asarray(table, dtype=np.float64)
    d = t[:, 3:6] / np.linalg.norm(t[:, 3:6], axis=1)[:, None]
    n = t[:, 0:3] / np.linalg.norm(t[:, 0:3], axis=1)[:, None]
    return np.ascontiguousarray(d), np.ascontiguousarray(n)


@dataclasses.dataclass
class Material:
    """Kalidindi model constants.

    layout: "4"  -> internal_vars [Fp_inv, g, slip, R]                     (models_copper.py:133)
            "9"  -> + per-point [gss_a, h, t_sat, xm, r]                   (calibration/case1.py:139)
            "10" -> + per-point stiffness C_gp (cells, 8, 3, 3, 3, 3)     (models_DPsteel_inhomo.py:229)
    self_block: latent matrix q = 1 inside blocks of 3 (True) or q = r everywhere (False: the
            discarded ``q.at[].set`` of the 9-array files, calibration/case1.py:175).
    """
    slip: str = "fcc12"
    g0: float = 60.8
    h: float = 541.5
    t_sat: float = 109.8
    gss_a: float = 2.5
    xm: float = 0.1
    r: float = 1.0
    C11: float = 1.684e5
    C12: float = 1.214e5
    C44: float = 0.754e5
    ao: float = 1e-3
    tol: float = 1e-8
    max_sub_step: int = 5
    layout: str = "4"
    self_block: bool = True
    max_iter: int = 250

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


# singlecrystal_copper/models_copper.py:54-149
COPPER = Material()
# singlecrystal_tantalum/models_tantalum.py:54-151
TANTALUM = Material(slip="bcc12", g0=67.4641, h=1959.1320, t_sat=7295.1754, gss_a=200.0, xm=1.0 / 45.2726,
                    C11=2.670e5, C12=1.610e5, C44=0.825e5)
# polycrystal_304steel/models_304steel.py:54-151, max_sub_step 8 (:232)
STEEL304 = Material(g0=90.0, h=392.9772, t_sat=7295.1754, gss_a=8.0, xm=1.0 / 120.0,
                    C11=2.622e5, C12=1.120e5, C44=0.746e5, max_sub_step=8)
# calibration/case1.py:57-64,100-102 (9-array layout, q no-op)
CALIB_COPPER = COPPER.replace(layout="9", self_block=False)
# calibration/case2.py:53-60,104-106
CALIB_TANTALUM = TANTALUM.replace(layout="9", self_block=False)
# calibration/case4.py:53-60,93-95, max_sub_step 8 (:219)
CALIB_STEEL304 = STEEL304.replace(layout="9", self_block=False)
# crystal_plasticity_OR_design/models_sens_rot.py:53-60,100-102; tol 1e-6 (:211), max_sub_step 10 (:230)
SENS_ROT = COPPER.replace(layout="9", self_block=False, tol=1e-6, max_sub_step=10)

# polycrystal_DPsteel/models_DPsteel_inhomo.py:76-102 (ferrite = phase 0, martensite = phase 1;
# BCC24 with the self block of :262-266, tol 1e-8 :302, max_sub_step 5 :321); used as
# custom_init(quat, cell_ori_inds, phase_inds, (DP_FERRITE, DP_MARTENSITE))
DP_FERRITE = Material(slip="bcc24", g0=170.0, h=400.0, t_sat=2500.0, gss_a=4.0, xm=0.05, r=1.0,
                      C11=2.314e5, C12=1.347e5, C44=1.164e5, layout="10")
DP_MARTENSITE = DP_FERRITE.replace(g0=435.0, h=950.0, t_sat=5300.0, C11=4.174e5, C12=2.424e5, C44=2.111e5)

PRESETS = dict(copper=COPPER, tantalum=TANTALUM, steel304=STEEL304, calib_copper=CALIB_COPPER,
               calib_tantalum=CALIB_TANTALUM, calib_steel304=CALIB_STEEL304, sens_rot=SENS_ROT,
               dp_ferrite=DP_FERRITE, dp_martensite=DP_MARTENSITE)


def cubic_stiffness(C11, C12, C44):
    """Full C_ijkl of a cubic crystal (models_copper.py:92-130)."""
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        C[i, i, i, i] = C11
    for i, j in ((0, 1), (0, 2), (1, 2)):
        C[i, i, j, j] = C[j, j, i, i] = C12
        C[i, j, i, j] = C[i, j, j, i] = C[j, i, i, j] = C[j, i, j, i] = C44
    return C
