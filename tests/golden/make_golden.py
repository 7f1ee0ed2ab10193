"""Extract the reference-held golden data for the crystal-plasticity path into small fixtures.

Reads text data files of the reference (slip tables and stress curves written by the
reference's own runs) from /root/reference and writes them, unchanged in value, under
tests/golden/.  No reference code is imported or executed.

Sources:
  slip_fcc12.txt   calibration/data/csv/input_slip_sys.txt          (FCC {111}<110>)
  slip_bcc12.txt   calibration/data/csv/input_slip_sys_bcc12.txt    (BCC {110}<111>)
  slip_bcc24.txt   polycrystal_DPsteel/data/csv/input_slip_sys_bcc24.txt
  cu_tension_szz.txt      calibration/data/csv/calibration_case1/stress_strain_curve_copper.txt
  ta_compression_szz.txt  calibration/data/csv/calibration_case2/ss_curve_e-2.txt
  cu_objective.txt        calibration/data/csv/calibration_case1/obj_val.txt (alpha.txt:1 = 1.4 x6)
"""
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = {
    "slip_fcc12.txt": "calibration/data/csv/input_slip_sys.txt",
    "slip_bcc12.txt": "calibration/data/csv/input_slip_sys_bcc12.txt",
    "slip_bcc24.txt": "polycrystal_DPsteel/data/csv/input_slip_sys_bcc24.txt",
    "cu_tension_szz.txt": "calibration/data/csv/calibration_case1/stress_strain_curve_copper.txt",
    "ta_compression_szz.txt": "calibration/data/csv/calibration_case2/ss_curve_e-2.txt",
    "cu_objective.txt": "calibration/data/csv/calibration_case1/obj_val.txt",
}


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are already committed")
    for dst, src in FILES.items():
        a = np.atleast_1d(np.loadtxt(os.path.join(REF, src)))
        fmt = "%d" if dst.startswith("slip") else "%.17g"
        np.savetxt(os.path.join(HERE, dst), a, fmt=fmt)
        print(dst, a.shape)


if __name__ == "__main__":
    main()
