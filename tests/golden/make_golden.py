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

The 304-steel polycrystal run of the reference (polycrystal_304steel.py), from its own files:
  steel304_mesh16.msh.gz  polycrystal_304steel/data/neper/polycrystal_304steel/domain0_mesh16.msh
                          (the Neper mesh the driver reads, gzip-compressed, content unchanged)
  steel304_quat.txt       polycrystal_304steel/data/csv/polycrystal_304steel/quat.txt
  steel304_vtu.npz        from the driver's output polycrystal_304steel/data/vtk/polycrystal_304steel/
                          u_000..u_015.vtu (load steps 1-16): per-step cell means of sigma_xx,
                          sigma_yy, sigma_zz, von Mises (float64 means of the float32 cell data),
                          the full float32 cell fields of steps 1, 5, 10 and 16, the cell
                          orientation indices and the connectivity (to cross-check mesh ingest).
                          VTU arrays are decoded with xml.etree + base64 + zlib (no code runs).
"""
import base64
import gzip
import os
import struct
import sys
import xml.etree.ElementTree as ET
import zlib

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


STEEL = "polycrystal_304steel/data"
STEEL_FULL_STEPS = (0, 4, 9, 15)


def vtu_arrays(path):
    """name -> numpy array of every DataArray of a meshio-written zlib-compressed binary VTU
    (UInt32 block header: number of blocks, block size, last block size, compressed sizes)."""
    types = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32, "UInt8": np.uint8}
    out = {}
    for da in ET.parse(path).getroot().iter("DataArray"):
        txt = da.text.strip()
        nb = struct.unpack("<I", base64.b64decode(txt[:8])[:4])[0]
        hlen = 4 * ((4 * (3 + nb) + 2) // 3)
        hdr = np.frombuffer(base64.b64decode(txt[:hlen]), dtype=np.uint32)
        data, o, parts = base64.b64decode(txt[hlen:]), 0, []
        for c in hdr[3:3 + nb]:
            parts.append(zlib.decompress(data[o:o + int(c)]))
            o += int(c)
        out[da.attrib["Name"]] = np.frombuffer(b"".join(parts), dtype=types[da.attrib["type"]])
    return out


def steel304():
    src = os.path.join(REF, STEEL, "neper/polycrystal_304steel/domain0_mesh16.msh")
    with open(src, "rb") as f, open(os.path.join(HERE, "steel304_mesh16.msh.gz"), "wb") as raw, \
            gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as g:
        g.write(f.read())
    q = np.loadtxt(os.path.join(REF, STEEL, "csv/polycrystal_304steel/quat.txt"))
    np.savetxt(os.path.join(HERE, "steel304_quat.txt"), q, fmt="%.17g")
    means, full = [], []
    for i in range(16):
        a = vtu_arrays(os.path.join(REF, STEEL, f"vtk/polycrystal_304steel/u_{i:03d}.vtu"))
        cols = [a["sigma_xx"], a["sigma_yy"], a["sigma_zz"], a["von_Mises_stress"]]
        means.append([c.astype(np.float64).mean() for c in cols])
        if i in STEEL_FULL_STEPS:
            full.append(np.stack(cols[:3], axis=1))
        if i == 0:
            ori = a["cell_ori_inds"].astype(np.int64)
            conn = a["connectivity"].reshape(-1, 8).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "steel304_vtu.npz"), mean_sigma=np.array(means), full_steps=np.array(STEEL_FULL_STEPS),
                        full_sigma=np.array(full, dtype=np.float32), cell_ori_inds=ori, connectivity=conn)
    print("steel304: mean szz", np.array(means)[:, 2])


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are already committed")
    for dst, src in FILES.items():
        a = np.atleast_1d(np.loadtxt(os.path.join(REF, src)))
        fmt = "%d" if dst.startswith("slip") else "%.17g"
        np.savetxt(os.path.join(HERE, dst), a, fmt=fmt)
        print(dst, a.shape)
    steel304()


if __name__ == "__main__":
    main()
