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
  cu_calib_alpha.txt      calibration/data/csv/calibration_case1/alpha.txt (the L-BFGS-B points the
                          calibration run evaluated; no objective is stored for the second)

Three reference runs, from their own files (the Neper mesh each driver reads, gzip-compressed
with content unchanged; its quaternion table; and the driver's VTU output):
  steel304_*   polycrystal_304steel.py: data/neper/polycrystal_304steel/domain0_mesh16.msh,
               data/csv/polycrystal_304steel/quat.txt, data/vtk/polycrystal_304steel/u_000..u_015.vtu
  ta_*         singlecrystal_tantalum.py: data/neper/singlecrystal_tantalum/mesh10.msh,
               data/vtk/singlecrystal_tantalum/u_000..u_049.vtu (identity orientation, no table).
               The driver text names mesh16.msh, but the stored VTUs have 1000 cells and the
               connectivity of mesh10.msh, so that is the mesh of the run they record.
  dp_*         polycrystal_DPsteel_inhomo.py: data/neper/polycrystal_DPsteel/n1-id1-mesh10.msh,
               data/csv/polycrystal_DPsteel/quat.txt,
               data/vtk/polycrystal_DPsteel/u_inhomo_000..u_inhomo_049.vtu
Each <case>_vtu.npz holds per-step cell means of sigma_xx, sigma_yy, sigma_zz and von Mises
(float64 means of the float32 cell data), the full float32 cell fields of four steps, the
nodal displacement "sol" of the same four steps, the per-cell orientation (and, for dp,
phase indices and the cubic constants C11, C12, C44) the run used (the dual-phase driver draws
both at random without a seed; the VTU records the draw), and the connectivity (to cross-check
mesh ingest); dp_vtu_skeleton.txt lists the element tags and attributes of
u_inhomo_000.vtu (no payload), the layout cpfem_amd.utils.save_sol reproduces.  VTU arrays are decoded with xml.etree + base64 + zlib (no code runs).
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
    "cu_calib_alpha.txt": "calibration/data/csv/calibration_case1/alpha.txt",
}


CASES = {
    "steel304": dict(mesh="polycrystal_304steel/data/neper/polycrystal_304steel/domain0_mesh16.msh",
                     quat="polycrystal_304steel/data/csv/polycrystal_304steel/quat.txt",
                     vtu="polycrystal_304steel/data/vtk/polycrystal_304steel/u_{:03d}.vtu", steps=16,
                     full=(0, 4, 9, 15), cell_fields=("cell_ori_inds",)),
    "ta": dict(mesh="singlecrystal_tantalum/data/neper/singlecrystal_tantalum/mesh10.msh", quat=None,
               vtu="singlecrystal_tantalum/data/vtk/singlecrystal_tantalum/u_{:03d}.vtu", steps=50,
               full=(0, 9, 24, 49), cell_fields=("cell_ori_inds",)),
    "dp": dict(mesh="polycrystal_DPsteel/data/neper/polycrystal_DPsteel/n1-id1-mesh10.msh",
               quat="polycrystal_DPsteel/data/csv/polycrystal_DPsteel/quat.txt",
               vtu="polycrystal_DPsteel/data/vtk/polycrystal_DPsteel/u_inhomo_{:03d}.vtu", steps=50,
               full=(0, 9, 24, 49), cell_fields=("cell_ori_inds", "phase_inds", "C11", "C12", "C44")),
}


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


def vtu_case(name, mesh, quat, vtu, steps, full, cell_fields):
    with open(os.path.join(REF, mesh), "rb") as f, open(os.path.join(HERE, f"{name}_mesh.msh.gz"), "wb") as raw, \
            gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as g:
        g.write(f.read())
    if quat is not None:
        np.savetxt(os.path.join(HERE, f"{name}_quat.txt"), np.loadtxt(os.path.join(REF, quat)), fmt="%.17g")
    means, fields, sols, extra = [], [], [], {}
    for i in range(steps):
        a = vtu_arrays(os.path.join(REF, vtu.format(i)))
        cols = [a["sigma_xx"], a["sigma_yy"], a["sigma_zz"], a["von_Mises_stress"]]
        means.append([c.astype(np.float64).mean() for c in cols])
        if i in full:
            fields.append(np.stack(cols[:3], axis=1))
            sols.append(a["sol"].reshape(-1, 3))
        if i == 0:
            for k in cell_fields:
                extra[k] = a[k].astype(np.int64 if k.endswith("_inds") else np.float64)
            extra["connectivity"] = a["connectivity"].reshape(-1, 8).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, f"{name}_vtu.npz"), mean_sigma=np.array(means), full_steps=np.array(full),
                        full_sigma=np.array(fields, dtype=np.float32), full_sol=np.array(sols), **extra)
    print(f"{name}: mean szz", np.array(means)[:, 2])


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are already committed")
    for dst, src in FILES.items():
        a = np.atleast_1d(np.loadtxt(os.path.join(REF, src)))
        fmt = "%d" if dst.startswith("slip") else "%.17g"
        np.savetxt(os.path.join(HERE, dst), a, fmt=fmt)
        print(dst, a.shape)
    for name, case in CASES.items():
        vtu_case(name, **case)
    vtu_skeleton(CASES["dp"]["vtu"].format(0), os.path.join(HERE, "dp_vtu_skeleton.txt"))


def vtu_skeleton(src, dst):
    """Element tags and attributes of one reference VTU, in document order (array payloads
    dropped): the layout a compatible writer must reproduce."""
    lines = []
    for el in ET.parse(os.path.join(REF, src)).getroot().iter():
        lines.append(el.tag + " " + " ".join(f'{k}="{v}"' for k, v in el.attrib.items()))
    with open(dst, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(dst, len(lines), "elements")


if __name__ == "__main__":
    main()
