"""VTU output (cpfem_amd.utils): the layout of the reference's own output files, exact
round trips of every array through an independent decoder, multi-block payloads, and the
post-processing fields of the drivers."""
import base64
import os
import struct
import types
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

from conftest import GOLDEN

pytest.importorskip("torch")

_NP = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32, "UInt8": np.uint8}


def _decode(path):
    """name -> array of a zlib-compressed binary VTU (UInt32 block header), decoded
    independently of the writer."""
    out = {}
    for da in ET.parse(path).getroot().iter("DataArray"):
        txt = "".join(da.text.split())
        nb = struct.unpack("<I", base64.b64decode(txt[:8])[:4])[0]
        hlen = 4 * ((4 * (3 + nb) + 2) // 3)
        hdr = struct.unpack(f"<{3 + nb}I", base64.b64decode(txt[:hlen])[:4 * (3 + nb)])
        payload = base64.b64decode(txt[hlen:])
        raw, o = b"", 0
        for c in hdr[3:]:
            raw += zlib.decompress(payload[o:o + c])
            o += c
        assert len(raw) == (nb - 1) * hdr[1] + hdr[2]
        out[da.attrib["Name"]] = np.frombuffer(raw, dtype=_NP[da.attrib["type"]])
    return out


def _skeleton(path):
    return [el.tag + " " + " ".join(f'{k}="{v}"' for k, v in el.attrib.items())
            for el in ET.parse(path).getroot().iter()]


def _dp_fe():
    from cpfem_amd import read_gmsh
    m = read_gmsh(os.path.join(GOLDEN, "dp_mesh.msh.gz"))
    return types.SimpleNamespace(points=m.points, cells=m.cells_dict["hexahedron"])


def test_layout_matches_reference_output(tmp_path):
    """The dual-phase driver's save_sol call (polycrystal_DPsteel_inhomo.py:259-260), fed the
    fields its stored u_inhomo_000.vtu holds, gives that file's element/attribute layout, and
    every array decodes to exactly the stored float32 values."""
    from cpfem_amd.utils import save_sol
    ref = np.load(os.path.join(GOLDEN, "dp_vtu.npz"))
    fe = _dp_fe()
    sig = ref["full_sigma"][0]
    cell_infos = [("cell_ori_inds", ref["cell_ori_inds"]), ("sigma_xx", sig[:, 0]), ("sigma_yy", sig[:, 1]),
                  ("sigma_zz", sig[:, 2]), ("von_Mises_stress", np.zeros(len(sig))),
                  ("phase_inds", ref["phase_inds"]), ("C11", ref["C11"]), ("C12", ref["C12"]), ("C44", ref["C44"])]
    path = tmp_path / "u_inhomo_000.vtu"
    save_sol(fe, ref["full_sol"][0], str(path), cell_infos=cell_infos)
    want = open(os.path.join(GOLDEN, "dp_vtu_skeleton.txt")).read().splitlines()
    assert _skeleton(path) == want
    a = _decode(path)
    np.testing.assert_array_equal(a["Points"].reshape(-1, 3), fe.points)
    np.testing.assert_array_equal(a["connectivity"].reshape(-1, 8), ref["connectivity"])
    np.testing.assert_array_equal(a["offsets"], 8 * np.arange(1, 1001))
    assert np.all(a["types"] == 12)
    np.testing.assert_array_equal(a["sol"].reshape(-1, 3), ref["full_sol"][0])
    for k in range(3):
        np.testing.assert_array_equal(a[("sigma_xx", "sigma_yy", "sigma_zz")[k]], sig[:, k])
    np.testing.assert_array_equal(a["phase_inds"], ref["phase_inds"].astype(np.float32))
    np.testing.assert_array_equal(a["C11"], ref["C11"].astype(np.float32))


def test_multiblock_payload_and_errors(tmp_path):
    """Arrays longer than one 32 KiB zlib block (a 16^3 mesh: 117 KB of points), a block that
    ends exactly at the boundary, and a cell field of the wrong length."""
    from cpfem_amd import box_mesh
    from cpfem_amd.utils import save_sol, write_vtu
    m = box_mesh(16, 16, 16, 1.0, 1.0, 1.0)
    fe = types.SimpleNamespace(points=m.points, cells=m.cells_dict["hexahedron"])
    rng = np.random.default_rng(0)
    u = rng.standard_normal((len(fe.points), 3))
    c = rng.standard_normal(len(fe.cells))
    path = str(tmp_path / "u.vtu")
    save_sol(fe, u, path, cell_infos=[("f", c)], point_infos=[("p", u[:, 0])])
    a = _decode(path)
    np.testing.assert_array_equal(a["Points"].reshape(-1, 3), fe.points)
    np.testing.assert_array_equal(a["sol"].reshape(-1, 3), u.astype(np.float32))
    np.testing.assert_array_equal(a["f"], c.astype(np.float32))
    np.testing.assert_array_equal(a["p"], u[:, 0].astype(np.float32))
    exact = np.arange(2 * 32768 // 8, dtype=np.float64)  # two full blocks, no partial one
    write_vtu(path, np.zeros((len(exact), 3)), np.zeros((1, 8), dtype=np.int32), point_data=[("e", exact)])
    np.testing.assert_array_equal(_decode(path)["e"], exact)
    with pytest.raises(ValueError):
        save_sol(fe, u, path, cell_infos=[("bad", c[:-1])])


def test_cell_stress_fields_von_mises():
    """Uniaxial sigma_zz = s gives von Mises s; pure shear tau gives sqrt(3) tau; a pressure
    gives 0."""
    from cpfem_amd.utils import cell_stress_fields
    s = np.zeros((3, 3, 3))
    s[0, 2, 2] = 250.0
    s[1, 0, 1] = s[1, 1, 0] = 10.0
    s[2] = -7.0 * np.eye(3)
    f = cell_stress_fields(s)
    np.testing.assert_allclose(f["von_Mises_stress"], [250.0, np.sqrt(3.0) * 10.0, 0.0], atol=1e-12)
    np.testing.assert_array_equal(f["sigma_zz"], [250.0, 0.0, -7.0])
