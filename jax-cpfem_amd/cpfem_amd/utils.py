"""Post-processing outputs of the example drivers: per-cell stress components, von Mises
stress and a VTU writer whose files have the layout of the reference's outputs.

The drivers (singlecrystal_copper.py:203-233, polycrystal_304steel.py:230-260,
polycrystal_DPsteel_inhomo.py:229-260) take compute_avg_stress per cell, form sigma_xx/yy/zz
and the von Mises stress, and call save_sol(fe, sol, path, cell_infos=[(name, per-cell
array), ...]).  save_sol here writes the same VTK XML UnstructuredGrid the reference's files
hold: zlib-compressed base64 binary arrays with a UInt32 block header, Points as Float64,
connectivity and offsets as Int32, cell types as Int64 (12 = hexahedron), the nodal solution
as a Float32 "sol" with 3 components, and every cell field as Float32, so ParaView state files
and scripts written for the reference's outputs read these unchanged.
"""
import base64
import os
import struct
import zlib

import numpy as np

VTK_HEXAHEDRON = 12
_VTK_TYPE = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int32): "Int32",
             np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}
_BLOCK = 32768  # uncompressed bytes per zlib block


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def cell_stress_fields(sigma):
    """(cells, 3, 3) Cauchy stress -> dict of the per-cell fields the drivers write:
    sigma_xx, sigma_yy, sigma_zz and von_Mises_stress
    (sqrt(((sxx-syy)^2 + (syy-szz)^2 + (szz-sxx)^2)/2 + 3 (sxy^2 + syz^2 + sxz^2)),
    singlecrystal_copper.py:206-214)."""
    s = _host(sigma).astype(np.float64)
    xx, yy, zz = s[:, 0, 0], s[:, 1, 1], s[:, 2, 2]
    xy, yz, xz = s[:, 0, 1], s[:, 1, 2], s[:, 0, 2]
    vm = np.sqrt(0.5 * ((xx - yy) ** 2 + (yy - zz) ** 2 + (zz - xx) ** 2) + 3.0 * (xy ** 2 + yz ** 2 + xz ** 2))
    return {"sigma_xx": xx, "sigma_yy": yy, "sigma_zz": zz, "von_Mises_stress": vm}


def _encode(arr):
    """base64 of the UInt32 block header (blocks, block size, size of the last block,
    compressed size of each block) followed by base64 of the concatenated zlib blocks."""
    raw = np.ascontiguousarray(arr).tobytes()
    chunks = [raw[i:i + _BLOCK] for i in range(0, len(raw), _BLOCK)] or [b""]
    comp = [zlib.compress(c) for c in chunks]
    last = len(chunks[-1]) if len(raw) % _BLOCK else (_BLOCK if raw else 0)
    header = struct.pack(f"<{3 + len(comp)}I", len(comp), _BLOCK, last, *[len(c) for c in comp])
    return base64.b64encode(header).decode() + base64.b64encode(b"".join(comp)).decode()


def _data_array(name, arr, ncomp=None):
    arr = np.asarray(arr)
    attrs = f'type="{_VTK_TYPE[arr.dtype]}" Name="{name}"'
    if ncomp is not None:
        attrs += f' NumberOfComponents="{ncomp}"'
    return f'<DataArray {attrs} format="binary">\n{_encode(arr)}\n</DataArray>\n'


def write_vtu(path, points, cells, point_data=(), cell_data=()):
    """Hexahedral unstructured grid to ``path``.  ``point_data`` / ``cell_data`` are sequences
    of (name, array) with arrays of shape (n,) or (n, k); arrays keep their dtype."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    cells = np.asarray(cells).reshape(len(cells), -1)
    nc, npe = cells.shape
    out = ['<?xml version="1.0"?>\n',
           '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
           'compressor="vtkZLibDataCompressor">\n',
           '<UnstructuredGrid>\n', f'<Piece NumberOfPoints="{len(points)}" NumberOfCells="{nc}">\n',
           '<Points>\n', _data_array("Points", points, 3), '</Points>\n', '<Cells>\n',
           _data_array("connectivity", cells.astype(np.int32).ravel()),
           _data_array("offsets", (npe * np.arange(1, nc + 1)).astype(np.int32)),
           _data_array("types", np.full(nc, VTK_HEXAHEDRON, dtype=np.int64)), '</Cells>\n']
    for tag, items, n in (("PointData", point_data, len(points)), ("CellData", cell_data, nc)):
        if not items:
            continue
        out.append(f"<{tag}>\n")
        for name, arr in items:
            arr = np.asarray(arr)
            if arr.shape[0] != n:
                raise ValueError(f"{tag} '{name}' has {arr.shape[0]} entries, expected {n}")
            out.append(_data_array(name, arr.reshape(n, -1).ravel() if arr.ndim > 1 else arr,
                                   arr.shape[1] if arr.ndim > 1 else None))
        out.append(f"</{tag}>\n")
    out += ['</Piece>\n', '</UnstructuredGrid>\n', '</VTKFile>\n']
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("".join(out))


def save_sol(fe, sol, sol_file, cell_infos=None, point_infos=None):
    """The drivers' VTU output call: nodal solution ``sol`` (num_total_nodes, 3) as Float32
    point data "sol", each (name, per-cell array) of ``cell_infos`` as Float32 cell data, each
    (name, per-node array) of ``point_infos`` as Float32 point data; the mesh is fe.points /
    fe.cells (hex8)."""
    n, nc = len(fe.points), len(fe.cells)
    pd = [("sol", _host(sol).reshape(n, -1).astype(np.float32))]
    for name, data in point_infos or ():
        pd.append((name, _host(data).astype(np.float32)))
    cd = []
    for name, data in cell_infos or ():
        data = _host(data).astype(np.float32)
        if data.shape[0] != nc:
            raise ValueError(f"cell field '{name}' has shape {data.shape}, expected ({nc},)")
        cd.append((name, data))
    write_vtu(sol_file, fe.points, fe.cells, pd, cd)
