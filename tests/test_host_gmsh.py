"""Gmsh 2.2 ingest (cpfem_amd.read_gmsh) on the reference's Neper mesh and on a small file
written here: node renumbering, element filtering, physical tags, gzip input."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytest.importorskip("torch")


def test_reference_neper_mesh():
    from cpfem_amd import read_gmsh
    m = read_gmsh(os.path.join(GOLDEN, "steel304_mesh.msh.gz"))
    cells = m.cells_dict["hexahedron"]
    assert m.points.shape == (17 ** 3, 3) and cells.shape == (16 ** 3, 8)
    np.testing.assert_allclose(m.points.max(axis=0), 0.016)
    grains = m.cell_data["gmsh:physical"][0]
    assert grains.min() == 1 and grains.max() == 8
    p = m.points[cells]  # positive volumes in VTK hexahedron order
    e1, e2, e3 = p[:, 1] - p[:, 0], p[:, 3] - p[:, 0], p[:, 4] - p[:, 0]
    assert np.all(np.einsum("ij,ij->i", np.cross(e1, e2), e3) > 0)
    ref = np.load(os.path.join(GOLDEN, "steel304_vtu.npz"))  # what the reference's meshio read gave
    np.testing.assert_array_equal(cells, ref["connectivity"])
    np.testing.assert_array_equal(grains - 1, ref["cell_ori_inds"])


@pytest.mark.parametrize("case", ["ta", "dp"])
def test_reference_runs_mesh_matches_their_output(case):
    """The tantalum (mesh10.msh) and dual-phase (n1-id1-mesh10.msh) meshes read here have the
    connectivity the reference's VTU output records."""
    from cpfem_amd import read_gmsh
    m = read_gmsh(os.path.join(GOLDEN, f"{case}_mesh.msh.gz"))
    ref = np.load(os.path.join(GOLDEN, f"{case}_vtu.npz"))
    np.testing.assert_array_equal(m.cells_dict["hexahedron"], ref["connectivity"])
    assert m.points.shape == (11 ** 3, 3)
    assert np.all(m.cell_data["gmsh:physical"][0] == 1)  # one Neper grain; orientations are per cell


def test_small_file_renumbering_and_filtering(tmp_path):
    from cpfem_amd import read_gmsh
    text = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
9
10 0 0 0
11 1 0 0
12 1 1 0
13 0 1 0
14 0 0 1
15 1 0 1
16 1 1 1
17 0 1 1
99 5 5 5
$EndNodes
$Elements
3
1 15 2 0 99 99
2 3 2 1 1 10 11 12 13
3 5 3 4 7 0 10 11 12 13 14 15 16 17
$EndElements
"""
    path = tmp_path / "one.msh"
    path.write_text(text)
    m = read_gmsh(str(path))
    np.testing.assert_array_equal(m.cells_dict["hexahedron"], [[0, 1, 2, 3, 4, 5, 6, 7]])
    assert m.cell_data["gmsh:physical"][0].tolist() == [4] and m.cell_data["gmsh:geometrical"][0].tolist() == [7]
    np.testing.assert_allclose(m.points[8], [5, 5, 5])
    bad = tmp_path / "v4.msh"
    bad.write_text(text.replace("2.2 0 8", "4.1 0 8"))
    with pytest.raises(ValueError):
        read_gmsh(str(bad))
