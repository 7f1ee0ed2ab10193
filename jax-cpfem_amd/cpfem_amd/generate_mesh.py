"""Structured hex8 box meshes, Gmsh 2.2 mesh ingest and the Mesh container (the
jax_fem.generate_mesh calls of the reference drivers, e.g.
calibration_case1_singleCrystalCopper_GB.py:106-108, and their meshio.read of Neper meshes,
polycrystal_304steel.py:83-103)."""
import gzip

import numpy as np


class Mesh:
    """points (n, 3) float64 and cells (c, 8) int hexahedra in VTK node order."""

    def __init__(self, points, cells, ele_type="HEX8"):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells, dtype=np.int64)
        self.ele_type = ele_type


class _MeshioLike:
    def __init__(self, points, cells, cell_data=None):
        self.points = points
        self.cells_dict = {"hexahedron": cells}
        self.cell_data = cell_data or {}


def read_gmsh(path):
    """Gmsh 2.2 ASCII mesh (what Neper's ``-M -elttype hex`` writes) -> an object with the
    attributes the reference drivers take from ``meshio.read``: ``points`` (n, 3),
    ``cells_dict['hexahedron']`` (c, 8, 0-based, Gmsh hexahedron order = VTK order) and
    ``cell_data['gmsh:physical']`` / ``['gmsh:geometrical']`` as one-element lists (the grain id
    of each cell is ``cell_data['gmsh:physical'][0] - 1``, polycrystal_304steel.py:85).
    Only 8-node hexahedra (element type 5) are read; other elements are skipped.  Nodes are
    renumbered densely in file order.  ``path`` may be gzip-compressed (``.gz``)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        lines = iter(f.read().splitlines())
    nodes, ids, hexes, phys, geom = None, None, [], [], []
    version = None
    for line in lines:
        tag = line.strip()
        if tag == "$MeshFormat":
            version = next(lines).split()[0]
            if not version.startswith("2"):
                raise ValueError(f"{path}: Gmsh format {version}, only 2.x ASCII is read")
        elif tag == "$Nodes":
            n = int(next(lines))
            ids = np.empty(n, dtype=np.int64)
            nodes = np.empty((n, 3))
            for k in range(n):
                v = next(lines).split()
                ids[k] = int(v[0])
                nodes[k] = (float(v[1]), float(v[2]), float(v[3]))
        elif tag == "$Elements":
            n = int(next(lines))
            for _ in range(n):
                v = next(lines).split()
                if int(v[1]) != 5:
                    continue
                ntag = int(v[2])
                tags = [int(t) for t in v[3:3 + ntag]]
                hexes.append([int(t) for t in v[3 + ntag:3 + ntag + 8]])
                phys.append(tags[0] if ntag > 0 else 0)
                geom.append(tags[1] if ntag > 1 else 0)
    if nodes is None or not hexes:
        raise ValueError(f"{path}: no nodes or no hexahedra")
    index = np.full(ids.max() + 1, -1, dtype=np.int64)
    index[ids] = np.arange(len(ids))
    cells = index[np.asarray(hexes, dtype=np.int64)]
    if (cells < 0).any():
        raise ValueError(f"{path}: element refers to an undefined node")
    data = {"gmsh:physical": [np.asarray(phys, dtype=np.int64)], "gmsh:geometrical": [np.asarray(geom, dtype=np.int64)]}
    return _MeshioLike(nodes, cells, data)


def get_meshio_cell_type(ele_type):
    if ele_type.upper() != "HEX8":
        raise NotImplementedError("only HEX8 is supported")
    return "hexahedron"


def box_mesh(Nx, Ny, Nz, Lx, Ly, Lz):
    """Structured box: node id = i + (Nx+1)(j + (Ny+1)k); returns an object with
    ``.points`` and ``.cells_dict['hexahedron']`` like the meshio mesh jax-fem returns."""
    xs = np.linspace(0.0, Lx, Nx + 1)
    ys = np.linspace(0.0, Ly, Ny + 1)
    zs = np.linspace(0.0, Lz, Nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def nid(a, b, c):
        return a + (Nx + 1) * (b + (Ny + 1) * c)

    cells = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)],
                     axis=1).astype(np.int64)
    return _MeshioLike(points, cells)


def voronoi_grains(cells_centroids, n_grains, L, rng):
    """Nearest-seed Voronoi grain id per cell; seeds uniform in [0, L)^3, ties to the lowest id."""
    seeds = rng.random((n_grains, 3)) * np.asarray(L, dtype=np.float64)
    ids = np.empty(len(cells_centroids), dtype=np.int64)
    for s in range(0, len(cells_centroids), 65536):
        c = cells_centroids[s:s + 65536]
        d2 = ((c[:, None, :] - seeds[None, :, :]) ** 2).sum(-1)
        ids[s:s + 65536] = np.argmin(d2, axis=1)
    return ids, seeds
