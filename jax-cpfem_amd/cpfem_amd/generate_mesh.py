"""Structured hex8 box meshes and the Mesh container (the jax_fem.generate_mesh calls of the
reference drivers, e.g. calibration_case1_singleCrystalCopper_GB.py:106-108)."""
import numpy as np


class Mesh:
    """points (n, 3) float64 and cells (c, 8) int hexahedra in VTK node order."""

    def __init__(self, points, cells, ele_type="HEX8"):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells, dtype=np.int64)
        self.ele_type = ele_type


class _MeshioLike:
    def __init__(self, points, cells):
        self.points = points
        self.cells_dict = {"hexahedron": cells}


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
