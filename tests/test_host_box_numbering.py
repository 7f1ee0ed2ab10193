"""Box numbering of a block-ELL pattern (cpfem_box_numbering, host code of libcpfem_hip.so;
no device needed).  The one-row-per-thread SpMV computes the columns of the nodes off the
box's faces from it, so it must be found exactly when every such node's 27 columns are
i + dz nx ny + dy nx + dx: on box meshes in any axis order, and not on a shuffled numbering,
a numbering with two interior nodes swapped, or a box one element thick."""
import ctypes
import os

import numpy as np
import pytest

from helpers import block_ell_pattern

pytest.importorskip("torch")


def _lib():
    from cpfem_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libcpfem_hip.so not built")
    return N.load_library()


def _dims(L, cells, nn):
    col, nslot = block_ell_pattern(cells, nn)
    dims = np.full(3, -1, dtype=np.int64)
    assert L.cpfem_box_numbering(nn, col.ctypes.data, nslot.ctypes.data, dims.ctypes.data) == 0
    return [int(d) for d in dims]


def _box(nx, ny, nz):
    from cpfem_amd.generate_mesh import box_mesh
    m = box_mesh(nx, ny, nz, float(nx), float(ny), float(nz))
    return m.cells_dict["hexahedron"], len(m.points)


def _columns_hold(cells, nn, dims):
    """Every node off the faces has the columns i + dz A + dy B + dx, checked by brute force."""
    col, nslot = block_ell_pattern(cells, nn)
    B, A = dims[0], dims[0] * dims[1]
    n_in = 0
    for i in range(nn):
        ix, iy, iz = i % B, (i // B) % dims[1], i // A
        if 0 < ix < dims[0] - 1 and 0 < iy < dims[1] - 1 and 0 < iz < dims[2] - 1:
            want = [i + dz * A + dy * B + dx for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
            assert nslot[i] == 27 and list(col[:, i]) == want
            n_in += 1
    return n_in


def test_box_meshes_are_found():
    L = _lib()
    for nx, ny, nz in [(2, 2, 2), (5, 4, 3), (3, 6, 4)]:
        cells, nn = _box(nx, ny, nz)
        d = _dims(L, cells, nn)
        assert d == [nx + 1, ny + 1, nz + 1]  # x varies fastest in the box mesh's numbering
        assert _columns_hold(cells, nn, d) == (nx - 1) * (ny - 1) * (nz - 1)


def test_other_axis_orders_are_found():
    """Renumbering the 5x4x3 box with z fastest, then y, then x gives strides (4, 20)."""
    L = _lib()
    cells, nn = _box(5, 4, 3)
    ix, iy, iz = np.arange(nn) % 6, (np.arange(nn) // 6) % 5, np.arange(nn) // 30
    new = iz + 4 * (iy + 5 * ix)  # old node -> new id
    d = _dims(L, new[cells], nn)
    assert d == [4, 5, 6]
    assert _columns_hold(new[cells], nn, d) == 4 * 3 * 2


def test_not_found():
    L = _lib()
    cells, nn = _box(5, 4, 3)
    perm = np.random.default_rng(0).permutation(nn)
    assert _dims(L, perm[cells], nn) == [0, 0, 0]
    # two interior nodes swapped: the strides read off the first interior node still match,
    # but the swapped nodes' columns do not
    swap = np.arange(nn)
    a, b = 1 + 6 * 1 + 30 * 1, 2 + 6 * 2 + 30 * 1
    swap[a], swap[b] = b, a
    assert _dims(L, swap[cells], nn) == [0, 0, 0]
    # one element thick: no node has 27 columns
    cells, nn = _box(4, 4, 1)
    assert _dims(L, cells, nn) == [0, 0, 0]


def test_bad_arguments():
    L = _lib()
    dims = np.zeros(3, dtype=np.int64)
    assert L.cpfem_box_numbering(0, None, None, dims.ctypes.data) != 0
