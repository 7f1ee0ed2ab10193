"""Node order of the direct solver (cpfem_band_order, host code of libcpfem_hip.so; no device
needed): a permutation whose reported bandwidth is the true one, the mesh's own lexicographic
order kept on a box mesh, and reverse Cuthill-McKee recovering a narrow band after the nodes
are shuffled."""
import ctypes
import os

import numpy as np
import pytest

from helpers import block_ell_pattern as _pattern

pytest.importorskip("torch")


def _lib():
    from cpfem_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libcpfem_hip.so not built")
    return N.load_library()


def _order(L, col, nslot):
    nn = len(nslot)
    ip = np.empty(nn, dtype=np.int32)
    bw = ctypes.c_int32()
    assert L.cpfem_band_order(nn, col.ctypes.data, nslot.ctypes.data, ip.ctypes.data, ctypes.byref(bw)) == 0
    return ip, bw.value


def _true_bw(col, nslot, ip):
    return max(abs(int(ip[i]) - int(ip[j])) for i in range(len(nslot)) for j in col[: nslot[i], i])


def test_box_mesh_keeps_lexicographic_order_and_shuffle_is_recovered():
    from cpfem_amd.generate_mesh import box_mesh
    L = _lib()
    n = 6
    m = box_mesh(n, n, n, 1.0, 1.0, 1.0)
    cells = m.cells_dict["hexahedron"]
    nn = len(m.points)
    col, nslot = _pattern(cells, nn)
    ip, bw = _order(L, col, nslot)
    assert sorted(ip) == list(range(nn))
    assert bw == _true_bw(col, nslot, ip)
    lex = (n + 1) ** 2 + (n + 1) + 1  # node id = i + (n+1)(j + (n+1)k)
    assert bw <= lex
    # shuffled node numbering: reverse Cuthill-McKee's level sets from a corner are diagonal
    # planes of the cube, about 1.5-2.2x wider than a lexicographic plane; still a narrow band
    # compared with the shuffled order itself
    perm = np.random.default_rng(0).permutation(nn)
    col2, nslot2 = _pattern(perm[cells], nn)
    ip2, bw2 = _order(L, col2, nslot2)
    assert sorted(ip2) == list(range(nn))
    assert bw2 == _true_bw(col2, nslot2, ip2)
    assert bw2 <= 2.5 * lex < _true_bw(col2, nslot2, np.arange(nn))


def test_bad_arguments():
    L = _lib()
    ip = np.empty(1, dtype=np.int32)
    bw = ctypes.c_int32()
    assert L.cpfem_band_order(0, None, None, ip.ctypes.data, ctypes.byref(bw)) != 0
