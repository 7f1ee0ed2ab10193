"""Shared builders for the parity tests: the same seeded inputs for the oracle and the
HIP path."""
import os

import numpy as np

from conftest import GOLDEN

FCC_304 = dict(xm=0.02, gss_a=2.0, h=500.0, t_sat=300.0, r=1.4, C11=204.6e3, C12=137.7e3, C44=126.2e3)


def block_ell_pattern(cells, nn):
    """Block-ELL node pattern of a hex mesh as the library stores it: [slots][nn] node columns,
    ascending, padded with the row's own node, and the slot counts."""
    nb = [set() for _ in range(nn)]
    for c in cells:
        for a in c:
            nb[a].update(int(b) for b in c)
    ms = max(len(s) for s in nb)
    col = np.empty((ms, nn), dtype=np.int32)
    nslot = np.empty(nn, dtype=np.int32)
    for i, s in enumerate(nb):
        v = sorted(s)
        nslot[i] = len(v)
        col[:, i] = v + [i] * (ms - len(v))
    return col, nslot


def slip_table(name):
    return np.loadtxt(os.path.join(GOLDEN, name))


def random_states(Q, N, rng, scale=1e-3, g=88.0, table=None, C=(204.6e3, 137.7e3, 126.2e3), cap=1.6):
    """Seeded point states in the physical regime: the elastic-trial resolved shear stress
    max_a |tau_a| / g_a is drawn in (0, cap], so points range from elastic to strongly
    plastic without the overflowing trial states a 1/m = 50 power law cannot start from."""
    from oracle import cp_local
    H = scale * rng.normal(size=(Q, 3, 3))
    Fp = np.eye(3) + 0.05 * scale * rng.normal(size=(Q, 3, 3))
    gg = g * (1 + 0.1 * rng.random((Q, N)))
    slip = 1e-3 * rng.normal(size=(Q, N))
    q = rng.normal(size=(Q, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    if table is not None:
        P0 = cp_local.schmid_tensors(slip_table(table))
        R = cp_local.rot_mat_from_quat(q)
        Cr = cp_local.rotate_rank4(R, np.broadcast_to(cp_local.cubic_stiffness(*C), (Q, 3, 3, 3, 3)))
        Prot = np.einsum("qij,ajl,qkl->qaik", R, P0, R)
        target = cap * rng.random(Q) ** 0.5
        for _ in range(3):
            A = (np.eye(3) + H) @ Fp
            E = 0.5 * (np.swapaxes(A, 1, 2) @ A - np.eye(3))
            S = np.einsum("qijkl,qkl->qij", Cr, E)
            ratio = np.max(np.abs(np.einsum("qij,qaij->qa", S, Prot)) / gg, axis=1)
            H = H * (target / np.maximum(ratio, 1e-12))[:, None, None]
    return H, Fp, gg, slip, q


def quat_to_R(q):
    from oracle import cp_local
    return cp_local.rot_mat_from_quat(q)
