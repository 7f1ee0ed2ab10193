"""Pins of the oracle adjoint (oracle/ad.py, the reference's AD path restated with torch
autograd on the CPU): gradients of whole load histories against central differences of the
oracle forward, w.r.t. hardening / elastic scale factors (FCC12, the 4-array model) and
per-grain z-y-x Euler angles (BCC24 dual-phase), plus a planted
mistake (the local Newton's IFT term dropped) that the comparison must catch."""
import numpy as np
import pytest

from helpers import slip_table

torch = pytest.importorskip("torch")

TIGHT = dict(tol=1e-10, rel_tol=1e-14, krylov=dict(tol=1e-13, atol=1e-20))


def _box(nx, ny, nz, L):
    from oracle import ad, fe
    pts, cells = fe.box_mesh(nx, ny, nz, *L)
    F = fe.faces(pts, L)
    bcs = [(F["left"], 0, 0.0), (F["front"], 1, 0.0), (F["bottom"], 2, 0.0), (F["top"], 2, None)]
    return ad.Mesh(pts, cells), bcs


def _fcc_case():
    from oracle import ad, cp_local
    mesh, bcs = _box(2, 1, 1, (2.0, 1.0, 1.0))
    consts = ad.Consts(cp_local.schmid_tensors(slip_table("slip_fcc12.txt")), tol=1e-10)
    rng = np.random.default_rng(1)
    q = rng.normal(size=(2, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    R = torch.as_tensor(cp_local.rot_mat_from_quat(q)).repeat_interleave(8, 0)
    Q = 8 * mesh.nc
    disps, ts = np.linspace(0, 3e-3, 3), np.linspace(0, 3.0, 3)
    base = torch.tensor([0.05, 2.0, 500.0, 300.0, 1.4, 204.6e3, 137.7e3, 126.2e3], dtype=torch.float64)
    W = torch.as_tensor(rng.normal(size=(2, 3, 3)))

    def loss(theta):  # theta: scales of g0, xm, gss_a, h, t_sat, r, C11, C12, C44
        mat8 = (base * theta[1:]).expand(Q, 8)
        g0 = (88.0 * theta[0]).expand(Q, 12)
        return (ad.simulate(mesh, consts, bcs, disps, ts, g0, mat8, R, TIGHT) * W).sum()

    return loss, torch.ones(9, dtype=torch.float64)


def _bcc_euler_case():
    from oracle import ad, cp_local
    mesh, bcs = _box(2, 1, 1, (2.0, 1.0, 1.0))
    consts = ad.Consts(cp_local.schmid_tensors(slip_table("slip_bcc24.txt")), tol=1e-10)
    rng = np.random.default_rng(2)
    G = 2
    q = rng.normal(size=(G, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    Rq = torch.as_tensor(cp_local.rot_mat_from_quat(q))
    grain = torch.as_tensor(np.array([0, 1]).repeat(8))
    hard = grain == 1
    Q = 8 * mesh.nc
    base = torch.tensor([0.05, 2.0, 400.0, 250.0, 1.4, 231.0e3, 135.0e3, 116.0e3], dtype=torch.float64)
    mat8 = base.expand(Q, 8).clone()
    mat8[hard, 2] *= 2.0
    g0 = torch.where(hard, 400.0, 100.0).to(torch.float64)[:, None].expand(Q, 24)
    disps, ts = np.linspace(0, 2e-3, 3), np.linspace(0, 2.0, 3)
    kw = dict(TIGHT, tol=1e-9)

    def loss(euler):  # R = R_euler(theta) R_quat per grain (poly_sensitivity_rot_mesh2.py:41-65,203)
        R = (ad.euler_zyx(euler) @ Rq)[grain]
        szz = ad.simulate(mesh, consts, bcs, disps, ts, g0, mat8, R, kw)[-1, 2, 2]
        return (szz - 1.01 * 40.0) ** 2

    return loss, torch.as_tensor(rng.uniform(-30, 30, size=(G, 3)))


def _grad(loss, x0):
    x = x0.clone().requires_grad_(True)
    loss(x).backward()
    return x.grad.numpy().reshape(-1)


def _grad_and_fd(loss, x0, h):
    g = _grad(loss, x0)
    fd = np.zeros_like(g)
    with torch.no_grad():
        for k in range(g.size):
            xp, xm = x0.clone().reshape(-1), x0.clone().reshape(-1)
            xp[k] += h
            xm[k] -= h
            fd[k] = (loss(xp.reshape(x0.shape)) - loss(xm.reshape(x0.shape))).item() / (2 * h)
    return g, fd


@pytest.fixture(scope="module")
def fcc_fd():
    loss, x0 = _fcc_case()
    return _grad_and_fd(loss, x0, 1e-6)


def test_oracle_adjoint_scales_match_fd(fcc_fd):
    g, fd = fcc_fd
    assert np.abs(g - fd).max() <= 1e-6 * np.abs(fd).max(), (g, fd)


def test_oracle_adjoint_euler_angles_dual_phase_match_fd():
    loss, x0 = _bcc_euler_case()
    g, fd = _grad_and_fd(loss, x0, 1e-4)
    assert np.abs(g - fd).max() <= 1e-6 * np.abs(fd).max(), (g, fd)


def test_planted_missing_ift_term_is_caught(monkeypatch, fcc_fd):
    """With the local Newton's implicit dependence dropped (S treated as a constant, as a
    wrong custom rule would do), the gradient misses FD by far more than the bar above."""
    from oracle import ad
    _, fd = fcc_fd
    monkeypatch.setattr(ad._LocalSolve, "backward",
                        staticmethod(lambda ctx, Sbar: tuple(torch.zeros_like(t) for t in ctx.saved_tensors[1:]) +
                                     (None, None)))
    loss, x0 = _fcc_case()
    g = _grad(loss, x0)
    assert np.abs(g - fd).max() > 1e-3 * np.abs(fd).max()
