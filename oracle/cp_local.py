"""Oracle: per-quadrature-point Kalidindi crystal-plasticity update (FP64, NumPy).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, vectorised over a batch of Q quadrature points, the model of
``singlecrystal_copper/models_copper.py:21-319`` and its per-point-parameter
variants ``calibration/case1.py:145-270`` (9-array layout, q no-op quirk) and
``polycrystal_DPsteel/models_DPsteel_inhomo.py:232-330`` (10-array layout,
per-point stiffness, correct q).  All arrays are FP64; the batch axis is first.

Pins (tests/test_oracle_local.py, tests/test_oracle_golden.py):
  * rot_mat_from_quat / euler_zyx_to_rot   -> scipy.spatial.transform.Rotation
  * schmid_tensors                         -> n.d = 0, |P| = 1, trace 0 (closed form)
  * latent_matrix                          -> block-of-3 rule by brute force
  * cubic_stiffness / rotate_rank4         -> major/minor symmetry, R = I identity,
                                              90-degree z-rotation swap, isotropy
  * residual / local_newton                -> elastic limit (S = C:E, St.Venant-
                                              Kirchhoff), analytic Jacobian vs central
                                              FD, converged-S symmetry, golden curves
  * consistent tangent                     -> central FD of P(u_grad) through the
                                              full local solve
  * update / Cauchy stress                 -> golden single-crystal curves (Cu tension,
                                              objective scalar) end-to-end
"""
import numpy as np

AO = 1e-3  # reference strain rate gamma0_dot, models_copper.py:147 (`ao = 0.001`)

I3 = np.eye(3)


# --------------------------------------------------------------------------- a1/a2
def rot_mat_from_quat(q):
    """Scalar-first quaternion -> rotation matrix, no normalisation.

    models_copper.py:37-45 (`get_rot_mat`).
    """
    q = np.asarray(q, dtype=np.float64)
    q0, q1, q2, q3 = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3
    R[..., 0, 1] = 2 * q1 * q2 - 2 * q0 * q3
    R[..., 0, 2] = 2 * q1 * q3 + 2 * q0 * q2
    R[..., 1, 0] = 2 * q1 * q2 + 2 * q0 * q3
    R[..., 1, 1] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3
    R[..., 1, 2] = 2 * q2 * q3 - 2 * q0 * q1
    R[..., 2, 0] = 2 * q1 * q3 - 2 * q0 * q2
    R[..., 2, 1] = 2 * q2 * q3 + 2 * q0 * q1
    R[..., 2, 2] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3
    return R


def _rx(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(np.shape(t) + (3, 3))
    R[..., 0, 0] = 1.0
    R[..., 1, 1] = c
    R[..., 1, 2] = -s
    R[..., 2, 1] = s
    R[..., 2, 2] = c
    return R


def _ry(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(np.shape(t) + (3, 3))
    R[..., 1, 1] = 1.0
    R[..., 0, 0] = c
    R[..., 0, 2] = s
    R[..., 2, 0] = -s
    R[..., 2, 2] = c
    return R


def _rz(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(np.shape(t) + (3, 3))
    R[..., 2, 2] = 1.0
    R[..., 0, 0] = c
    R[..., 0, 1] = -s
    R[..., 1, 0] = s
    R[..., 1, 1] = c
    return R


def euler_zyx_to_rot(angles_deg):
    """Extrinsic z-y-x Euler angles in degrees -> rotation matrix.

    poly_sensitivity_rot_mesh2.py:41-53 uses
    ``Rotation.from_euler('zyx', [a, b, c], degrees=True).as_matrix()``: lower-case
    axes are extrinsic, so R = Rx(c) @ Ry(b) @ Rz(a).
    """
    ang = np.deg2rad(np.asarray(angles_deg, dtype=np.float64))
    return _rx(ang[..., 2]) @ _ry(ang[..., 1]) @ _rz(ang[..., 0])


# --------------------------------------------------------------------------- a3
def schmid_tensors(slip_table):
    """P0_a = d_a (x) n_a with normalised direction d (cols 3:6) and normal n (cols 0:3).

    models_copper.py:59-69.
    """
    t = np.asarray(slip_table, dtype=np.float64)
    d = t[:, 3:6] / np.linalg.norm(t[:, 3:6], axis=1)[:, None]
    n = t[:, 0:3] / np.linalg.norm(t[:, 0:3], axis=1)[:, None]
    return np.einsum("ai,aj->aij", d, n)


def latent_matrix(num_slip, r, self_block=True):
    """Latent-hardening matrix q.

    ``self_block=True``: q = r except q[i, 3*(i//3)+j] = 1 for j < 3
    (models_copper.py:71-76; models_DPsteel_inhomo.py:261-266).
    ``self_block=False``: q = r everywhere, the discarded ``q.at[].set`` of
    calibration/case1.py:170-175 (also case2.py, case4.py, models_sens_rot.py:175).
    """
    q = r * np.ones((num_slip, num_slip))
    if self_block:
        for i in range(num_slip):
            for j in range(3):
                q[i, i // 3 * 3 + j] = 1.0
    return q


def cubic_stiffness(C11, C12, C44):
    """Cubic stiffness C_ijkl from (C11, C12, C44); models_copper.py:92-130."""
    C = np.zeros((3, 3, 3, 3))
    for i in range(3):
        C[i, i, i, i] = C11
    for i, j in ((0, 1), (0, 2), (1, 2)):
        C[i, i, j, j] = C12
        C[j, j, i, i] = C12
        C[i, j, i, j] = C44
        C[i, j, j, i] = C44
        C[j, i, i, j] = C44
        C[j, i, j, i] = C44
    return C


# --------------------------------------------------------------------------- a4/a5
def rotate_rank2(R, T):
    """R T R^T; models_copper.py:29-32 (`rotate_tensor_rank_2`)."""
    return np.einsum("...ij,...kl,...jl->...ik", R, R, T)


def rotate_rank4(R, C):
    """C'_ikmo = R_ij R_kl R_mn R_op C_jlnp; models_copper.py:21-26."""
    return np.einsum("...ij,...kl,...mn,...op,...jlnp->...ikmo", R, R, R, R, C, optimize=True)


# --------------------------------------------------------------------------- material
class Material:
    """Per-point material description for a batch of Q points.

    Fields are arrays of shape (Q,) (per point, the 9/10-array layouts of
    calibration/case1.py:92-96 and models_DPsteel_inhomo.py:229) or scalars
    (the 4-array layout of models_copper.py:140-149, broadcast).

    ``C`` is either a (Q,3,3,3,3) / (3,3,3,3) stiffness or None with cubic
    constants ``C11, C12, C44``.
    """

    def __init__(self, schmid, xm, gss_a, h, t_sat, r, C=None, C11=None, C12=None,
                 C44=None, self_block=True, ao=AO):
        self.schmid = np.asarray(schmid, dtype=np.float64)
        self.N = self.schmid.shape[0]
        self.xm = xm
        self.gss_a = gss_a
        self.h = h
        self.t_sat = t_sat
        self.r = r
        if C is None:
            C = cubic_stiffness(C11, C12, C44)
        self.C = np.asarray(C, dtype=np.float64)
        self.self_block = self_block
        self.ao = ao

    def take(self, idx):
        """Material restricted to the points ``idx`` of the batch."""
        def sel(v):
            v = np.asarray(v)
            return v[idx] if v.ndim >= 1 and v.shape[0] > 1 and v.ndim != 4 else v

        C = self.C if self.C.ndim == 4 else self.C[idx]
        return Material(self.schmid, sel(self.xm), sel(self.gss_a), sel(self.h),
                        sel(self.t_sat), sel(self.r), C=C, self_block=self.self_block,
                        ao=self.ao)


def _col(v, Q):
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        return np.full((Q,), float(v))
    return v.reshape(Q)


def _q_matrix(mat, Q):
    """Per-point latent matrix (Q,N,N) built from the per-point r."""
    r = _col(mat.r, Q)
    N = mat.N
    q = r[:, None, None] * np.ones((Q, N, N))
    if mat.self_block:
        blk = np.zeros((N, N), dtype=bool)
        for i in range(N):
            for j in range(3):
                blk[i, i // 3 * 3 + j] = True
        q[:, blk] = 1.0
    return q


# --------------------------------------------------------------------------- a6/a7
class _Ctx:
    """Per-point constants of one local solve (everything except S)."""

    def __init__(self, H, Fpi_old, g_old, slip_old, R, mat, dt):
        Q = H.shape[0]
        self.Q = Q
        self.H = H
        self.Fpi_old = Fpi_old
        self.g_old = g_old
        self.slip_old = slip_old
        self.R = R
        self.dt = dt
        self.mat = mat
        # rotate_tensor_rank_2_vmap(rot_mat, Schmid_tensors): models_copper.py:173,189
        self.Prot = np.einsum("qij,ajl,qkl->qaik", R, mat.schmid, R)
        C = mat.C if mat.C.ndim == 5 else np.broadcast_to(mat.C, (Q, 3, 3, 3, 3))
        self.Crot = rotate_rank4(R, C)  # models_copper.py:199
        self.xm = _col(mat.xm, Q)
        self.gss_a = _col(mat.gss_a, Q)
        self.h = _col(mat.h, Q)
        self.t_sat = _col(mat.t_sat, Q)
        self.F = H + I3  # models_copper.py:188
        self.A = self.F @ Fpi_old

    def sub(self, m):
        c = _Ctx.__new__(_Ctx)
        c.Q = int(np.count_nonzero(m)) if m.dtype == bool else len(m)
        for k in ("H", "Fpi_old", "g_old", "slip_old", "R", "Prot", "Crot", "xm", "gss_a",
                  "h", "t_sat", "F", "A"):
            setattr(c, k, getattr(self, k)[m])
        c.dt = self.dt
        c.mat = self.mat
        return c


def _flow(ctx, S):
    """tau, gamma_inc, d(gamma_inc)/d(tau); models_copper.py:173-174."""
    tau = np.einsum("qij,qaij->qa", S, ctx.Prot)
    u = tau / ctx.g_old
    p = 1.0 / ctx.xm[:, None]
    c = ctx.mat.ao * ctx.dt
    gamma_inc = c * np.abs(u) ** p * np.sign(tau)
    # jacfwd of |u|**p*sign(tau): p*|u|**(p-1)*sign(u)/g*sign(tau)  (sign has zero jvp)
    dgamma = c * p * np.abs(u) ** (p - 1.0) * np.sign(u) * np.sign(tau) / ctx.g_old
    return tau, gamma_inc, dgamma


def helper(ctx, S):
    """tau -> gamma_inc -> Lp -> Fp_inv_new -> Fe; models_copper.py:172-192 (no hardening)."""
    tau, gamma_inc, dgamma = _flow(ctx, S)
    Lp = np.einsum("qa,qaij->qij", gamma_inc, ctx.Prot)  # :189
    Fpi_new = ctx.Fpi_old @ (I3 - Lp)  # :190
    Fe = ctx.F @ Fpi_new  # :191
    return tau, gamma_inc, dgamma, Lp, Fpi_new, Fe


def residual(ctx, S):
    """r(S) = S - C_rot : 1/2 (Fe^T Fe - I), row-major 9-vector; models_copper.py:195-201."""
    tau, gamma_inc, dgamma, Lp, Fpi_new, Fe = helper(ctx, S)
    E = 0.5 * (np.swapaxes(Fe, 1, 2) @ Fe - I3)
    S_ = np.einsum("qijkl,qkl->qij", ctx.Crot, E)
    return (S - S_).reshape(ctx.Q, 9), (tau, gamma_inc, dgamma, Lp, Fpi_new, Fe)


def jac_S(ctx, S, aux):
    """Exact d r / d S (9x9, row-major) — what jax.jacfwd(f_partial) yields (:226).

    dr = dS + C_rot : sym(Fe^T A dLp),  dLp = sum_a g'_a (P_a : dS) P_a, A = F Fp_inv_old.
    """
    tau, gamma_inc, dgamma, Lp, Fpi_new, Fe = aux
    M = np.swapaxes(Fe, 1, 2) @ ctx.A  # (Q,3,3)
    MP = np.einsum("qij,qajk->qaik", M, ctx.Prot)
    B = 0.5 * (MP + np.swapaxes(MP, 2, 3))
    Z = np.einsum("qijkl,qakl->qaij", ctx.Crot, B)  # (Q,N,3,3)
    J = np.einsum("qa,qai,qaj->qij", dgamma, Z.reshape(ctx.Q, -1, 9),
                  ctx.Prot.reshape(ctx.Q, -1, 9))
    return J + np.eye(9)


def jac_H(ctx, S, aux):
    """Exact d r / d u_grad (9x9) at fixed S — the u_grad block of jac_x (:256)."""
    tau, gamma_inc, dgamma, Lp, Fpi_new, Fe = aux
    # dFe = dH Fpi_new -> dE = sym(Fe^T dH Fpi_new); dr = -C_rot : dE
    X = np.einsum("qki,qlj->qklij", Fe, Fpi_new)  # Fe^T e_k (x) Fpi_new^T e_l
    X = 0.5 * (X + np.swapaxes(X, 3, 4))
    dr = -np.einsum("qijmn,qklmn->qijkl", ctx.Crot, X)
    return dr.reshape(ctx.Q, 9, 9)


def _solve(J, b):
    """Batched dense solve; a singular point yields NaN as jnp.linalg.solve does instead of
    raising (the local Newton of the reference has no error path)."""
    try:
        return np.linalg.solve(J, b)
    except np.linalg.LinAlgError:
        out = np.full(np.broadcast_shapes(J.shape[:-2] + (J.shape[-1],), b.shape[:-1]) + b.shape[-1:], np.nan)
        for i in range(J.shape[0]):
            try:
                out[i] = np.linalg.solve(J[i], b[i])
            except np.linalg.LinAlgError:
                pass
        return out


# --------------------------------------------------------------------------- a8
def local_newton(ctx, tol=1e-8, max_sub_step=5, max_iter=200):
    """Local Newton from S = 0 with the "cut-half" backtracking; models_copper.py:204-249.

    Per point: while |r| > tol: solve J d = -r; trial y + lam d for lam = 1, 1/2, ...
    until |r_trial| < |r| or max_sub_step trials; accept the last trial
    (y + 2*relax_f*d).  No iteration cap in the reference; ``max_iter`` only guards
    the oracle (the return flags points that hit it).
    Returns (S (Q,3,3), r (Q,9), iters (Q,), trials (Q,), capped (Q,) bool).
    """
    Q = ctx.Q
    y = np.zeros((Q, 9))
    r, _ = residual(ctx, y.reshape(Q, 3, 3))
    nr = np.linalg.norm(r, axis=1)
    iters = np.zeros(Q, dtype=np.int64)
    trials = np.zeros(Q, dtype=np.int64)
    active = nr > tol
    while np.any(active):
        idx = np.nonzero(active)[0]
        c = ctx.sub(idx)
        ya = y[idx]
        ra = r[idx]
        nra = nr[idx]
        _, aux = residual(c, ya.reshape(-1, 3, 3))
        J = jac_S(c, ya.reshape(-1, 3, 3), aux)
        dy = _solve(J, -ra[..., None])[..., 0]
        relax = np.ones(len(idx))
        crt = ra.copy()
        ncrt = nra.copy()
        sub = np.zeros(len(idx), dtype=np.int64)
        go = (ncrt >= nra) & (sub < max_sub_step)
        while np.any(go):
            gi = np.nonzero(go)[0]
            cg = c.sub(gi)
            rt, _ = residual(cg, (ya[gi] + relax[gi, None] * dy[gi]).reshape(-1, 3, 3))
            crt[gi] = rt
            ncrt[gi] = np.linalg.norm(rt, axis=1)
            relax[gi] *= 0.5
            sub[gi] += 1
            go = (ncrt >= nra) & (sub < max_sub_step)
        y[idx] = ya + 2.0 * relax[:, None] * dy
        r[idx] = crt
        nr[idx] = ncrt
        iters[idx] += 1
        trials[idx] += sub
        active = (nr > tol) & (iters < max_iter)
    capped = (nr > tol)
    return y.reshape(Q, 3, 3), r, iters, trials, capped


# --------------------------------------------------------------------------- a9/a10
def pk1(ctx, S, aux):
    """sigma = Fe S Fe^T / det Fe ; P = det F sigma F^-T; models_copper.py:155-162."""
    Fe = aux[5]
    sigma = (Fe @ S @ np.swapaxes(Fe, 1, 2)) / np.linalg.det(Fe)[:, None, None]
    P = np.linalg.det(ctx.F)[:, None, None] * sigma @ np.swapaxes(np.linalg.inv(ctx.F), 1, 2)
    return P


def tangent(ctx, S, aux):
    """Consistent tangent dP/du_grad (Q,9,9) via the IFT of f_jvp (models_copper.py:251-259).

    dS = -J_S^{-1} J_H dH, then the exact derivative of P(H, S(H)) through the
    flow rule.  Uses the identity P = F Y with Y = Fpi S Fpi^T / det Fpi (equal to
    det F * (Fe S Fe^T / det Fe) F^-T since Fe = F Fpi).
    """
    Q = ctx.Q
    tau, gamma_inc, dgamma, Lp, Fpi, Fe = aux
    JS = jac_S(ctx, S, aux)
    JH = jac_H(ctx, S, aux)
    dS = -_solve(JS, JH).reshape(Q, 3, 3, 9)  # dS[i,j] per direction
    detFpi = np.linalg.det(Fpi)
    Y = Fpi @ S @ np.swapaxes(Fpi, 1, 2) / detFpi[:, None, None]
    invFpi = np.linalg.inv(Fpi)
    out = np.empty((Q, 3, 3, 9))
    for m in range(9):
        dSm = dS[..., m]
        dtau = np.einsum("qij,qaij->qa", dSm, ctx.Prot)
        dLp = np.einsum("qa,qaij->qij", dgamma * dtau, ctx.Prot)
        dFpi = -ctx.Fpi_old @ dLp
        tr = np.einsum("qij,qji->q", invFpi, dFpi)
        dY = (dFpi @ S @ np.swapaxes(Fpi, 1, 2) + Fpi @ dSm @ np.swapaxes(Fpi, 1, 2)
              + Fpi @ S @ np.swapaxes(dFpi, 1, 2)) / detFpi[:, None, None] - Y * tr[:, None, None]
        dH = np.zeros((Q, 3, 3))
        dH[:, m // 3, m % 3] = 1.0
        out[..., m] = dH @ Y + ctx.F @ dY
    return out.reshape(Q, 9, 9)


def hardening(ctx, gamma_inc, q):
    """g_inc = q @ (h |dgamma| |1 - g/g_sat|^a sign(1 - g/g_sat)); models_copper.py:178-179."""
    x = 1.0 - ctx.g_old / ctx.t_sat[:, None]
    tmp = ctx.h[:, None] * np.abs(gamma_inc) * np.abs(x) ** ctx.gss_a[:, None] * np.sign(x)
    return np.einsum("qab,qb->qa", q, tmp)


# --------------------------------------------------------------------------- driver
def local_update(u_grad, Fp_inv, g, slip, R, mat, dt, tol=1e-8, max_sub_step=5,
                 want_tangent=True, max_iter=200):
    """Full per-point update for a batch: the maps of models_copper.py:151-170.

    Returns a dict with S, P (first_PK_stress), A = dP/du_grad (if asked),
    Fp_inv_new, g_new, slip_new (update_int_vars), sigma (P F^T / det F as in
    compute_avg_stress :308-313), iters, trials, capped.
    """
    H = np.asarray(u_grad, dtype=np.float64).reshape(-1, 3, 3)
    Q = H.shape[0]
    ctx = _Ctx(H, np.asarray(Fp_inv, dtype=np.float64).reshape(Q, 3, 3),
               np.asarray(g, dtype=np.float64).reshape(Q, -1),
               np.asarray(slip, dtype=np.float64).reshape(Q, -1),
               np.asarray(R, dtype=np.float64).reshape(Q, 3, 3), mat, float(dt))
    S, r, iters, trials, capped = local_newton(ctx, tol=tol, max_sub_step=max_sub_step,
                                               max_iter=max_iter)
    _, aux = residual(ctx, S)
    tau, gamma_inc, dgamma, Lp, Fpi_new, Fe = aux
    P = pk1(ctx, S, aux)
    q = _q_matrix(mat, Q)
    g_new = ctx.g_old + hardening(ctx, gamma_inc, q)  # :186
    slip_new = ctx.slip_old + gamma_inc  # :187
    sigma = P @ np.swapaxes(ctx.F, 1, 2) / np.linalg.det(ctx.F)[:, None, None]
    out = dict(S=S, P=P, Fp_inv_new=Fpi_new, g_new=g_new, slip_new=slip_new, sigma=sigma,
               iters=iters, trials=trials, capped=capped, residual=r)
    if want_tangent:
        out["A"] = tangent(ctx, S, aux)
    return out


def pk1_only(u_grad, Fp_inv, g, slip, R, mat, dt, tol=1e-8, max_sub_step=5):
    """first_PK_stress for a batch (no tangent)."""
    return local_update(u_grad, Fp_inv, g, slip, R, mat, dt, tol, max_sub_step,
                        want_tangent=False)["P"]
