// Reverse-mode (adjoint) kernels: the per-point vector-Jacobian product through the converged
// local Newton (implicit-function theorem), the residual cotangent field, and the transpose of
// the block-ELL Jacobian.
//
// Reference: the backward pass of the custom_jvp local solver (models_copper.py:251-259, as
// transposed by JAX when differentiating update_int_vars_gp / compute_avg_stress) and
// implicit_vjp (crystal_plasticity_OR_design/solver.py:801-853): A^T lambda = v, then
// -(dc/dp)^T lambda.
#include "cp_dual.hpp"
#include "cp_internal.hpp"

namespace cpfem {

constexpr int VBS = 64;  // threads per block of the point VJP kernel

__device__ __forceinline__ double vjp_sg(const Geom& G, int comp, int64_t qp) {
  return G.uniform ? G.sg[comp * 8 + (qp & 7)] : G.sg[comp * G.n_qp + qp];
}
__device__ __forceinline__ double vjp_w(const Geom& G, int64_t qp) { return G.uniform ? G.jxw[qp & 7] : G.jxw[qp]; }

// Per-point VJP.  Output cotangents: Pbar [9][Q] (may be null), sigma cotangent given per
// cell (sbar_cell [9][n_cells], spread to points with the JxW cell-average weights; may be
// null), Fbar [9][Q], gbar [N][Q] (may be null).  Input cotangents written (or added when
// accumulate) for the direction groups in `mask`: 1 u_grad -> Hbar [9][Q], 2 Fp_inv_old ->
// Fpibar [9][Q], 4 g_old -> gbar_old [N][Q], 8 material -> pbar [8][Q], 16 R -> Rbar [9][Q].
template <int N>
__global__ void __launch_bounds__(VBS) k_local_vjp(Geom G, const double* __restrict__ u, StateView sv, MatView mv,
                                                  SlipTable<N> tab, SolveCfg cfg, int self_block, VjpArgs va) {
  // The re-solve's per-slip-system columns (2N) and, once it has finished, the VJP's 42-double
  // Jacobian column share one LDS allocation: each lane only touches its own column.
  __shared__ double sh[(2 * N > 42 ? 2 * N : 42) * VBS];
  const int64_t q = (int64_t)blockIdx.x * VBS + threadIdx.x;
  if (q >= G.n_qp) return;
  const int64_t Q = G.n_qp;
  double H[9], F[9], Fpi[9], R[9];
  {
    const int64_t cell = q >> 3;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t nd = G.cells[cell * 8 + a];
      const double g0 = vjp_sg(G, 3 * a, q), g1 = vjp_sg(G, 3 * a + 1, q), g2 = vjp_sg(G, 3 * a + 2, q);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double ui = u[3 * nd + i];
        H[3 * i] += ui * g0;
        H[3 * i + 1] += ui * g1;
        H[3 * i + 2] += ui * g2;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
    Fpi[i] = sv.Fpi[i * sv.ld + q];
    R[i] = sv.rot[i * sv.ld + q];
  }
  double mp[NPAR];
#pragma unroll
  for (int m = 0; m < NPAR; ++m) mp[m] = mv.ptr[m][q * mv.stride[m]];
  PointMat pm;
  pm.p = 1.0 / mp[MP_XM];
  pm.c = mv.ao * mv.dt;
  pm.gss_a = mp[MP_GSS_A];
  pm.h = mp[MP_H];
  pm.tsat = mp[MP_TSAT];
  pm.r = mp[MP_R];
  pm.C11 = mp[MP_C11];
  pm.C12 = mp[MP_C12];
  pm.C44 = mp[MP_C44];
  double s[6];
  {
    Point<N> pt;
    pt.setup(F, Fpi, R, sv.g + q, sv.ld, pm, sh + threadIdx.x, sh + N * VBS + threadIdx.x, VBS);
    int32_t st = 0;
    pt.solve(tab, cfg.tol, cfg.max_sub_step, cfg.max_iter, s, &st);
  }
  // output cotangents
  double Pb[9], sb[9], Fb[9];
  const double w = vjp_w(G, q);
  double wsum = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) wsum += vjp_w(G, (q & ~(int64_t)7) + j);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    Pb[i] = va.Pbar ? va.Pbar[i * Q + q] : 0.0;
    sb[i] = va.sbar_cell ? va.sbar_cell[i * G.n_cells + (q >> 3)] * w / wsum : 0.0;
    Fb[i] = va.Fbar ? va.Fbar[i * Q + q] : 0.0;
  }
  SeededIn<N> xin{H, Fpi, sv.g + q, sv.ld, mp, R, -1};
  const double* gbq = va.gbar ? va.gbar + q : nullptr;
  const double* slq = va.slipbar ? va.slipbar + q : nullptr;
  int mask = va.mask;
  if (!va.Hbar) mask &= ~1;
  if (!va.Fpibar) mask &= ~2;
  if (!va.gbar_old) mask &= ~4;
  if (!va.pbar) mask &= ~8;
  if (!va.Rbar) mask &= ~16;
  point_vjp<N>(xin, s, tab, mv.ao, mv.dt, self_block, Pb, sb, Fb, gbq, slq, Q, sh + threadIdx.x, VBS, mask,
               [&](int grp, int off, double xb) {
                 double* out = grp == 1 ? va.Hbar : grp == 2 ? va.Fpibar : grp == 4 ? va.gbar_old : grp == 8 ? va.pbar : va.Rbar;
                 double* dst = out + (int64_t)off * Q + q;
                 *dst = va.accumulate ? *dst + xb : xb;
               });
}

// Residual cotangent of implicit_vjp: Pbar_q = -JxW_q * grad(lambda)_q with lambda zero on
// Dirichlet dofs (those rows of the constraint do not depend on the parameters).
__global__ void __launch_bounds__(256) k_grad_field(Geom G, const double* __restrict__ lam, const uint8_t* __restrict__ is_bc,
                                                   double scale_by_w, double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= G.n_qp) return;
  const int64_t cell = q >> 3;
  double Gq[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Gq[i] = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t nd = G.cells[cell * 8 + a];
    const double g0 = vjp_sg(G, 3 * a, q), g1 = vjp_sg(G, 3 * a + 1, q), g2 = vjp_sg(G, 3 * a + 2, q);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double li = is_bc[3 * nd + i] ? 0.0 : lam[3 * nd + i];
      Gq[3 * i] += li * g0;
      Gq[3 * i + 1] += li * g1;
      Gq[3 * i + 2] += li * g2;
    }
  }
  const double sc = scale_by_w != 0.0 ? scale_by_w * vjp_w(G, q) : 1.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i * G.n_qp + q] = sc * Gq[i];
}

// A^T in the same block-ELL layout: block (j, s') of A^T is the transpose of block
// (i = col[s'][j], tslot[s'][j]) of A; padding slots stay zero.
__global__ void __launch_bounds__(256) k_transpose(BellView M, const int32_t* __restrict__ tslot, const double* __restrict__ A,
                                                  double* __restrict__ At) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nrow = 3 * M.n_nodes;
  if (t >= nrow * M.n_slots) return;
  const int sp = (int)(t / nrow);
  const int64_t r = t - (int64_t)sp * nrow;
  const int64_t j = r / 3;
  const int c = (int)(r - 3 * j);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (sp < M.nslot[j]) {
    const int64_t i = M.col[(int64_t)sp * M.n_nodes + j];
    const int s = tslot[(int64_t)sp * M.n_nodes + j];
    const double* src = A + (3 * (int64_t)s + c) * nrow + 3 * i;
    v0 = src[0];
    v1 = src[1];
    v2 = src[2];
  }
  At[(3 * (int64_t)sp + 0) * nrow + r] = v0;
  At[(3 * (int64_t)sp + 1) * nrow + r] = v1;
  At[(3 * (int64_t)sp + 2) * nrow + r] = v2;
}

template <int N>
void launch_local_vjp(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                      const SolveCfg& cfg, int self_block, const VjpArgs& va, hipStream_t st) {
  const int64_t nb = (G.n_qp + VBS - 1) / VBS;
  hipLaunchKernelGGL(k_local_vjp<N>, dim3((unsigned)nb), dim3(VBS), 0, st, G, u, sv, mv, tab, cfg, self_block, va);
}

void launch_grad_field(const Geom& G, const double* lam, const uint8_t* is_bc, double scale_by_w, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_grad_field, dim3((unsigned)((G.n_qp + 255) / 256)), dim3(256), 0, st, G, lam, is_bc, scale_by_w, out);
}

void launch_transpose(const BellView& M, const int32_t* tslot, const double* A, double* At, hipStream_t st) {
  const int64_t n = 3 * M.n_nodes * M.n_slots;
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M, tslot, A, At);
}

template void launch_local_vjp<12>(const Geom&, const double*, const StateView&, const MatView&, const SlipTable<12>&,
                                   const SolveCfg&, int, const VjpArgs&, hipStream_t);
template void launch_local_vjp<24>(const Geom&, const double*, const StateView&, const MatView&, const SlipTable<24>&,
                                   const SolveCfg&, int, const VjpArgs&, hipStream_t);

}  // namespace cpfem
