// Element-level kernels of the CP hot path: quadrature-point constitutive update,
// element residual / stiffness, scatter-free gathers into the block-ELL matrix and the
// residual vector, cell-averaged Cauchy stress and the history update.
//
// Upstream element semantics (jax-fem, used through solver.py:392 and restated by
// models_copper.py:277-278,315): u_grad_q = sum_a u_a (x) dN_a(q);
// r_e[a,i] = sum_q P_q,ij dN_a,j JxW_q; K_e[a i, b k] = sum_q dN_a,j A_q[ij,kl] dN_b,l JxW_q.
#include <algorithm>

#include "cp_internal.hpp"

namespace cpfem {

__device__ __forceinline__ PointMat load_mat(const MatView& mv, int64_t q) {
  PointMat m;
  m.p = 1.0 / mv.ptr[MP_XM][q * mv.stride[MP_XM]];
  m.c = mv.ao * mv.dt;
  m.gss_a = mv.ptr[MP_GSS_A][q * mv.stride[MP_GSS_A]];
  m.h = mv.ptr[MP_H][q * mv.stride[MP_H]];
  m.tsat = mv.ptr[MP_TSAT][q * mv.stride[MP_TSAT]];
  m.r = mv.ptr[MP_R][q * mv.stride[MP_R]];
  m.C11 = mv.ptr[MP_C11][q * mv.stride[MP_C11]];
  m.C12 = mv.ptr[MP_C12][q * mv.stride[MP_C12]];
  m.C44 = mv.ptr[MP_C44][q * mv.stride[MP_C44]];
  return m;
}

__device__ __forceinline__ double sg_at(const Geom& G, int comp, int64_t qp) {
  return G.uniform ? G.sg[comp * 8 + (qp & 7)] : G.sg[comp * G.n_qp + qp];
}
__device__ __forceinline__ double jxw_at(const Geom& G, int64_t qp) { return G.uniform ? G.jxw[qp & 7] : G.jxw[qp]; }

// u_grad at quadrature point qp from the nodal solution (AoS dofs, 3 per node)
__device__ __forceinline__ void ugrad_from_dofs(const Geom& G, const double* __restrict__ u, int64_t qp, double* H) {
  const int64_t cell = qp >> 3;
#pragma unroll
  for (int i = 0; i < 9; ++i) H[i] = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t nd = G.cells[cell * 8 + a];
    const double g0 = sg_at(G, 3 * a, qp), g1 = sg_at(G, 3 * a + 1, qp), g2 = sg_at(G, 3 * a + 2, qp);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double ui = u[3 * nd + i];
      H[3 * i] += ui * g0;
      H[3 * i + 1] += ui * g1;
      H[3 * i + 2] += ui * g2;
    }
  }
}

__device__ __forceinline__ void load_point(const StateView& sv, int64_t q, double* Fpi, double* R) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    Fpi[i] = sv.Fpi[i * sv.ld + q];
    R[i] = sv.rot[i * sv.ld + q];
  }
}

constexpr int QBS = 128;  // threads per block of the point kernels (LDS: 2 * N * QBS doubles)

// Sum of a 32-bit count over the 64 lanes of a wave (every lane takes part).
__device__ __forceinline__ uint32_t wave_count(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Point q's inputs: F = I + u_grad from the dofs, Fp_inv_old and R, and the solver set-up
// (rotated stiffness, 1/g_old in LDS).
// With gi (a [N][n_qp] device array) the 1/g_old table lives there instead of in LDS, and
// sh holds only the d dgamma / d tau table.
template <int N>
__device__ __forceinline__ void start_point(const Geom& G, const double* __restrict__ u, const StateView& sv,
                                            const MatView& mv, int64_t q, double* sh, Point<N>& pt, double* F,
                                            double* Fpi, double* gi = nullptr) {
  double H[9], R[9];
  ugrad_from_dofs(G, u, q, H);
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
  load_point(sv, q, Fpi, R);
  const PointMat m = load_mat(mv, q);
  if (gi)
    pt.setup(F, Fpi, R, sv.g + q, sv.ld, m, sh + threadIdx.x, gi + q, QBS, G.n_qp);
  else
    pt.setup(F, Fpi, R, sv.g + q, sv.ld, m, sh + threadIdx.x, sh + N * QBS + threadIdx.x, QBS);
}

// Outputs of a converged point: JxW P, JxW dP/du_grad (wA), the converged stress (sout,
// [6][Q]), iterations, status bits.
template <int N>
__device__ __forceinline__ void finish_point(const Geom& G, const SlipTable<N>& tab, int64_t q, Point<N>& pt,
                                             const double* F, const double* Fpi, const double* s, int it, int32_t st,
                                             double* __restrict__ wP, double* __restrict__ wA,
                                             int32_t* __restrict__ iters, int32_t* __restrict__ status,
                                             double* __restrict__ sout = nullptr) {
  const double w = jxw_at(G, q);
  double P[9];
  pk1<N>(pt, F, Fpi, s, P);
#pragma unroll
  for (int i = 0; i < 9; ++i) wP[i * G.n_qp + q] = w * P[i];
  if (sout) {
#pragma unroll
    for (int j = 0; j < 6; ++j) sout[j * G.n_qp + q] = s[j];
  }
  if (wA) {
    tangent<N>(pt, tab, F, Fpi, s, [&](int m9, const double* dP) {
#pragma unroll
      for (int i = 0; i < 9; ++i) wA[(int64_t)(9 * i + m9) * G.n_qp + q] = w * dP[i];
    });
  }
  if (iters) iters[q] = it;
  if (st) atomicOr(status, st);
}

__device__ __forceinline__ void add_counts(unsigned long long* counts, uint32_t n_it, uint32_t n_ev) {
  if (!counts) return;
  n_it = wave_count(n_it);
  n_ev = wave_count(n_ev);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counts[0], (unsigned long long)n_it);
    atomicAdd(&counts[1], (unsigned long long)n_ev);
  }
}

constexpr double DBL_MAXV = 1.7976931348623157e308;

// ---------------------------------------------------------------------------------
// Quadrature-point kernels of the assembly.  k_element_res: H from dofs, local Newton,
// JxW-weighted P (SoA [9][Q]) and, for Jacobian assemblies, the converged stress (sout,
// [6][Q]); k_point_tangent then forms the JxW-weighted consistent tangent ([81][Q]: each of
// the 81 stores of a wave writes 512 contiguous bytes) from the stored stress.  Without the
// tangent the Newton kernel fits 256 VGPRs (a few spilled) and keeps the 1/g_old table in a
// device array instead of LDS, so two waves share a SIMD; the tangent kernel needs ~440 VGPRs
// (one wave per SIMD) but runs one residual evaluation and the tangent per point, no Newton
// loop whose lanes wait on each other.  The tangent kernel re-evaluates the residual at the
// stored stress, which gives bitwise the state the Newton kernel ended with (same inline
// code, multiply-adds fused per source expression).
// The lanes of a wave run their Newton iterations in lockstep; a wave whose points need very
// different iteration counts would otherwise run until its slowest lane converges.  With
// parking (pk.lanes > 0), once at most pk.lanes lanes of the wave are still iterating, those
// lanes store their point's iterate in the park queue and the wave finishes the others;
// k_element_qp_park completes the parked points.  The iterates, and so every output, are the
// same either way: a parked point resumes from its stored stress with the residual
// re-evaluated there.  counts (optional): [0] += local Newton iterations, [1] += residual
// evaluations (one atomic per wave), the measured inputs of the algorithmic flop count.
template <int N>
__device__ __forceinline__ void element_qp_body(const Geom& G, const double* __restrict__ u, const StateView& sv,
                                                const MatView& mv, const SlipTable<N>& tab, const SolveCfg& cfg,
                                                double* __restrict__ wP, double* __restrict__ sout,
                                                int32_t* __restrict__ iters, int32_t* __restrict__ status,
                                                unsigned long long* __restrict__ counts, const PointPark& pk, double* sh,
                                                double* gi) {
  const int64_t q = (int64_t)blockIdx.x * QBS + threadIdx.x;
  const bool valid = q < G.n_qp;
  Point<N> pt;
  double F[9], Fpi[9], s[6];
  int it = 0;
  int32_t st = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k] = 0.0;
  pt.nev = 0;
  pt.nr = 0.0;
  if (valid) {
    start_point<N>(G, u, sv, mv, q, sh, pt, F, Fpi, gi);
    pt.eval(tab, s);
  }
  bool act = valid && pt.nr > cfg.tol;
  for (;;) {  // wave-uniform exits: every lane sees the same ballot
    const int nact = __popcll(__ballot(act));
    if (nact == 0 || nact <= pk.lanes) break;
    if (act) {
      if (it >= cfg.max_iter) {
        st |= PT_CAPPED;
        act = false;
      } else {
        pt.newton_step(tab, cfg.max_sub_step, s);
        ++it;
        act = pt.nr > cfg.tol;
      }
    }
  }
  const unsigned long long park = __ballot(act);
  if (park) {  // park the lanes still iterating (one queue reservation per wave)
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(park) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&pk.ctr[0], (uint32_t)__popcll(park));
    base = __shfl(base, leader, 64);
    if (act) {
      const int64_t k = (int64_t)base + __popcll(park & ((1ull << lane) - 1ull));
      pk.idx[k] = (int32_t)q;
      pk.it[k] = it;
#pragma unroll
      for (int j = 0; j < 6; ++j) pk.s[j * pk.cap + k] = s[j];
    }
  }
  uint32_t n_it = 0;
  if (valid && !act) {
    if (!(fabs(pt.nr) <= DBL_MAXV)) st |= PT_NONFINITE;
    finish_point<N>(G, tab, q, pt, F, Fpi, s, it, st, wP, nullptr, iters, status, sout);
    n_it = (uint32_t)it;
  }
  add_counts(counts, n_it, valid ? (uint32_t)pt.nev : 0u);
}

template <int N>
__global__ void __launch_bounds__(QBS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_element_res(Geom G, const double* __restrict__ u, StateView sv, MatView mv, SlipTable<N> tab, SolveCfg cfg,
              double* __restrict__ wP, double* __restrict__ sout, int32_t* __restrict__ iters,
              int32_t* __restrict__ status, unsigned long long* __restrict__ counts, PointPark pk,
              double* __restrict__ gi) {
  __shared__ double sh[N * QBS];
  element_qp_body<N>(G, u, sv, mv, tab, cfg, wP, sout, iters, status, counts, pk, sh, gi);
}

// Consistent tangent of every point from its converged stress sin ([6][Q]).
template <int N>
__global__ void __launch_bounds__(QBS) k_point_tangent(Geom G, const double* __restrict__ u, StateView sv, MatView mv,
                                                      SlipTable<N> tab, const double* __restrict__ sin,
                                                      double* __restrict__ wA) {
  __shared__ double sh[2 * N * QBS];
  const int64_t q = (int64_t)blockIdx.x * QBS + threadIdx.x;
  if (q >= G.n_qp) return;
  Point<N> pt;
  double F[9], Fpi[9], s[6];
  start_point<N>(G, u, sv, mv, q, sh, pt, F, Fpi);
#pragma unroll
  for (int j = 0; j < 6; ++j) s[j] = sin[j * G.n_qp + q];
  pt.eval(tab, s);
  const double w = jxw_at(G, q);
  tangent<N>(pt, tab, F, Fpi, s, [&](int m9, const double* dP) {
#pragma unroll
    for (int i = 0; i < 9; ++i) wA[(int64_t)(9 * i + m9) * G.n_qp + q] = w * dP[i];
  });
}

// The parked points of k_element_res, one queue entry per lane: each point's stress iterate is
// restored, its residual re-evaluated there and its Newton iterations continued to
// convergence.  The grid covers the longest possible queue; blocks past the queue's length
// (final when this kernel starts: k_element_res ran before it on the same stream) exit at once,
// so only as many blocks as parked points run, dispatched as earlier blocks retire.  (Looping
// each wave over the queue instead costs this kernel 64 spilled registers per lane.)
template <int N>
__global__ void __launch_bounds__(QBS) k_element_qp_park(Geom G, const double* __restrict__ u, StateView sv, MatView mv,
                                                        SlipTable<N> tab, SolveCfg cfg, double* __restrict__ wP,
                                                        double* __restrict__ sout, int32_t* __restrict__ iters,
                                                        int32_t* __restrict__ status,
                                                        unsigned long long* __restrict__ counts, PointPark pk) {
  __shared__ double sh[2 * N * QBS];
  const uint32_t total = __hip_atomic_load(&pk.ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((int64_t)blockIdx.x * QBS >= (int64_t)total) return;  // uniform over the block
  if (counts && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counts[2], (unsigned long long)total);
  const int64_t k = (int64_t)blockIdx.x * QBS + threadIdx.x;
  uint32_t n_it = 0, n_ev = 0;
  if (k < (int64_t)total) {
    const int64_t q = pk.idx[k];
    Point<N> pt;
    double F[9], Fpi[9], s[6];
    start_point<N>(G, u, sv, mv, q, sh, pt, F, Fpi);
#pragma unroll
    for (int j = 0; j < 6; ++j) s[j] = pk.s[j * pk.cap + k];
    pt.eval(tab, s);
    int it = pk.it[k];
    int32_t st = 0;
    pt.iterate(tab, cfg.tol, cfg.max_sub_step, cfg.max_iter, s, &it, &st);
    finish_point<N>(G, tab, q, pt, F, Fpi, s, it, st, wP, nullptr, iters, status, sout);
    n_it = (uint32_t)it;
    n_ev = (uint32_t)pt.nev;
  }
  add_counts(counts, n_it, n_ev);
}

// ---------------------------------------------------------------------------------
// Element stiffness / residual: one thread per (cell, local node a); writes the 8 node-pair
// blocks of row-block a (KE[cell][a][b][i][k]) and r_e[cell][a][i].  A block takes 16 cells
// (128 points, contiguous in the [component][Q] layouts) and stages their JxW P and, one
// row index i at a time, the 27 tangent components dP_(ij)/dH_(kl) of all 128 points in LDS
// with coalesced loads: each value is read from HBM once, where eight threads of a cell
// reading it from global memory re-fetched the tangent ~7 times (L2 misses) per launch.
constexpr int EK_CELLS = 16;
constexpr int EK_BS = 8 * EK_CELLS;  // threads = points per block

// Two waves per SIMD: 256 VGPRs with 12 B of spills (one wave per SIMD at 258 registers
// without the bound).
__global__ void __launch_bounds__(EK_BS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_element_K(Geom G, const double* __restrict__ wP, const double* __restrict__ wA, double* __restrict__ KE,
            double* __restrict__ RE) {
  __shared__ double sP[9][EK_BS];
  __shared__ double sA[27][EK_BS];
  const int64_t q0 = (int64_t)blockIdx.x * EK_BS;  // first point of the block
  const int tid = threadIdx.x;
  const int64_t t = q0 + tid;                      // (cell, a) of this thread
  const bool valid = t < G.n_qp;
  const int64_t cell = t >> 3;
  const int a = tid & 7;
  const int cl = tid >> 3;  // cell within the block
#pragma unroll
  for (int c = 0; c < 9; ++c) sP[c][tid] = valid ? wP[c * G.n_qp + t] : 0.0;
  double K[72];
#pragma unroll
  for (int k = 0; k < 72; ++k) K[k] = 0.0;
  double re[3] = {0.0, 0.0, 0.0};
#pragma unroll  // i indexes the register arrays K and re: unrolled, no scratch
  for (int i = 0; i < 3; ++i) {
    __syncthreads();  // sA free (previous i), sP written
#pragma unroll
    for (int c = 0; c < 27; ++c) sA[c][tid] = valid ? wA[(int64_t)(27 * i + c) * G.n_qp + t] : 0.0;
    __syncthreads();
    if (valid) {
#pragma unroll 1
      for (int q = 0; q < 8; ++q) {
        const int p = 8 * cl + q;
        const int64_t qp = cell * 8 + q;
        double ga[3], gb[24];
#pragma unroll
        for (int c = 0; c < 24; ++c) gb[c] = sg_at(G, c, qp);
        ga[0] = sg_at(G, 3 * a, qp);
        ga[1] = sg_at(G, 3 * a + 1, qp);
        ga[2] = sg_at(G, 3 * a + 2, qp);
        double sr = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) sr += sP[3 * i + j][p] * ga[j];
        re[i] += sr;
        double Z[9];  // Z[k][l] = sum_j ga_j A[(i j),(k l)]
#pragma unroll
        for (int kl = 0; kl < 9; ++kl) {
          double z = 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) z += ga[j] * sA[9 * j + kl][p];
          Z[kl] = z;
        }
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            K[b * 9 + i * 3 + k] += Z[3 * k] * gb[3 * b] + Z[3 * k + 1] * gb[3 * b + 1] + Z[3 * k + 2] * gb[3 * b + 2];
      }
    }
  }
  if (!valid) return;
  double2* out = reinterpret_cast<double2*>(KE + t * 72);
#pragma unroll
  for (int k = 0; k < 36; ++k) out[k] = make_double2(K[2 * k], K[2 * k + 1]);
#pragma unroll
  for (int k = 0; k < 3; ++k) RE[t * 3 + k] = re[k];
}

// ---------------------------------------------------------------------------------
// Block-ELL gather: thread per (slot s, scalar row r = 3 i + c), padding slots included.
// Sums the element blocks
// that touch (node i, node col[s][i]) in ascending cell order (the order in which scipy's
// COO->CSR sum_duplicates adds them, solver.py:281-282), then applies the row elimination
// of get_A (solver.py:290-293): a Dirichlet row becomes zero with a unit diagonal.
__global__ void __launch_bounds__(256) k_gather_matrix(BellView M, const double* __restrict__ KE,
                                                      const int32_t* __restrict__ cptr, const int32_t* __restrict__ clist,
                                                      const uint8_t* __restrict__ is_bc, double* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nrow = 3 * M.n_nodes;
  if (t >= nrow * M.n_slots) return;
  const int s = (int)(t / nrow);
  const int64_t r = t - (int64_t)s * nrow;
  const int64_t i = r / 3;
  const int c = (int)(r - 3 * i);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
  if (s >= M.nslot[i]) {
    // padding slot (column = own node): explicit zeros so the SpMV can run a fixed slot count
  } else if (is_bc[r]) {
    const int64_t j = M.col[(int64_t)s * M.n_nodes + i];
    if (j == i) {
      v0 = (c == 0) ? 1.0 : 0.0;
      v1 = (c == 1) ? 1.0 : 0.0;
      v2 = (c == 2) ? 1.0 : 0.0;
    }
  } else {
    const int64_t e0 = cptr[(int64_t)s * M.n_nodes + i], e1 = cptr[(int64_t)s * M.n_nodes + i + 1];
    for (int64_t e = e0; e < e1; ++e) {
      const double* blk = KE + (int64_t)clist[e] * 9 + 3 * c;
      v0 += blk[0];
      v1 += blk[1];
      v2 += blk[2];
    }
  }
  vals[(3 * (int64_t)s + 0) * nrow + r] = v0;
  vals[(3 * (int64_t)s + 1) * nrow + r] = v1;
  vals[(3 * (int64_t)s + 2) * nrow + r] = v2;
}

// Residual gather + apply_bc_vec (solver.py:119-133): res[bc] = u[bc] - ubar.
// Also writes per-block partial sums of res^2 (deterministic reduction finished on device).
__global__ void __launch_bounds__(256) k_gather_residual(int64_t n_nodes, const double* __restrict__ RE,
                                                        const int32_t* __restrict__ nptr, const int32_t* __restrict__ nlist,
                                                        const uint8_t* __restrict__ is_bc, const double* __restrict__ ubar,
                                                        const double* __restrict__ u, double* __restrict__ res) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * n_nodes) return;
  const int64_t i = r / 3;
  const int c = (int)(r - 3 * i);
  double v = 0.0;
  if (is_bc[r]) {
    v = u[r] - ubar[r];
  } else {
    for (int64_t e = nptr[i]; e < nptr[i + 1]; ++e) v += RE[(int64_t)nlist[e] * 3 + c];
  }
  res[r] = v;
}

// Element residual only: thread per (cell, a)
__global__ void __launch_bounds__(256) k_element_R(Geom G, const double* __restrict__ wP, double* __restrict__ RE) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.n_qp) return;
  const int64_t cell = t >> 3;
  const int a = (int)(t & 7);
  double re[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    const int64_t qp = cell * 8 + q;
    const double g0 = sg_at(G, 3 * a, qp), g1 = sg_at(G, 3 * a + 1, qp), g2 = sg_at(G, 3 * a + 2, qp);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      re[i] += wP[(3 * i) * G.n_qp + qp] * g0 + wP[(3 * i + 1) * G.n_qp + qp] * g1 + wP[(3 * i + 2) * G.n_qp + qp] * g2;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) RE[t * 3 + i] = re[i];
}

// ---------------------------------------------------------------------------------
// Point-level update (tensor_map / update_int_vars_map, models_copper.py:263-269) on explicit
// u_grad arrays [9][Q] (or from dofs when Hd == nullptr).  Any output pointer may be null.
template <int N, bool TANGENT>
__global__ void __launch_bounds__(QBS) k_local(int64_t n_qp, const double* __restrict__ Hd, Geom G,
                                              const double* __restrict__ u, StateView sv, MatView mv, SlipTable<N> tab,
                                              SolveCfg cfg, int self_block, double* __restrict__ P_out,
                                              double* __restrict__ A_out, double* __restrict__ sig_out,
                                              StateOut so, int32_t* __restrict__ iters, int32_t* __restrict__ status) {
  __shared__ double sh[2 * N * QBS];
  const int64_t q = (int64_t)blockIdx.x * QBS + threadIdx.x;
  if (q >= n_qp) return;
  double H[9], F[9], Fpi[9], R[9];
  if (Hd) {
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = Hd[i * n_qp + q];
  } else {
    ugrad_from_dofs(G, u, q, H);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
  load_point(sv, q, Fpi, R);
  const PointMat m = load_mat(mv, q);
  double* gps = sh + threadIdx.x;
  Point<N> pt;
  pt.setup(F, Fpi, R, sv.g + q, sv.ld, m, gps, sh + N * QBS + threadIdx.x, QBS);
  double s[6];
  int32_t st = 0;
  const int it = pt.solve(tab, cfg.tol, cfg.max_sub_step, cfg.max_iter, s, &st);
  double P[9];
  pk1<N>(pt, F, Fpi, s, P);
  if (TANGENT && A_out) {  // a separate instantiation: the tangent's registers cost the others occupancy
    tangent<N>(pt, tab, F, Fpi, s, [&](int m9, const double* dP) {
#pragma unroll
      for (int i = 0; i < 9; ++i) A_out[(9 * i + m9) * n_qp + q] = dP[i];
    });
  }
  if (P_out) {
#pragma unroll
    for (int i = 0; i < 9; ++i) P_out[i * n_qp + q] = P[i];
  }
  if (sig_out) {  // sigma = P F^T / det F (compute_avg_stress, models_copper.py:308-313)
    double sg[9];
    mm3t(P, F, sg);
    const double idF = 1.0 / det3(F);
#pragma unroll
    for (int i = 0; i < 9; ++i) sig_out[i * n_qp + q] = sg[i] * idF;
  }
  if (so.Fpi) {  // update_int_vars (models_copper.py:164-169,186-187)
    double Fn[9];
    fp_inv_new(Fpi, R, pt.Lpc, Fn);
#pragma unroll
    for (int i = 0; i < 9; ++i) so.Fpi[i * sv.ld + q] = Fn[i];
    double Sc[9];
    pt.crystal_stress(s, Sc);
    for (int a = 0; a < N; ++a) {
      double dg;
      pt.slip_increment(tab, Sc, a, &dg);
      gps[a * QBS] = dg;
      so.slip[a * sv.ld + q] = sv.slip[a * sv.ld + q] + dg;
    }
    hardening<N>(sv.g + q, sv.ld, gps, QBS, m, self_block, so.g + q, sv.ld);
  }
  if (iters) iters[q] = it;
  if (st) atomicOr(status, st);
}

// Cell average of sigma with JxW weights (models_copper.py:315)
__global__ void k_cell_average(Geom G, const double* __restrict__ sig, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.n_cells * 9) return;
  const int64_t cell = t / 9;
  const int comp = (int)(t - cell * 9);
  double num = 0.0, den = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double w = jxw_at(G, cell * 8 + q);
    num += sig[comp * G.n_qp + cell * 8 + q] * w;
    den += w;
  }
  out[cell * 9 + comp] = num / den;
}

// x0 of linear_incremental_solver (solver.py:221-227): ubar - u on Dirichlet dofs, 0 elsewhere
__global__ void k_bc_x0(int64_t n, const uint8_t* __restrict__ is_bc, const double* __restrict__ ubar,
                        const double* __restrict__ u, double* __restrict__ x0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x0[i] = is_bc[i] ? ubar[i] - u[i] : 0.0;
}

// ---------------------------------------------------------------------------------
// Launch helpers (host)
template <int N>
void launch_element_qp(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                       const SolveCfg& cfg, double* wP, double* sout, int32_t* iters, int32_t* status,
                       unsigned long long* counts, const PointPark& park, double* gi, hipStream_t st) {
  const int64_t nb = (G.n_qp + QBS - 1) / QBS;
  hipLaunchKernelGGL(k_element_res<N>, dim3((unsigned)nb), dim3(QBS), 0, st, G, u, sv, mv, tab, cfg, wP, sout, iters, status,
                     counts, park, gi);
}

template <int N>
void launch_element_qp_park(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                            const SolveCfg& cfg, double* wP, double* sout, int32_t* iters, int32_t* status,
                            unsigned long long* counts, const PointPark& park, hipStream_t st) {
  const int64_t nb = (park.cap + QBS - 1) / QBS;
  hipLaunchKernelGGL(k_element_qp_park<N>, dim3((unsigned)nb), dim3(QBS), 0, st, G, u, sv, mv, tab, cfg, wP, sout, iters,
                     status, counts, park);
}

template <int N>
void launch_point_tangent(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                          const double* sin, double* wA, hipStream_t st) {
  const int64_t nb = (G.n_qp + QBS - 1) / QBS;
  hipLaunchKernelGGL(k_point_tangent<N>, dim3((unsigned)nb), dim3(QBS), 0, st, G, u, sv, mv, tab, sin, wA);
}

template <int N>
void launch_local(int64_t n_qp, const double* Hd, const Geom& G, const double* u, const StateView& sv, const MatView& mv,
                  const SlipTable<N>& tab, const SolveCfg& cfg, int self_block, double* P, double* A, double* sig,
                  const StateOut& so, int32_t* iters, int32_t* status, hipStream_t st) {
  const int64_t nb = (n_qp + QBS - 1) / QBS;
  if (A)
    hipLaunchKernelGGL((k_local<N, true>), dim3((unsigned)nb), dim3(QBS), 0, st, n_qp, Hd, G, u, sv, mv, tab, cfg, self_block,
                       P, A, sig, so, iters, status);
  else
    hipLaunchKernelGGL((k_local<N, false>), dim3((unsigned)nb), dim3(QBS), 0, st, n_qp, Hd, G, u, sv, mv, tab, cfg,
                       self_block, P, A, sig, so, iters, status);
}

void launch_element_K(const Geom& G, const double* wP, const double* wA, double* KE, double* RE, hipStream_t st) {
  const int bs = 256;
  const int64_t nb = (G.n_qp + bs - 1) / bs;
  if (wA)
    hipLaunchKernelGGL(k_element_K, dim3((unsigned)((G.n_qp + EK_BS - 1) / EK_BS)), dim3(EK_BS), 0, st, G, wP, wA, KE, RE);
  else
    hipLaunchKernelGGL(k_element_R, dim3((unsigned)nb), dim3(bs), 0, st, G, wP, RE);
}

void launch_gather_matrix(const BellView& M, const double* KE, const int32_t* cptr, const int32_t* clist, const uint8_t* is_bc,
                          double* vals, hipStream_t st) {
  const int bs = 256;
  const int64_t n = 3 * M.n_nodes * M.n_slots;
  hipLaunchKernelGGL(k_gather_matrix, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, M, KE, cptr, clist, is_bc, vals);
}

void launch_gather_residual(int64_t n_nodes, const double* RE, const int32_t* nptr, const int32_t* nlist, const uint8_t* is_bc,
                            const double* ubar, const double* u, double* res, hipStream_t st) {
  const int bs = 256;
  const int64_t n = 3 * n_nodes;
  hipLaunchKernelGGL(k_gather_residual, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, n_nodes, RE, nptr, nlist, is_bc,
                     ubar, u, res);
}

void launch_bc_x0(int64_t n, const uint8_t* is_bc, const double* ubar, const double* u, double* x0, hipStream_t st) {
  hipLaunchKernelGGL(k_bc_x0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, is_bc, ubar, u, x0);
}

__global__ void k_zero_words(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

void launch_zero_words(uint32_t* p, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

void launch_cell_average(const Geom& G, const double* sig, double* out, hipStream_t st) {
  const int bs = 256;
  const int64_t n = G.n_cells * 9;
  hipLaunchKernelGGL(k_cell_average, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, st, G, sig, out);
}

#define CPFEM_INST(N)                                                                                                      \
  template void launch_element_qp<N>(const Geom&, const double*, const StateView&, const MatView&, const SlipTable<N>&,    \
                                     const SolveCfg&, double*, double*, int32_t*, int32_t*, unsigned long long*,           \
                                     const PointPark&, double*, hipStream_t);                                              \
  template void launch_element_qp_park<N>(const Geom&, const double*, const StateView&, const MatView&,                    \
                                          const SlipTable<N>&, const SolveCfg&, double*, double*, int32_t*, int32_t*,      \
                                          unsigned long long*, const PointPark&, hipStream_t);                             \
  template void launch_point_tangent<N>(const Geom&, const double*, const StateView&, const MatView&, const SlipTable<N>&, \
                                        const double*, double*, hipStream_t);                                              \
  template void launch_local<N>(int64_t, const double*, const Geom&, const double*, const StateView&, const MatView&,      \
                                const SlipTable<N>&, const SolveCfg&, int, double*, double*, double*, const StateOut&,     \
                                int32_t*, int32_t*, hipStream_t);
CPFEM_INST(12)
CPFEM_INST(24)

}  // namespace cpfem
