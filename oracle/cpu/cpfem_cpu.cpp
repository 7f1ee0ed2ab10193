// C++/OpenMP restatement of the reference algorithm on the host: the CPU baseline that the
// benchmark times beside the GPU.  TEST INFRASTRUCTURE ONLY: it is used by the benchmark's
// CPU-baseline leg (and its CPU tests), never by the product path.
//
// What it computes, on all host threads (OpenMP):
//  * the per-point Kalidindi update with consistent tangent (the local Newton from S = 0 with
//    cut-half backtracking of singlecrystal_copper/models_copper.py:204-249, the IFT tangent of
//    :251-259, P of :155-162) — the host build of the same per-point arithmetic the HIP kernels
//    compile (jax-cpfem_amd/csrc/cp_point.hpp), parallel over cells;
//  * the hex8 element residual and stiffness of the upstream FE kernel used at
//    crystal_plasticity_OR_design/solver.py:392 (r_e = sum_q P dN JxW,
//    K_e = sum_q dN A dN JxW), then the CSR assembly with summed duplicates and the row
//    elimination of get_A (solver.py:279-304) and apply_bc_vec (:119-133): each scalar row is
//    owned by one thread, which adds its incident cells' blocks in ascending cell order (no
//    atomics);
//  * Jacobi-preconditioned BiCGSTAB with the iteration of jax.scipy.sparse.linalg.bicgstab
//    (solver.py:19-48) on the CSR matrix;
//  * the per-point VJP of the adjoint (jax-cpfem_amd/csrc/cp_dual.hpp's host build), to time
//    the point part of the hand-written adjoint's CPU cost.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../jax-cpfem_amd/csrc/cp_point.hpp"
#include "../../jax-cpfem_amd/csrc/cp_dual.hpp"

using namespace cpfem;

namespace {

const double HEXN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct Ctx {
  int64_t nn = 0, nc = 0, nq = 0, nd = 0;
  int n_slip = 0;
  std::vector<int32_t> cells;   // (nc, 8)
  std::vector<double> sg, jxw;  // (nc, 8q, 8a, 3), (nc, 8q)
  std::vector<int64_t> indptr;  // CSR over scalar rows, columns sorted
  std::vector<int32_t> indices;
  std::vector<int32_t> nptr, nlist;  // node -> (cell * 8 + local a), ascending cell
  std::vector<uint8_t> slot;         // (nc, 8a, 8b): position of node b in node a's sorted neighbour list
  std::vector<int32_t> dpos;         // (nd): position of the diagonal in its row
  std::vector<double> KE, RE;        // (nc, 24, 24), (nc, 24)
  SlipTable<12> t12{};
  SlipTable<24> t24{};
};

void geometry(Ctx& c, const double* pts) {
  const double g[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  double dN[8][8][3];
  int q = 0;
  for (int a0 = 0; a0 < 2; ++a0)
    for (int a1 = 0; a1 < 2; ++a1)
      for (int a2 = 0; a2 < 2; ++a2, ++q) {
        const double xi[3] = {g[a0], g[a1], g[a2]};
        for (int a = 0; a < 8; ++a) {
          double f[3], df[3];
          for (int d = 0; d < 3; ++d) {
            f[d] = HEXN[a][d] == 1.0 ? xi[d] : 1.0 - xi[d];
            df[d] = HEXN[a][d] == 1.0 ? 1.0 : -1.0;
          }
          dN[q][a][0] = df[0] * f[1] * f[2];
          dN[q][a][1] = f[0] * df[1] * f[2];
          dN[q][a][2] = f[0] * f[1] * df[2];
        }
      }
  c.sg.assign((size_t)c.nc * 192, 0.0);
  c.jxw.assign((size_t)c.nc * 8, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < c.nc; ++e) {
    double X[8][3];
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) X[a][d] = pts[3 * (int64_t)c.cells[8 * e + a] + d];
    for (int qq = 0; qq < 8; ++qq) {
      double J[9] = {0};
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) J[3 * i + j] += X[a][i] * dN[qq][a][j];
      const double det = det3(J);
      double Ji[9];
      inv3(J, det, Ji);
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
          for (int j = 0; j < 3; ++j) s += dN[qq][a][j] * Ji[3 * j + i];
          c.sg[((e * 8 + qq) * 8 + a) * 3 + i] = s;
        }
      c.jxw[e * 8 + qq] = det * 0.125;
    }
  }
}

void pattern(Ctx& c) {
  c.nptr.assign(c.nn + 1, 0);
  for (int64_t e = 0; e < c.nc; ++e)
    for (int a = 0; a < 8; ++a) c.nptr[c.cells[8 * e + a] + 1]++;
  for (int64_t i = 0; i < c.nn; ++i) c.nptr[i + 1] += c.nptr[i];
  c.nlist.assign((size_t)c.nc * 8, 0);
  {
    std::vector<int32_t> fill(c.nptr.begin(), c.nptr.end() - 1);
    for (int64_t e = 0; e < c.nc; ++e)
      for (int a = 0; a < 8; ++a) c.nlist[fill[c.cells[8 * e + a]]++] = (int32_t)(8 * e + a);
  }
  std::vector<std::vector<int32_t>> nb(c.nn);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nn; ++i) {
    auto& v = nb[i];
    for (int32_t k = c.nptr[i]; k < c.nptr[i + 1]; ++k) {
      const int64_t e = c.nlist[k] / 8;
      for (int b = 0; b < 8; ++b) v.push_back(c.cells[8 * e + b]);
    }
    if (v.empty()) v.push_back((int32_t)i);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  c.indptr.assign(c.nd + 1, 0);
  for (int64_t i = 0; i < c.nn; ++i)
    for (int k = 0; k < 3; ++k) c.indptr[3 * i + k + 1] = 3 * (int64_t)nb[i].size();
  for (int64_t r = 0; r < c.nd; ++r) c.indptr[r + 1] += c.indptr[r];
  c.indices.assign(c.indptr[c.nd], 0);
  c.dpos.assign(c.nd, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nn; ++i) {
    const auto& v = nb[i];
    for (int k = 0; k < 3; ++k) {
      const int64_t r = 3 * i + k;
      int64_t o = c.indptr[r];
      for (size_t s = 0; s < v.size(); ++s)
        for (int kk = 0; kk < 3; ++kk) {
          if (v[s] == i && kk == k) c.dpos[r] = (int32_t)(o - c.indptr[r]);
          c.indices[o++] = 3 * v[s] + kk;
        }
    }
  }
  c.slot.assign((size_t)c.nc * 64, 0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < c.nc; ++e)
    for (int a = 0; a < 8; ++a) {
      const auto& v = nb[c.cells[8 * e + a]];
      for (int b = 0; b < 8; ++b)
        c.slot[e * 64 + 8 * a + b] = (uint8_t)(std::lower_bound(v.begin(), v.end(), c.cells[8 * e + b]) - v.begin());
    }
}

struct State {
  const double* Fpi;  // [9][Q]
  const double* g;    // [N][Q]
  const double* R;    // [9][Q]
  const double* mat;  // [8][Q]: xm, gss_a, h, t_sat, r, C11, C12, C44
  double ao, dt, tol;
  int max_sub, max_iter, self_block;
};

template <int N>
const SlipTable<N>& table(const Ctx& c);
template <>
const SlipTable<12>& table<12>(const Ctx& c) { return c.t12; }
template <>
const SlipTable<24>& table<24>(const Ctx& c) { return c.t24; }

inline PointMat point_mat(const State& s, int64_t Q, int64_t q) {
  PointMat m;
  m.p = 1.0 / s.mat[0 * Q + q];
  m.c = s.ao * s.dt;
  m.gss_a = s.mat[1 * Q + q];
  m.h = s.mat[2 * Q + q];
  m.tsat = s.mat[3 * Q + q];
  m.r = s.mat[4 * Q + q];
  m.C11 = s.mat[5 * Q + q];
  m.C12 = s.mat[6 * Q + q];
  m.C44 = s.mat[7 * Q + q];
  return m;
}

// u_grad at point (e, q) from the nodal dofs
inline void ugrad(const Ctx& c, const double* u, int64_t e, int q, double* H) {
  for (int i = 0; i < 9; ++i) H[i] = 0.0;
  const double* g = &c.sg[(e * 8 + q) * 24];
  for (int a = 0; a < 8; ++a) {
    const double* ua = u + 3 * (int64_t)c.cells[8 * e + a];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) H[3 * i + j] += ua[i] * g[3 * a + j];
  }
}

inline void gather_point(const State& s, int64_t Q, int64_t qp, int N, double* Fpi, double* R, double* g) {
  for (int i = 0; i < 9; ++i) {
    Fpi[i] = s.Fpi[i * Q + qp];
    R[i] = s.R[i * Q + qp];
  }
  for (int a = 0; a < N; ++a) g[a] = s.g[a * Q + qp];
}

// Per-cell element residual and (if tangent) stiffness; iteration counters.
template <int N>
void element_pass(Ctx& c, const double* u, const State& s, bool tangent_on, int64_t* counts) {
  const SlipTable<N>& tab = table<N>(c);
  const int64_t Q = c.nq;
  int64_t it_sum = 0, ev_sum = 0;
#pragma omp parallel reduction(+ : it_sum, ev_sum)
  {
    double gp[N], gi[N], gold[N];
#pragma omp for schedule(dynamic, 16)
    for (int64_t e = 0; e < c.nc; ++e) {
      double Ke[576];
      double re[24];
      std::memset(Ke, 0, sizeof(Ke));
      std::memset(re, 0, sizeof(re));
      for (int q = 0; q < 8; ++q) {
        const int64_t qp = 8 * e + q;
        double H[9], F[9], Fpi[9], R[9], P[9], A[81];
        ugrad(c, u, e, q, H);
        for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
        gather_point(s, Q, qp, N, Fpi, R, gold);
        const PointMat m = point_mat(s, Q, qp);
        Point<N> pt;
        pt.setup(F, Fpi, R, gold, 1, m, gp, gi, 1);
        double S[6];
        int32_t st = 0;
        it_sum += pt.solve(tab, s.tol, s.max_sub, s.max_iter, S, &st);
        ev_sum += pt.nev;
        pk1<N>(pt, F, Fpi, S, P);
        const double w = c.jxw[qp];
        const double* dn = &c.sg[qp * 24];
        for (int a = 0; a < 8; ++a)
          for (int i = 0; i < 3; ++i)
            re[3 * a + i] += w * (P[3 * i] * dn[3 * a] + P[3 * i + 1] * dn[3 * a + 1] + P[3 * i + 2] * dn[3 * a + 2]);
        if (!tangent_on) continue;
        tangent<N>(pt, tab, F, Fpi, S, [&](int m9, const double* dP) {
          for (int i = 0; i < 9; ++i) A[9 * i + m9] = w * dP[i];
        });
        // K[a i][b k] += sum_{j,l} dN_a,j A[(i j),(k l)] dN_b,l
        for (int a = 0; a < 8; ++a)
          for (int i = 0; i < 3; ++i) {
            double Z[9];
            for (int kl = 0; kl < 9; ++kl)
              Z[kl] = dn[3 * a] * A[9 * (3 * i) + kl] + dn[3 * a + 1] * A[9 * (3 * i + 1) + kl] +
                      dn[3 * a + 2] * A[9 * (3 * i + 2) + kl];
            double* Krow = Ke + (3 * a + i) * 24;
            for (int b = 0; b < 8; ++b)
              for (int k = 0; k < 3; ++k)
                Krow[3 * b + k] += Z[3 * k] * dn[3 * b] + Z[3 * k + 1] * dn[3 * b + 1] + Z[3 * k + 2] * dn[3 * b + 2];
          }
      }
      if (tangent_on) std::memcpy(&c.KE[e * 576], Ke, sizeof(Ke));
      std::memcpy(&c.RE[e * 24], re, sizeof(re));
    }
  }
  if (counts) {
    counts[0] = it_sum;
    counts[1] = ev_sum;
  }
}

// Row-owner CSR gather with the row elimination of Dirichlet rows; residual with
// apply_bc_vec.
void gather(Ctx& c, const double* u, const uint8_t* is_bc, const double* ubar, double* res, double* vals, bool want_mat) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c.nn; ++i) {
    for (int ci = 0; ci < 3; ++ci) {
      const int64_t r = 3 * i + ci;
      double* row = vals + c.indptr[r];
      const int64_t len = c.indptr[r + 1] - c.indptr[r];
      double rv = 0.0;
      if (want_mat) std::fill(row, row + len, 0.0);
      if (is_bc[r]) {
        if (want_mat) row[c.dpos[r]] = 1.0;
        res[r] = u[r] - ubar[r];
        continue;
      }
      for (int32_t k = c.nptr[i]; k < c.nptr[i + 1]; ++k) {
        const int64_t e = c.nlist[k] / 8;
        const int a = c.nlist[k] % 8;
        rv += c.RE[e * 24 + 3 * a + ci];
        if (!want_mat) continue;
        const double* Krow = &c.KE[e * 576 + (3 * a + ci) * 24];
        const uint8_t* sl = &c.slot[e * 64 + 8 * a];
        for (int b = 0; b < 8; ++b)
          for (int kk = 0; kk < 3; ++kk) row[3 * sl[b] + kk] += Krow[3 * b + kk];
      }
      res[r] = rv;
    }
  }
}

inline void spmv(const Ctx& c, const double* vals, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < c.nd; ++r) {
    double s = 0.0;
    for (int64_t k = c.indptr[r]; k < c.indptr[r + 1]; ++k) s += vals[k] * x[c.indices[k]];
    y[r] = s;
  }
}

}  // namespace

extern "C" {

int cpu_threads(void) { return omp_get_max_threads(); }

void* cpu_create(int64_t n_nodes, int64_t n_cells, const double* points, const int32_t* cells, int n_slip, const double* dirs,
                 const double* normals) {
  if (n_slip != 12 && n_slip != 24) return nullptr;
  Ctx* c = new Ctx();
  c->nn = n_nodes;
  c->nc = n_cells;
  c->nq = 8 * n_cells;
  c->nd = 3 * n_nodes;
  c->n_slip = n_slip;
  c->cells.assign(cells, cells + 8 * n_cells);
  for (int a = 0; a < n_slip; ++a)
    for (int d = 0; d < 3; ++d) {
      if (n_slip == 12) {
        c->t12.d[a][d] = dirs[3 * a + d];
        c->t12.n[a][d] = normals[3 * a + d];
      } else {
        c->t24.d[a][d] = dirs[3 * a + d];
        c->t24.n[a][d] = normals[3 * a + d];
      }
    }
  geometry(*c, points);
  pattern(*c);
  c->KE.assign((size_t)c->nc * 576, 0.0);
  c->RE.assign((size_t)c->nc * 24, 0.0);
  return c;
}

void cpu_destroy(void* h) { delete static_cast<Ctx*>(h); }

int64_t cpu_nnz(void* h) { return static_cast<Ctx*>(h)->indptr.back(); }

int cpu_csr(void* h, int64_t* indptr, int32_t* indices) {
  const Ctx* c = static_cast<Ctx*>(h);
  std::memcpy(indptr, c->indptr.data(), c->indptr.size() * sizeof(int64_t));
  std::memcpy(indices, c->indices.data(), c->indices.size() * sizeof(int32_t));
  return 0;
}

// Residual (n_dofs) and, if vals, the CSR Jacobian with Dirichlet row elimination, at dofs u
// with the history (Fpi, g) and rotations / parameters of the points (SoA [comp][Q]).
// counts[0..1] = total local Newton iterations, residual evaluations.
int cpu_assemble(void* h, const double* u, const double* Fpi, const double* g, const double* R, const double* mat, double ao,
                 double dt, double tol, int max_sub, int max_iter, int self_block, const uint8_t* is_bc, const double* ubar,
                 double* res, double* vals, int64_t* counts) {
  Ctx* c = static_cast<Ctx*>(h);
  State s{Fpi, g, R, mat, ao, dt, tol, max_sub, max_iter, self_block};
  if (c->n_slip == 12)
    element_pass<12>(*c, u, s, vals != nullptr, counts);
  else
    element_pass<24>(*c, u, s, vals != nullptr, counts);
  gather(*c, u, is_bc, ubar, res, vals, vals != nullptr);
  return 0;
}

// Point pass without tangent: Cauchy stress per point and the history update
// (compute_avg_stress and update_int_vars_gp, models_copper.py:273-319).
int cpu_point_pass(void* h, const double* u, const double* Fpi, const double* g, const double* R, const double* mat, double ao,
                   double dt, double tol, int max_sub, int max_iter, int self_block, double* sigma, double* Fpi_new,
                   double* g_new) {
  Ctx* c = static_cast<Ctx*>(h);
  State s{Fpi, g, R, mat, ao, dt, tol, max_sub, max_iter, self_block};
  const int64_t Q = c->nq;
  auto run = [&](auto tag) {
    constexpr int N = decltype(tag)::value;
    const SlipTable<N>& tab = table<N>(*c);
#pragma omp parallel
    {
      double gp[N], gi[N], gold[N], tmp[N], gn[N];
#pragma omp for schedule(dynamic, 128)
      for (int64_t qp = 0; qp < Q; ++qp) {
        double H[9], F[9], Fi[9], Rq[9], P[9], Fn[9];
        ugrad(*c, u, qp / 8, (int)(qp % 8), H);
        for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
        gather_point(s, Q, qp, N, Fi, Rq, gold);
        const PointMat m = point_mat(s, Q, qp);
        Point<N> pt;
        pt.setup(F, Fi, Rq, gold, 1, m, gp, gi, 1);
        double S[6];
        int32_t st = 0;
        pt.solve(tab, s.tol, s.max_sub, s.max_iter, S, &st);
        pk1<N>(pt, F, Fi, S, P);
        double sg[9];
        mm3t(P, F, sg);
        const double idF = 1.0 / det3(F);
        for (int i = 0; i < 9; ++i) sigma[i * Q + qp] = sg[i] * idF;
        fp_inv_new(Fi, pt.R, pt.Lpc, Fn);
        for (int i = 0; i < 9; ++i) Fpi_new[i * Q + qp] = Fn[i];
        double Sc[9];
        pt.crystal_stress(S, Sc);
        for (int a = 0; a < N; ++a) pt.slip_increment(tab, Sc, a, &tmp[a]);
        hardening<N>(gold, 1, tmp, 1, m, s.self_block, gn, 1);
        for (int a = 0; a < N; ++a) g_new[a * Q + qp] = gn[a];
      }
    }
  };
  if (c->n_slip == 12)
    run(std::integral_constant<int, 12>());
  else
    run(std::integral_constant<int, 24>());
  return 0;
}

// Per-point VJP pass of the adjoint (cotangent Pbar [9][Q] on P, all input groups out),
// as the hand-written adjoint's second point VJP of each load step.
int cpu_vjp_pass(void* h, const double* u, const double* Fpi, const double* g, const double* R, const double* mat, double ao,
                 double dt, double tol, int max_sub, int max_iter, int self_block, const double* Pbar, double* Hbar,
                 double* pbar) {
  Ctx* c = static_cast<Ctx*>(h);
  State s{Fpi, g, R, mat, ao, dt, tol, max_sub, max_iter, self_block};
  const int64_t Q = c->nq;
  auto run = [&](auto tag) {
    constexpr int N = decltype(tag)::value;
    const SlipTable<N>& tab = table<N>(*c);
#pragma omp parallel
    {
      double gp[N], gi[N], gold[N], zN[N], Jl[42];
#pragma omp for schedule(dynamic, 128)
      for (int64_t qp = 0; qp < Q; ++qp) {
        double H[9], F[9], Fi[9], Rq[9], mq[8], Pb[9];
        const double z9[9] = {0};
        for (int a = 0; a < N; ++a) zN[a] = 0.0;
        ugrad(*c, u, qp / 8, (int)(qp % 8), H);
        for (int i = 0; i < 9; ++i) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
        gather_point(s, Q, qp, N, Fi, Rq, gold);
        for (int k = 0; k < 8; ++k) mq[k] = s.mat[k * Q + qp];
        for (int i = 0; i < 9; ++i) Pb[i] = Pbar[i * Q + qp];
        const PointMat m = point_mat(s, Q, qp);
        Point<N> pt;
        pt.setup(F, Fi, Rq, gold, 1, m, gp, gi, 1);
        double S[6];
        int32_t st = 0;
        pt.solve(tab, s.tol, s.max_sub, s.max_iter, S, &st);
        SeededIn<N> xin{H, Fi, gold, 1, mq, Rq, -1};
        point_vjp<N>(xin, S, tab, s.ao, s.dt, s.self_block, Pb, z9, z9, zN, zN, 1, Jl, 1, 31,
                     [&](int grp, int off, double xb) {
                       if (grp == 1) Hbar[off * Q + qp] = xb;
                       if (grp == 8) pbar[off * Q + qp] = xb;
                     });
      }
    }
  };
  if (c->n_slip == 12)
    run(std::integral_constant<int, 12>());
  else
    run(std::integral_constant<int, 24>());
  return 0;
}

// Jacobi-BiCGSTAB with the iteration of jax.scipy.sparse.linalg.bicgstab on the CSR matrix
// (x holds x0).  Stops when <r,r> <= max(tol^2 <b,b>, atol^2) or after maxiter iterations;
// returns the iteration counter (negative on breakdown, as the reference's solver).
int cpu_bicgstab(void* h, const double* vals, const double* b, double* x, double tol, double atol, int maxiter) {
  Ctx* c = static_cast<Ctx*>(h);
  const int64_t n = c->nd;
  std::vector<double> r(n), rhat(n), p(n), q(n), phat(n), s(n), shat(n), t(n), invd(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) invd[i] = 1.0 / vals[c->indptr[i] + c->dpos[i]];
  spmv(*c, vals, x, q.data());
  double bs = 0.0, rs = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : bs, rs)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    rhat[i] = p[i] = q[i] = r[i];
    bs += b[i] * b[i];
    rs += r[i] * r[i];
  }
  const double atol2 = std::max(tol * tol * bs, atol * atol);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int k = 0;
  while (rs > atol2 && k < maxiter && k >= 0) {
    double rho_ = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rho_)
    for (int64_t i = 0; i < n; ++i) rho_ += rhat[i] * r[i];
    const double beta = rho_ / rho * alpha / omega;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      p[i] = r[i] + beta * (p[i] - omega * q[i]);
      phat[i] = p[i] * invd[i];
    }
    double rq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rq)
    for (int64_t i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int64_t kk = c->indptr[i]; kk < c->indptr[i + 1]; ++kk) acc += vals[kk] * phat[c->indices[kk]];
      q[i] = acc;
      rq += rhat[i] * acc;
    }
    const double alpha_ = rho_ / rq;
    double ss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ss)
    for (int64_t i = 0; i < n; ++i) {
      s[i] = r[i] - alpha_ * q[i];
      shat[i] = s[i] * invd[i];
      ss += s[i] * s[i];
    }
    const bool early = ss < atol2;
    double ts = 0.0, tt = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ts, tt)
    for (int64_t i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int64_t kk = c->indptr[i]; kk < c->indptr[i + 1]; ++kk) acc += vals[kk] * shat[c->indices[kk]];
      t[i] = acc;
      ts += acc * s[i];
      tt += acc * acc;
    }
    const double omega_ = ts / tt;
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
    for (int64_t i = 0; i < n; ++i) {
      if (early) {
        x[i] += alpha_ * phat[i];
        r[i] = s[i];
      } else {
        x[i] += alpha_ * phat[i] + omega_ * shat[i];
        r[i] = s[i] - omega_ * t[i];
      }
      rr += r[i] * r[i];
    }
    int kn = (omega_ == 0.0 || alpha_ == 0.0) ? -11 : k + 1;
    if (rho_ == 0.0) kn = -10;
    k = kn;
    rho = rho_;
    alpha = alpha_;
    omega = omega_;
    rs = rr;
  }
  return k;
}

}  // extern "C"
