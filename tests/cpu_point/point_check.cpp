// CPU build of the kernel's per-point arithmetic (jax-cpfem_amd/csrc/cp_point.hpp) for the
// CPU test suite: the same source the HIP kernels compile, run on host doubles, so the
// 6-unknown reduction, the closed-form rotated stiffness and the Ycal tangent can be
// checked against the oracle without a GPU.  Test infrastructure only.
#include <stdint.h>

#include <vector>

#include "../../jax-cpfem_amd/csrc/cp_point.hpp"

using namespace cpfem;

template <int N>
static void run(int64_t Q, const double* H, const double* Fpi, const double* g, const double* R, const double* mat,
                const double* dirs, const double* nrms, double ao, double dt, double tol, int max_sub, int max_iter,
                int self_block, double* P, double* A, double* Fpi_new, double* g_new, double* dgam, int32_t* iters,
                int32_t* status) {
  SlipTable<N> tab;
  for (int a = 0; a < N; ++a)
    for (int d = 0; d < 3; ++d) {
      tab.d[a][d] = dirs[3 * a + d];
      tab.n[a][d] = nrms[3 * a + d];
    }
  std::vector<double> gp(N), gi(N);
  for (int64_t q = 0; q < Q; ++q) {
    PointMat m;
    const double* mq = mat + 8 * q;
    m.p = 1.0 / mq[0];
    m.c = ao * dt;
    m.gss_a = mq[1];
    m.h = mq[2];
    m.tsat = mq[3];
    m.r = mq[4];
    m.C11 = mq[5];
    m.C12 = mq[6];
    m.C44 = mq[7];
    double F[9];
    for (int i = 0; i < 9; ++i) F[i] = H[9 * q + i] + ((i % 4) == 0 ? 1.0 : 0.0);
    Point<N> pt;
    pt.setup(F, Fpi + 9 * q, R + 9 * q, g + N * q, 1, m, gp.data(), gi.data(), 1);
    double s[6];
    int32_t st = 0;
    iters[q] = pt.solve(tab, tol, max_sub, max_iter, s, &st);
    status[q] = st;
    pk1<N>(pt, F, Fpi + 9 * q, s, P + 9 * q);
    if (A) {
      double* Aq = A + 81 * q;
      tangent<N>(pt, tab, F, Fpi + 9 * q, s, [&](int m9, const double* dP) {
        for (int i = 0; i < 9; ++i) Aq[9 * i + m9] = dP[i];
      });
    }
    fp_inv_new(Fpi + 9 * q, pt.R, pt.Lpc, Fpi_new + 9 * q);
    double Sc[9];
    pt.crystal_stress(s, Sc);
    std::vector<double> tmp(N);
    for (int a = 0; a < N; ++a) {
      pt.slip_increment(tab, Sc, a, &tmp[a]);
      dgam[N * q + a] = tmp[a];
    }
    hardening<N>(g + N * q, 1, tmp.data(), 1, m, self_block, g_new + N * q, 1);
  }
}

extern "C" int cpu_point_local(int n_slip, int64_t Q, const double* H, const double* Fpi, const double* g, const double* R,
                               const double* mat, const double* dirs, const double* nrms, double ao, double dt, double tol,
                               int max_sub, int max_iter, int self_block, double* P, double* A, double* Fpi_new,
                               double* g_new, double* dgam, int32_t* iters, int32_t* status) {
  if (n_slip == 12)
    run<12>(Q, H, Fpi, g, R, mat, dirs, nrms, ao, dt, tol, max_sub, max_iter, self_block, P, A, Fpi_new, g_new, dgam, iters,
            status);
  else if (n_slip == 24)
    run<24>(Q, H, Fpi, g, R, mat, dirs, nrms, ao, dt, tol, max_sub, max_iter, self_block, P, A, Fpi_new, g_new, dgam, iters,
            status);
  else
    return 1;
  return 0;
}

#include "../../jax-cpfem_amd/csrc/cp_dual.hpp"

// Per-point VJP on the CPU (same code path as the device kernel).  Cotangents per point:
// Pb, sb (point Cauchy stress), Fb [Q][9], gb, slb [Q][N]; outputs Hb, Fpib [Q][9],
// gob [Q][N], pb [Q][8], Rb [Q][9].
template <int N>
static void run_vjp(int64_t Q, const double* H, const double* Fpi, const double* g, const double* R, const double* mat,
                    const double* dirs, const double* nrms, double ao, double dt, double tol, int max_sub, int max_iter,
                    int self_block, const double* Pb, const double* sb, const double* Fb, const double* gb,
                    const double* slb, double* Hb, double* Fpib, double* gob, double* pb, double* Rb) {
  SlipTable<N> tab;
  for (int a = 0; a < N; ++a)
    for (int d = 0; d < 3; ++d) {
      tab.d[a][d] = dirs[3 * a + d];
      tab.n[a][d] = nrms[3 * a + d];
    }
  std::vector<double> gp(N), gi(N), Jl(42);
  for (int64_t q = 0; q < Q; ++q) {
    PointMat m;
    const double* mq = mat + 8 * q;
    m.p = 1.0 / mq[0];
    m.c = ao * dt;
    m.gss_a = mq[1];
    m.h = mq[2];
    m.tsat = mq[3];
    m.r = mq[4];
    m.C11 = mq[5];
    m.C12 = mq[6];
    m.C44 = mq[7];
    double F[9];
    for (int i = 0; i < 9; ++i) F[i] = H[9 * q + i] + ((i % 4) == 0 ? 1.0 : 0.0);
    Point<N> pt;
    pt.setup(F, Fpi + 9 * q, R + 9 * q, g + N * q, 1, m, gp.data(), gi.data(), 1);
    double s[6];
    int32_t st = 0;
    pt.solve(tab, tol, max_sub, max_iter, s, &st);
    SeededIn<N> xin{H + 9 * q, Fpi + 9 * q, g + N * q, 1, mq, R + 9 * q, -1};
    point_vjp<N>(xin, s, tab, ao, dt, self_block, Pb + 9 * q, sb + 9 * q, Fb + 9 * q, gb + N * q, slb + N * q, 1, Jl.data(),
                 1, 31, [&](int grp, int off, double xb) {
                   double* out = grp == 1 ? Hb + 9 * q : grp == 2 ? Fpib + 9 * q : grp == 4 ? gob + N * q
                                 : grp == 8 ? pb + 8 * q : Rb + 9 * q;
                   out[off] = xb;
                 });
  }
}

extern "C" int cpu_point_vjp(int n_slip, int64_t Q, const double* H, const double* Fpi, const double* g, const double* R,
                             const double* mat, const double* dirs, const double* nrms, double ao, double dt, double tol,
                             int max_sub, int max_iter, int self_block, const double* Pb, const double* sb, const double* Fb,
                             const double* gb, const double* slb, double* Hb, double* Fpib, double* gob, double* pb,
                             double* Rb) {
  if (n_slip == 12)
    run_vjp<12>(Q, H, Fpi, g, R, mat, dirs, nrms, ao, dt, tol, max_sub, max_iter, self_block, Pb, sb, Fb, gb, slb, Hb, Fpib,
                gob, pb, Rb);
  else if (n_slip == 24)
    run_vjp<24>(Q, H, Fpi, g, R, mat, dirs, nrms, ao, dt, tol, max_sub, max_iter, self_block, Pb, sb, Fb, gb, slb, Hb, Fpib,
                gob, pb, Rb);
  else
    return 1;
  return 0;
}

// The kernel's logarithm, exponential and power (cp_point.hpp) on n arguments.
extern "C" void cpu_fastmath(int64_t n, const double* x, const double* p, double* ln, double* ex, double* pw) {
  for (int64_t i = 0; i < n; ++i) {
    ln[i] = cpfem::ln_pos(x[i]);
    ex[i] = cpfem::exp_arg(x[i]);
    pw[i] = cpfem::pow_pos(x[i], p[i]);
  }
}
