// Forward-mode dual numbers and the per-point map written once for any scalar type, used by
// the reverse-mode (adjoint) kernel: with T = Dual, one evaluation gives the derivative of the
// residual and of every output in one seeded input direction.
//
// local_fn follows singlecrystal_copper/models_copper.py:155-200 (and the per-point-parameter
// variants calibration/case1.py:152-200, models_DPsteel_inhomo.py:240-294):
//   tau_a = S : R P0_a R^T; dgam_a = gamma0 dt |tau_a/g_a|^(1/m) sign(tau_a)
//   Lp = sum dgam_a R P0_a R^T; Fp_inv_new = Fp_inv_old (I - Lp); Fe = F Fp_inv_new
//   r = S - C_rot : 1/2 (Fe^T Fe - I)
//   P = F Fp_inv_new S Fp_inv_new^T / det Fp_inv_new (= det F (Fe S Fe^T / det Fe) F^-T)
//   sigma = P F^T / det F
//   g_new = g + q (h |dgam| |1 - g/gsat|^a sign(1 - g/gsat))
#pragma once
#include "cp_point.hpp"

namespace cpfem {

struct Dual {
  double v, d;
  CPFEM_HD Dual(double v_ = 0.0, double d_ = 0.0) : v(v_), d(d_) {}
};

constexpr int NPAR = 8;  // xm, gss_a, h, t_sat, r, C11, C12, C44
CPFEM_HD Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
CPFEM_HD Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
CPFEM_HD Dual operator-(Dual a) { return {-a.v, -a.d}; }
CPFEM_HD Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
CPFEM_HD Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
CPFEM_HD Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
CPFEM_HD Dual operator/(Dual a, Dual b) {
  const double iv = 1.0 / b.v;
  return {a.v * iv, (a.d - a.v * iv * b.d) * iv};
}
CPFEM_HD Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
CPFEM_HD Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
CPFEM_HD Dual& operator+=(Dual& a, Dual b) {
  a.v += b.v;
  a.d += b.d;
  return a;
}
CPFEM_HD Dual dabs(Dual a) { return a.v < 0.0 ? Dual{-a.v, -a.d} : (a.v > 0.0 ? a : Dual{0.0, 0.0}); }
// x^y for x >= 0 with the conventions of lax.pow's jvp: d/dx = y x^(y-1) (0 at x = 0 for y > 1),
// d/dy = x^y log x (0 at x = 0).
// For x > 0 the value is e^(y ln x) (ln_pos, exp_arg of cp_point.hpp) with the logarithm
// shared by both derivative terms.
CPFEM_HD Dual dpow(Dual x, Dual y) {
  if (x.v > 0.0) {
    const double lx = ln_pos(x.v);
    const double pv = exp_arg(y.v * lx);
    return {pv, pv * (y.d * lx + y.v * x.d / x.v)};
  }
  return {pow(x.v, y.v), y.v == 1.0 ? x.d : 0.0};
}
CPFEM_HD double sgnv(Dual a) { return sgn(a.v); }

CPFEM_HD double dval(double a) { return a; }
CPFEM_HD double dval(Dual a) { return a.v; }
CPFEM_HD double dabs(double a) { return fabs(a); }
CPFEM_HD double dpow(double x, double y) { return x > 0.0 ? pow_pos(x, y) : pow(x, y); }
CPFEM_HD double sgnv(double a) { return sgn(a); }

template <typename T>
CPFEM_HD void tmm3(const T* a, const T* b, T* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
template <typename T>
CPFEM_HD void tmm3t(const T* a, const T* b, T* c) {  // a b^T
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[3 * j] + a[3 * i + 1] * b[3 * j + 1] + a[3 * i + 2] * b[3 * j + 2];
}
template <typename T>
CPFEM_HD void tmtm3(const T* a, const T* b, T* c) {  // a^T b
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[i] * b[j] + a[3 + i] * b[3 + j] + a[6 + i] * b[6 + j];
}
template <typename T>
CPFEM_HD T tdet3(const T* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Input directions (used to seed derivatives): H[0..8], Fp_inv_old[9..17], g_old[18..18+N),
// xm, gss_a, h, t_sat, r, C11, C12, C44 [18+N .. 25+N], R[26+N .. 34+N].
// Inputs are read through an accessor In with H(i), Fpi(i), g(a), mp(m), R(i) returning T;
// outputs go to a sink Out with r(i, T), P(i, T), sig(i, T), Fpi(i, T), g(a, T), dgam(a, T), so neither
// side needs to hold all of them at once.
template <typename T, int N, class In, class Out>
CPFEM_HD void local_fn(const In& x, const T* s, const SlipTable<N>& tab, double ao, double dt, int self_block, Out& o) {
  T F[9], R[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    F[i] = x.H(i) + ((i % 4) == 0 ? 1.0 : 0.0);
    R[i] = x.R(i);
  }
  // crystal-frame stress R^T S R
  T S[9] = {s[0], s[5], s[4], s[5], s[1], s[3], s[4], s[3], s[2]};
  T Sc[9];
  {
    T U[9];
    tmm3(S, R, U);
    tmtm3(R, U, Sc);
  }
  const T p = T(1.0) / x.mp(0);
  const double c = ao * dt;
  T Lpc[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) Lpc[i] = T(0.0);
  T tot = T(0.0);
  T blk[N / 3];
#pragma unroll
  for (int b = 0; b < N / 3; ++b) blk[b] = T(0.0);
  const T hh = x.mp(2), tsat = x.mp(3), ga = x.mp(1);
#pragma unroll 1
  for (int a = 0; a < N; ++a) {
    const double d0 = tab.d[a][0], d1 = tab.d[a][1], d2 = tab.d[a][2];
    const double n0 = tab.n[a][0], n1 = tab.n[a][1], n2 = tab.n[a][2];
    const T tau = d0 * (Sc[0] * n0 + Sc[1] * n1 + Sc[2] * n2) + d1 * (Sc[3] * n0 + Sc[4] * n1 + Sc[5] * n2) +
                  d2 * (Sc[6] * n0 + Sc[7] * n1 + Sc[8] * n2);
    const T ga_ = x.g(a);
    const T dg = c * dpow(dabs(tau / ga_), p) * sgnv(tau);
    o.dgam(a, dg);  // slip_new = slip_old + dgam
    Lpc[0] += dg * (d0 * n0); Lpc[1] += dg * (d0 * n1); Lpc[2] += dg * (d0 * n2);
    Lpc[3] += dg * (d1 * n0); Lpc[4] += dg * (d1 * n1); Lpc[5] += dg * (d1 * n2);
    Lpc[6] += dg * (d2 * n0); Lpc[7] += dg * (d2 * n1); Lpc[8] += dg * (d2 * n2);
    const T xh = 1.0 - ga_ / tsat;
    const T t = hh * dabs(dg) * dpow(dabs(xh), ga) * sgnv(xh);
    tot += t;
#pragma unroll
    for (int b = 0; b < N / 3; ++b)
      if (b == a / 3) blk[b] += t;
  }
  // g_new = g + r tot + (1 - r) blk  (q @ tmp with the latent matrix, models_copper.py:179)
  {
    const T rr = x.mp(4);
#pragma unroll 1
    for (int a = 0; a < N; ++a) {
      T bl = T(0.0);
#pragma unroll
      for (int b = 0; b < N / 3; ++b)
        if (b == a / 3) bl = blk[b];
      T inc = rr * tot;
      if (self_block) inc += (1.0 - rr) * bl;
      o.g(a, x.g(a) + inc);
    }
  }
  // Fp_inv_new = Fp_inv_old (I - R Lpc R^T); Fe = F Fp_inv_new
  T Fpn[9], Fe[9];
  {
    T X[9], Lp[9];
    tmm3(R, Lpc, X);
    tmm3t(X, R, Lp);
    T Fo[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Fo[i] = x.Fpi(i);
    T Z[9];
    tmm3(Fo, Lp, Z);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      Fpn[i] = Fo[i] - Z[i];
      o.Fpi(i, Fpn[i]);
    }
  }
  tmm3(F, Fpn, Fe);
  // residual r = s - C_rot : E with the closed-form rotated cubic stiffness (exact for any R)
  {
    T G[9];
    tmm3t(R, R, G);
    const T C11 = x.mp(5), C12 = x.mp(6), C44 = x.mp(7);
    const T Ac = C11 - C12 - 2.0 * C44;
    T e[6];
    e[0] = 0.5 * (Fe[0] * Fe[0] + Fe[3] * Fe[3] + Fe[6] * Fe[6] + (-1.0));
    e[1] = 0.5 * (Fe[1] * Fe[1] + Fe[4] * Fe[4] + Fe[7] * Fe[7] + (-1.0));
    e[2] = 0.5 * (Fe[2] * Fe[2] + Fe[5] * Fe[5] + Fe[8] * Fe[8] + (-1.0));
    e[3] = Fe[1] * Fe[2] + Fe[4] * Fe[5] + Fe[7] * Fe[8];
    e[4] = Fe[0] * Fe[2] + Fe[3] * Fe[5] + Fe[6] * Fe[8];
    e[5] = Fe[0] * Fe[1] + Fe[3] * Fe[4] + Fe[6] * Fe[7];
#pragma unroll
    for (int pI = 0; pI < 6; ++pI) {
      const int i = vi(pI), j = vj(pI);
      T acc = T(0.0);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int k = vi(q), l = vj(q);
        T cub = T(0.0);
#pragma unroll
        for (int aa = 0; aa < 3; ++aa) cub += R[3 * i + aa] * R[3 * j + aa] * R[3 * k + aa] * R[3 * l + aa];
        const T Cpq = C12 * G[3 * i + j] * G[3 * k + l] + C44 * (G[3 * i + k] * G[3 * j + l] + G[3 * i + l] * G[3 * j + k]) + Ac * cub;
        acc += Cpq * e[q];
      }
      o.r(pI, s[pI] - acc);
    }
  }
  // P = F Fpi S Fpi^T / det Fpi ; sigma = P F^T / det F
  T P[9];
  {
    T SFt[9], Y[9];
    tmm3t(S, Fpn, SFt);
    tmm3(Fpn, SFt, Y);
    const T idFp = T(1.0) / tdet3(Fpn);
#pragma unroll
    for (int i = 0; i < 9; ++i) Y[i] = Y[i] * idFp;
    tmm3(F, Y, P);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) o.P(i, P[i]);
  T PFt[9];
  tmm3t(P, F, PFt);
  const T idF = T(1.0) / tdet3(F);
#pragma unroll
  for (int i = 0; i < 9; ++i) o.sig(i, PFt[i] * idF);
}

// Accessor over plain-double inputs that seeds the derivative of one input direction.
template <int N>
struct SeededIn {
  const double* Hv;
  const double* Fv;
  const double* gv;  // g_old of this point, stride gs
  int64_t gs;
  const double* mv;  // 8 material values
  const double* Rv;
  int seed;          // input direction with unit derivative (-1: none)
  CPFEM_HD Dual H(int i) const { return Dual(Hv[i], seed == i ? 1.0 : 0.0); }
  CPFEM_HD Dual Fpi(int i) const { return Dual(Fv[i], seed == 9 + i ? 1.0 : 0.0); }
  CPFEM_HD Dual g(int a) const { return Dual(gv[a * gs], seed == 18 + a ? 1.0 : 0.0); }
  CPFEM_HD Dual mp(int m) const { return Dual(mv[m], seed == 18 + N + m ? 1.0 : 0.0); }
  CPFEM_HD Dual R(int i) const { return Dual(Rv[i], seed == 26 + N + i ? 1.0 : 0.0); }
};

// Sink that accumulates obar . d(outputs) and keeps d(residual).
template <int N>
struct VjpSink {
  const double* Pb;
  const double* sb;
  const double* Fb;
  const double* gb;  // may be null, stride gs
  const double* sl;  // cotangent of slip_new, may be null, stride gs
  int64_t gs;
  double acc;
  double dr[6];
  CPFEM_HD void dgam(int a, Dual v) {
    if (sl) acc += sl[a * gs] * v.d;
  }
  CPFEM_HD void r(int i, Dual v) { dr[i] = v.d; }
  CPFEM_HD void P(int i, Dual v) { acc += Pb[i] * v.d; }
  CPFEM_HD void sig(int i, Dual v) { acc += sb[i] * v.d; }
  CPFEM_HD void Fpi(int i, Dual v) { acc += Fb[i] * v.d; }
  CPFEM_HD void g(int a, Dual v) {
    if (gb) acc += gb[a * gs] * v.d;
  }
};

// Solve J^T mu = b with the partial-pivoting LU of lu6 (P J = L U).
CPFEM_HD void lu6_solve_T(const double* J, const int* piv, double* b) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {  // U^T y = b
#pragma unroll
    for (int j = 0; j < i; ++j) b[i] -= J[6 * j + i] * b[j];
    b[i] /= J[6 * i + i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i)  // L^T z = y
#pragma unroll
    for (int j = i + 1; j < 6; ++j) b[i] -= J[6 * j + i] * b[j];
#pragma unroll
  for (int k = 5; k >= 0; --k)  // undo the row swaps in reverse order
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const bool sw = (i == piv[k]);
      const double x = b[k], y = b[i];
      b[k] = sw ? y : x;
      b[i] = sw ? x : y;
    }
}

// Reverse-mode product of one point through its converged local solve s (implicit-function
// theorem on r(s, x) = 0):  xbar_k = obar . do/dx_k - mu . dr/dx_k,  J_s^T mu = (do/ds)^T obar.
// Jl is scratch for 42 doubles with stride js (LDS column on the device).  For every input
// direction k in the groups of `mask` (1 H, 2 Fp_inv_old, 4 g_old, 8 material, 16 R) calls
// emit(group, offset, value).
template <int N, class Emit>
CPFEM_HD void point_vjp(const SeededIn<N>& xin, const double* s, const SlipTable<N>& tab, double ao, double dt, int self_block,
                        const double* Pb, const double* sb, const double* Fb, const double* gbq, const double* slq, int64_t gs,
                        double* Jl, int js, int mask, Emit emit) {
  double* gl = Jl + 36 * js;
#pragma unroll 1
  for (int qd = 0; qd < 6; ++qd) {
    Dual sd[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) sd[k] = Dual(s[k], k == qd ? 1.0 : 0.0);
    VjpSink<N> o{Pb, sb, Fb, gbq, slq, gs, 0.0, {}};
    local_fn<Dual, N>(xin, sd, tab, ao, dt, self_block, o);
#pragma unroll
    for (int i = 0; i < 6; ++i) Jl[(6 * i + qd) * js] = o.dr[i];
    gl[qd * js] = o.acc;
  }
  double J[36], mu[6];
  int piv[6];
#pragma unroll
  for (int i = 0; i < 36; ++i) J[i] = Jl[i * js];
#pragma unroll
  for (int i = 0; i < 6; ++i) mu[i] = gl[i * js];
  lu6(J, piv);
  lu6_solve_T(J, piv, mu);
  Dual sc[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sc[k] = Dual(s[k]);
  const int nin = 35 + N;
#pragma unroll 1
  for (int k = 0; k < nin; ++k) {
    int grp, off;
    if (k < 9) { grp = 1; off = k; }
    else if (k < 18) { grp = 2; off = k - 9; }
    else if (k < 18 + N) { grp = 4; off = k - 18; }
    else if (k < 26 + N) { grp = 8; off = k - 18 - N; }
    else { grp = 16; off = k - 26 - N; }
    if (!(mask & grp)) continue;
    SeededIn<N> xk = xin;
    xk.seed = k;
    VjpSink<N> o{Pb, sb, Fb, gbq, slq, gs, 0.0, {}};
    local_fn<Dual, N>(xk, sc, tab, ao, dt, self_block, o);
    double xb = o.acc;
#pragma unroll
    for (int i = 0; i < 6; ++i) xb -= mu[i] * o.dr[i];
    emit(grp, off, xb);
  }
}

}  // namespace cpfem
