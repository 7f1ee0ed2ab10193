// Per-quadrature-point Kalidindi crystal-plasticity update, FP64, one lane per point.
//
// Follows singlecrystal_copper/models_copper.py:151-261 (and the per-point-parameter
// variants calibration/case1.py:148-260, polycrystal_DPsteel/models_DPsteel_inhomo.py:232-330):
//   tau_a   = S : (R P0_a R^T)                                    (:173)
//   dgam_a  = gamma0 dt |tau_a / g_a|^(1/m) sign(tau_a)           (:174, g = old resistance)
//   Lp      = sum_a dgam_a R P0_a R^T ; Fp_inv = Fp_inv_old (I - Lp) ; Fe = F Fp_inv   (:189-191)
//   r(S)    = S - C_rot : 1/2 (Fe^T Fe - I)                       (:199-200)
//   Newton from S = 0, |r| > tol, cut-half backtracking accepting the last trial   (:204-249)
//   P       = det F * (Fe S Fe^T / det Fe) F^-T                   (:160-161)
//   tangent = IFT: dS = -J_S^-1 J_H dH                            (:251-259)
//
// Design (not a translation): the reference solves a 9-unknown system with a 9x9
// forward-mode Jacobian and recomputes the rank-4 rotation every residual.  Here
//  * S stays symmetric (S0 = 0, symmetric residual range, skew part of every step is 0),
//    so the unknown is the 6-vector (S00,S11,S22,S12,S02,S01); the convergence and
//    line-search decisions use the 9-component norm (off-diagonals counted twice);
//  * the rotated stiffness is formed once per point in closed form from the cubic
//    constants: C_ijkl = C12 G_ij G_kl + C44 (G_ik G_jl + G_il G_jk)
//    + (C11 - C12 - 2 C44) sum_a R_ia R_ja R_ka R_la, G = R R^T (exact for any R);
//  * tau uses the crystal-frame stress R^T S R against the unrotated slip vectors, which
//    live in the kernel argument block (uniform across lanes -> scalar loads);
//  * the Jacobian is I + C6 * sum_a g'_a W_a b_a^T (one 6x6x6 product per Jacobian);
//  * P = F (Fp_inv S Fp_inv^T) / det Fp_inv, identical to the reference expression.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CPFEM_HD __host__ __device__ __forceinline__
#else
// Plain C++ build of the same math (used by the CPU check of the kernel arithmetic).
#define CPFEM_HD inline
#endif

namespace cpfem {

constexpr int MAX_SLIP = 48;

template <int N>
struct SlipTable {
  double d[N][3];  // normalised slip directions (crystal frame)
  double n[N][3];  // normalised slip-plane normals
};

// Natural logarithm and exponential for the slip-rate power law, written out so that a power
// costs ~55 VALU instructions on gfx950 instead of the library pow's 226 (exp(p log x) with
// the library log and exp: 140).  Methods: the logarithm reduces x = m 2^e with m in
// [sqrt(1/2), sqrt(2)) and sums the series ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^23/23),
// s = (m - 1) / (m + 1), |s| <= 0.1716, so the first omitted term is below 2^-57 relative;
// the exponential reduces x = k ln 2 + r with ln 2 split into a head and a tail (Cody and
// Waite's additive reduction), |r| <= ln 2 / 2, and sums the Taylor series of e^r to r^13
// (first omitted term ~4e-18).  Both stay within a few units in the last place of the
// correctly rounded values (checked against the C library over wide ranges by the CPU tests).
constexpr double LN2_HEAD = 6.93147180369123816490e-01;  // ln 2 to 32 bits: k * LN2_HEAD is exact
constexpr double LN2_TAIL = 1.90821492927058770002e-10;  // ln 2 - LN2_HEAD
constexpr double LN2_D = 0.6931471805599453094;         // ln 2 rounded to double

// ln x for x > 0; ln 0 = -inf, ln inf = inf, NaN and negative arguments give NaN.
CPFEM_HD double ln_pos(double x) {
  if (!(x > 0.0) || x > 1.7976931348623157e308) return x == 0.0 ? -INFINITY : x + (x < 0.0 ? NAN : 0.0);
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    --e;
  }
  const double s = (m - 1.0) / (m + 1.0), z = s * s;
  double t = 1.0 / 23.0;
  t = fma(t, z, 1.0 / 21.0);
  t = fma(t, z, 1.0 / 19.0);
  t = fma(t, z, 1.0 / 17.0);
  t = fma(t, z, 1.0 / 15.0);
  t = fma(t, z, 1.0 / 13.0);
  t = fma(t, z, 1.0 / 11.0);
  t = fma(t, z, 1.0 / 9.0);
  t = fma(t, z, 1.0 / 7.0);
  t = fma(t, z, 1.0 / 5.0);
  t = fma(t, z, 1.0 / 3.0);
  const double lm = 2.0 * fma(s * z, t, s);  // 2 atanh(s)
  return fma((double)e, LN2_D, lm);
}

// e^x; underflows to 0 below -745.2 and overflows to inf above 709.8; NaN propagates.
CPFEM_HD double exp_arg(double x) {
  if (!(x > -745.2)) return x != x ? x : 0.0;
  if (x > 709.8) return INFINITY;
  const double k = rint(x * (1.0 / LN2_D));
  const double r = fma(-k, LN2_TAIL, fma(-k, LN2_HEAD, x));
  double q = 1.0 / 6227020800.0;     // 1/13!
  q = fma(q, r, 1.0 / 479001600.0);  // 1/12!
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, (int)k);
}

// x^p for x >= 0 and p > 0 as e^(p ln x).  Relative error within (4 |p ln x| + 8) 2^-52 (the
// logarithm's few units in the last place, amplified by the exponent: < 1e-13 for the slip
// rates here); x = 0 gives e^-inf = 0, and inf and NaN propagate, as pow does for p > 0.
CPFEM_HD double pow_pos(double x, double p) { return exp_arg(p * ln_pos(x)); }

// Per-point material constants
struct PointMat {
  double p;      // 1/m
  double c;      // gamma0 * dt
  double gss_a;  // hardening exponent a
  double h;      // h0
  double tsat;   // saturation resistance
  double r;      // latent ratio
  double C11, C12, C44;
};

// Status bits per point
enum : int32_t {
  PT_CAPPED = 1,     // hit the iteration cap (no cap in the reference)
  PT_NONFINITE = 2,  // residual norm became NaN/Inf
};

CPFEM_HD double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

// 3x3 helpers, row-major
CPFEM_HD void mm3(const double* a, const double* b, double* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
// c = a * b^T
CPFEM_HD void mm3t(const double* a, const double* b, double* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[3 * i] * b[3 * j] + a[3 * i + 1] * b[3 * j + 1] + a[3 * i + 2] * b[3 * j + 2];
}
// c = a^T * b
CPFEM_HD void mtm3(const double* a, const double* b, double* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) c[3 * i + j] = a[i] * b[j] + a[3 + i] * b[3 + j] + a[6 + i] * b[6 + j];
}
CPFEM_HD double det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
CPFEM_HD void inv3(const double* a, double det, double* b) {
  const double id = 1.0 / det;
  b[0] = (a[4] * a[8] - a[5] * a[7]) * id;
  b[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  b[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  b[3] = (a[5] * a[6] - a[3] * a[8]) * id;
  b[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  b[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  b[6] = (a[3] * a[7] - a[4] * a[6]) * id;
  b[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  b[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// Voigt index pairs: 0:(0,0) 1:(1,1) 2:(2,2) 3:(1,2) 4:(0,2) 5:(0,1)
CPFEM_HD int vi(int p) { return p < 3 ? p : (p == 3 ? 1 : 0); }
CPFEM_HD int vj(int p) { return p < 3 ? p : (p == 5 ? 1 : 2); }
// packed symmetric 6x6 (upper triangle, row-major): index of (p,q), p<=q
CPFEM_HD int sym6(int p, int q) {
  if (p > q) { int t = p; p = q; q = t; }
  return p * 6 - (p * (p - 1)) / 2 + (q - p);
}

// Rotated cubic stiffness in Voigt form (21 packed entries), C6[p][q] = C_{ij kl}.
CPFEM_HD void rotated_stiffness(const double* R, double C11, double C12, double C44, double* C6) {
  double G[9];
  mm3t(R, R, G);  // G = R R^T
  const double Ac = C11 - C12 - 2.0 * C44;
#pragma unroll
  for (int p = 0; p < 6; ++p)
#pragma unroll
    for (int q = p; q < 6; ++q) {
      const int i = vi(p), j = vj(p), k = vi(q), l = vj(q);
      double cub = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) cub += R[3 * i + a] * R[3 * j + a] * R[3 * k + a] * R[3 * l + a];
      C6[sym6(p, q)] = C12 * G[3 * i + j] * G[3 * k + l] + C44 * (G[3 * i + k] * G[3 * j + l] + G[3 * i + l] * G[3 * j + k]) + Ac * cub;
    }
}

// y = C6 * x (x in engineering/Voigt form with doubled shears)
CPFEM_HD void c6mul(const double* C6, const double* x, double* y) {
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) s += C6[sym6(p, q)] * x[q];
    y[p] = s;
  }
}

CPFEM_HD void sym_full(const double* s, double* S) {
  S[0] = s[0]; S[1] = s[5]; S[2] = s[4];
  S[3] = s[5]; S[4] = s[1]; S[5] = s[3];
  S[6] = s[4]; S[7] = s[3]; S[8] = s[2];
}

// In-place LU with partial pivoting of a 6x6 matrix; piv[k] is the row swapped with k.
// Row swaps use compare-and-select over constant indices so that the matrix stays in
// registers (a runtime row index would force it into scratch memory).
CPFEM_HD void lu6(double* J, int* piv) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int p = k;
    double best = fabs(J[6 * k + k]);
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double v = fabs(J[6 * i + k]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const bool sw = (i == p);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double a = J[6 * k + j], b = J[6 * i + j];
        J[6 * k + j] = sw ? b : a;
        J[6 * i + j] = sw ? a : b;
      }
    }
    const double inv = 1.0 / J[6 * k + k];
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const double f = J[6 * i + k] * inv;
      J[6 * i + k] = f;
#pragma unroll
      for (int j = k + 1; j < 6; ++j) J[6 * i + j] -= f * J[6 * k + j];
    }
  }
}
CPFEM_HD void lu6_solve(const double* J, const int* piv, double* b) {
#pragma unroll
  for (int k = 0; k < 6; ++k) {
#pragma unroll
    for (int i = k + 1; i < 6; ++i) {
      const bool sw = (i == piv[k]);
      const double x = b[k], y = b[i];
      b[k] = sw ? y : x;
      b[i] = sw ? x : y;
    }
  }
#pragma unroll
  for (int i = 1; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) b[i] -= J[6 * i + j] * b[j];
#pragma unroll
  for (int i = 5; i >= 0; --i) {
#pragma unroll
    for (int j = i + 1; j < 6; ++j) b[i] -= J[6 * i + j] * b[j];
    b[i] /= J[6 * i + i];
  }
}

// Per-point solver state.  The per-slip-system arrays (1/g_old and d dgamma/d tau) live
// outside the register file: in LDS on the device (column [a * stride + lane]), in a plain
// array in the host build, so the slip loops need not be unrolled.
template <int N>
struct Point {
  double R[9];    // rotation (as given, not re-orthonormalised)
  double A[9];    // F Fp_inv_old
  double C6[21];  // rotated stiffness, packed symmetric Voigt
  double p, c;    // 1/m, gamma0 dt
  double* gpv;    // d dgamma_a / d tau_a   [a * st]
  double* giv;    // 1 / g_old_a            [a * sti]
  int st;
  int64_t sti;
  // results of the last residual evaluation
  double Lpc[9];  // sum_a dgamma_a d_a (x) n_a   (crystal frame)
  double Fe[9];
  double r[6];
  double nr;
  int nev;        // residual evaluations since setup (work counter)

  CPFEM_HD double& gp(int a) { return gpv[a * st]; }
  CPFEM_HD double& gi(int a) { return giv[a * sti]; }

  // gp_store / gi_store: the per-slip arrays, at stride `stride` and `stride_gi` (-1: the same)
  CPFEM_HD void setup(const double* F, const double* Fpi_old, const double* Rin, const double* g_old, int64_t gstride,
                      const PointMat& m, double* gp_store, double* gi_store, int stride, int64_t stride_gi = -1) {
    gpv = gp_store;
    giv = gi_store;
    st = stride;
    sti = stride_gi < 0 ? stride : stride_gi;
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = Rin[i];
    mm3(F, Fpi_old, A);
    rotated_stiffness(R, m.C11, m.C12, m.C44, C6);
#pragma unroll 2
    for (int a = 0; a < N; ++a) gi(a) = 1.0 / g_old[a * gstride];
    p = m.p;
    c = m.c;
    nev = 0;
  }

  // crystal-frame stress R^T S R of the symmetric Voigt stress s
  CPFEM_HD void crystal_stress(const double* s, double* Sc) const {
    double S[9], T[9];
    sym_full(s, S);
    mm3(S, R, T);
    mtm3(R, T, Sc);
  }

  // Residual at symmetric S (Voigt 6); fills gp, Lpc, Fe, r, nr.
  CPFEM_HD void eval(const SlipTable<N>& tab, const double* s) {
    ++nev;
    double Sc[9];
    crystal_stress(s, Sc);
#pragma unroll
    for (int i = 0; i < 9; ++i) Lpc[i] = 0.0;
#pragma unroll 2
    for (int a = 0; a < N; ++a) {
      const double d0 = tab.d[a][0], d1 = tab.d[a][1], d2 = tab.d[a][2];
      const double n0 = tab.n[a][0], n1 = tab.n[a][1], n2 = tab.n[a][2];
      const double v0 = Sc[0] * n0 + Sc[1] * n1 + Sc[2] * n2;
      const double v1 = Sc[3] * n0 + Sc[4] * n1 + Sc[5] * n2;
      const double v2 = Sc[6] * n0 + Sc[7] * n1 + Sc[8] * n2;
      const double tau = d0 * v0 + d1 * v1 + d2 * v2;
      const double ax = fabs(tau * gi(a));
      const double pw = pow_pos(ax, p);
      const double dg = c * pw * sgn(tau);
      gp(a) = (tau != 0.0) ? c * p * pw / fabs(tau) : (p == 1.0 ? c * gi(a) : 0.0);
      Lpc[0] += dg * d0 * n0; Lpc[1] += dg * d0 * n1; Lpc[2] += dg * d0 * n2;
      Lpc[3] += dg * d1 * n0; Lpc[4] += dg * d1 * n1; Lpc[5] += dg * d1 * n2;
      Lpc[6] += dg * d2 * n0; Lpc[7] += dg * d2 * n1; Lpc[8] += dg * d2 * n2;
    }
    // Fe = A (I - R Lpc R^T)
    double X[9], Y[9], Z[9];
    mm3(R, Lpc, X);
    mm3t(X, R, Y);
    mm3(A, Y, Z);
#pragma unroll
    for (int i = 0; i < 9; ++i) Fe[i] = A[i] - Z[i];
    // E = 1/2 (Fe^T Fe - I), engineering form
    double e[6];
    e[0] = 0.5 * (Fe[0] * Fe[0] + Fe[3] * Fe[3] + Fe[6] * Fe[6] - 1.0);
    e[1] = 0.5 * (Fe[1] * Fe[1] + Fe[4] * Fe[4] + Fe[7] * Fe[7] - 1.0);
    e[2] = 0.5 * (Fe[2] * Fe[2] + Fe[5] * Fe[5] + Fe[8] * Fe[8] - 1.0);
    e[3] = Fe[1] * Fe[2] + Fe[4] * Fe[5] + Fe[7] * Fe[8];
    e[4] = Fe[0] * Fe[2] + Fe[3] * Fe[5] + Fe[6] * Fe[8];
    e[5] = Fe[0] * Fe[1] + Fe[3] * Fe[4] + Fe[6] * Fe[7];
    double stv[6];
    c6mul(C6, e, stv);
#pragma unroll
    for (int q = 0; q < 6; ++q) r[q] = s[q] - stv[q];
    nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + 2.0 * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]));
  }

  // Slip increments at symmetric S (same arithmetic as eval).
  CPFEM_HD void slip_increment(const SlipTable<N>& tab, const double* Sc, int a, double* dg) {
    const double d0 = tab.d[a][0], d1 = tab.d[a][1], d2 = tab.d[a][2];
    const double n0 = tab.n[a][0], n1 = tab.n[a][1], n2 = tab.n[a][2];
    const double v0 = Sc[0] * n0 + Sc[1] * n1 + Sc[2] * n2;
    const double v1 = Sc[3] * n0 + Sc[4] * n1 + Sc[5] * n2;
    const double v2 = Sc[6] * n0 + Sc[7] * n1 + Sc[8] * n2;
    const double tau = d0 * v0 + d1 * v1 + d2 * v2;
    *dg = c * pow_pos(fabs(tau * gi(a)), p) * sgn(tau);
  }

  // Jacobian dr/ds (6x6, row-major) at the last evaluated point:
  // J = I + C6 * sum_a g'_a W_a b_a^T with W_a = 2 sym(M R d_a (x) R n_a) (engineering form),
  // b_a = d tau_a / d s, M = Fe^T A.
  CPFEM_HD void jac(const SlipTable<N>& tab, double* J) {
    double M[9], MR[9];
    mtm3(Fe, A, M);
    mm3(M, R, MR);
#pragma unroll
    for (int i = 0; i < 36; ++i) J[i] = 0.0;
#pragma unroll 2
    for (int a = 0; a < N; ++a) {
      const double d0 = tab.d[a][0], d1 = tab.d[a][1], d2 = tab.d[a][2];
      const double n0 = tab.n[a][0], n1 = tab.n[a][1], n2 = tab.n[a][2];
      const double m0 = MR[0] * d0 + MR[1] * d1 + MR[2] * d2;
      const double m1 = MR[3] * d0 + MR[4] * d1 + MR[5] * d2;
      const double m2 = MR[6] * d0 + MR[7] * d1 + MR[8] * d2;
      const double t0 = R[0] * d0 + R[1] * d1 + R[2] * d2;
      const double t1 = R[3] * d0 + R[4] * d1 + R[5] * d2;
      const double t2 = R[6] * d0 + R[7] * d1 + R[8] * d2;
      const double u0 = R[0] * n0 + R[1] * n1 + R[2] * n2;
      const double u1 = R[3] * n0 + R[4] * n1 + R[5] * n2;
      const double u2 = R[6] * n0 + R[7] * n1 + R[8] * n2;
      const double g = gp(a);
      double w[6], b[6];
      w[0] = g * m0 * u0; w[1] = g * m1 * u1; w[2] = g * m2 * u2;
      w[3] = g * (m1 * u2 + m2 * u1); w[4] = g * (m0 * u2 + m2 * u0); w[5] = g * (m0 * u1 + m1 * u0);
      b[0] = t0 * u0; b[1] = t1 * u1; b[2] = t2 * u2;
      b[3] = t1 * u2 + t2 * u1; b[4] = t0 * u2 + t2 * u0; b[5] = t0 * u1 + t1 * u0;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) J[6 * i + j] += w[i] * b[j];
    }
    // in place, column by column: J[:, j] <- e_j + C6 * G[:, j]
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double col[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double s = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) s += C6[sym6(i, k)] * J[6 * k + j];
        col[i] = s;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) J[6 * i + j] = col[i];
    }
  }

  // One local Newton iteration from the last evaluated point s (models_copper.py:218-247):
  // d = -J^-1 r, then trials s + lam d for lam = 1, 1/2, ... until |r_trial| < |r| or
  // max_sub_step trials; the last trial is accepted (s <- trial, r and nr are its residual).
  CPFEM_HD void newton_step(const SlipTable<N>& tab, int max_sub_step, double* s) {
    double dlt[6];
    {
      double J[36];
      int piv[6];
      jac(tab, J);
      lu6(J, piv);
#pragma unroll
      for (int q = 0; q < 6; ++q) dlt[q] = -r[q];
      lu6_solve(J, piv, dlt);
    }
    const double nr0 = nr;
    double lam = 1.0;
    int k = 0;
    double trial[6];
    do {
#pragma unroll
      for (int q = 0; q < 6; ++q) trial[q] = s[q] + lam * dlt[q];
      eval(tab, trial);
      lam *= 0.5;
      ++k;
    } while (nr >= nr0 && k < max_sub_step);
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = trial[q];
  }

  // Iterations from the current evaluated point while |r| > tol, counted in *it, up to it_end;
  // sets PT_CAPPED if it_end is reached unconverged and PT_NONFINITE on a non-finite |r|.
  CPFEM_HD void iterate(const SlipTable<N>& tab, double tol, int max_sub_step, int it_end, double* s, int* it,
                        int32_t* stat) {
    while (nr > tol) {
      if (*it >= it_end) { *stat |= PT_CAPPED; break; }
      newton_step(tab, max_sub_step, s);
      ++*it;
    }
    if (!(fabs(nr) <= 1.7976931348623157e308)) *stat |= PT_NONFINITE;
  }

  // Local Newton from S = 0 with cut-half backtracking (models_copper.py:204-249).
  // Returns the iteration count; status bits in *stat; s holds the converged stress.
  CPFEM_HD int solve(const SlipTable<N>& tab, double tol, int max_sub_step, int max_iter, double* s, int32_t* stat) {
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = 0.0;
    eval(tab, s);
    int it = 0;
    iterate(tab, tol, max_sub_step, max_iter, s, &it, stat);
    return it;
  }
};

// Plastic deformation after the converged solve: Fp_inv_new = Fp_inv_old (I - R Lpc R^T)
CPFEM_HD void fp_inv_new(const double* Fpi_old, const double* R, const double* Lpc, double* Fpi) {
  double X[9], Y[9], Z[9];
  mm3(R, Lpc, X);
  mm3t(X, R, Y);
  mm3(Fpi_old, Y, Z);
#pragma unroll
  for (int i = 0; i < 9; ++i) Fpi[i] = Fpi_old[i] - Z[i];
}

// First Piola-Kirchhoff stress P = F Fp_inv S Fp_inv^T / det Fp_inv at the converged point
// (equal to det F (Fe S Fe^T / det Fe) F^-T of models_copper.py:160-161 since Fe = F Fp_inv).
template <int N>
CPFEM_HD void pk1(const Point<N>& pt, const double* F, const double* Fpi_old, const double* s, double* P) {
  double Fpi[9], S[9], SFt[9], Y[9];
  fp_inv_new(Fpi_old, pt.R, pt.Lpc, Fpi);
  const double idFp = 1.0 / det3(Fpi);
  sym_full(s, S);
  mm3t(S, Fpi, SFt);
  mm3(Fpi, SFt, Y);
#pragma unroll
  for (int i = 0; i < 9; ++i) Y[i] *= idFp;
  mm3(F, Y, P);
}

// Consistent tangent dP/dH at the converged point (IFT of f_jvp, models_copper.py:251-259).
// With P = F Y, Y = Fpi S Fpi^T / det Fpi and ds = J^-1 C6 : 2 sym(Fe^T dH Fpi):
//   dP = dH Y + F Ycal(ds),  Ycal = dY/ds (6x6, through S and through Fpi(S)).
// Ycal is formed once from six unit stress directions; the nine dH = e_k e_l^T columns then
// cost one 6x6 solve each.  All loops over directions are unrolled so every array index is a
// compile-time constant (register resident).  Column m = 3k+l goes to sink(m, dP[9]).
template <int N, class Sink>
CPFEM_HD void tangent(Point<N>& pt, const SlipTable<N>& tab, const double* F, const double* Fpi_old, const double* s,
                      Sink sink) {
  double Fpi[9];
  fp_inv_new(Fpi_old, pt.R, pt.Lpc, Fpi);
  const double dFp = det3(Fpi);
  const double idFp = 1.0 / dFp;
  double U[9], Yv[9];
  {
    double S[9];
    sym_full(s, S);
    mm3t(S, Fpi, U);  // S Fpi^T
    mm3(Fpi, U, Yv);  // Fpi S Fpi^T (= det * Y)
  }
  double Yc[36];  // Ycal[p][q] at 6 * p + q
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int i = vi(q), j = vj(q);
    // dSc = R^T dS R for the unit symmetric direction q
    double dSc[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        dSc[3 * a + b] = (i == j) ? pt.R[3 * i + a] * pt.R[3 * i + b]
                                  : pt.R[3 * i + a] * pt.R[3 * j + b] + pt.R[3 * j + a] * pt.R[3 * i + b];
    double dLpc[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) dLpc[m] = 0.0;
#pragma unroll 2
    for (int a = 0; a < N; ++a) {
      const double d0 = tab.d[a][0], d1 = tab.d[a][1], d2 = tab.d[a][2];
      const double n0 = tab.n[a][0], n1 = tab.n[a][1], n2 = tab.n[a][2];
      const double v0 = dSc[0] * n0 + dSc[1] * n1 + dSc[2] * n2;
      const double v1 = dSc[3] * n0 + dSc[4] * n1 + dSc[5] * n2;
      const double v2 = dSc[6] * n0 + dSc[7] * n1 + dSc[8] * n2;
      const double gd = pt.gp(a) * (d0 * v0 + d1 * v1 + d2 * v2);
      dLpc[0] += gd * d0 * n0; dLpc[1] += gd * d0 * n1; dLpc[2] += gd * d0 * n2;
      dLpc[3] += gd * d1 * n0; dLpc[4] += gd * d1 * n1; dLpc[5] += gd * d1 * n2;
      dLpc[6] += gd * d2 * n0; dLpc[7] += gd * d2 * n1; dLpc[8] += gd * d2 * n2;
    }
    // dFpi = -Fpi_old R dLpc R^T
    double dFpi[9];
    {
      double X[9], Y[9];
      mm3(pt.R, dLpc, X);
      mm3t(X, pt.R, Y);
      mm3(Fpi_old, Y, dFpi);
    }
    // tr(Fpi^-1 dFpi) with the sign of dFpi folded in below
    double iFp[9];
    inv3(Fpi, dFp, iFp);
    double tr = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) tr -= iFp[3 * a + b] * dFpi[3 * b + a];
    double T1[9];
    mm3(dFpi, U, T1);  // -(dFpi S Fpi^T)
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const int a = vi(p), b = vj(p);
      const double mid = (i == j) ? Fpi[3 * a + i] * Fpi[3 * b + i]
                                  : Fpi[3 * a + i] * Fpi[3 * b + j] + Fpi[3 * a + j] * Fpi[3 * b + i];
      Yc[6 * p + q] = (-T1[3 * a + b] - T1[3 * b + a] + mid - Yv[3 * a + b] * tr) * idFp;
    }
  }
  double J[36];
  int piv[6];
  pt.jac(tab, J);
  lu6(J, piv);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double f0 = pt.Fe[3 * k], f1 = pt.Fe[3 * k + 1], f2 = pt.Fe[3 * k + 2];
      const double q0 = Fpi[3 * l], q1 = Fpi[3 * l + 1], q2 = Fpi[3 * l + 2];
      double w[6], ds[6];
      w[0] = f0 * q0; w[1] = f1 * q1; w[2] = f2 * q2;
      w[3] = f1 * q2 + f2 * q1; w[4] = f0 * q2 + f2 * q0; w[5] = f0 * q1 + f1 * q0;
      c6mul(pt.C6, w, ds);
      lu6_solve(J, piv, ds);
      double dy[6], dY[9], dP[9];
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) acc += Yc[6 * p + q] * ds[q];
        dy[p] = acc;
      }
      sym_full(dy, dY);
      mm3(F, dY, dP);
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) dP[3 * k + jj] += Yv[3 * l + jj] * idFp;
      sink(3 * k + l, dP);
    }
  }
}

// Kalidindi hardening after convergence (models_copper.py:178-187):
// g_new_a = g_a + sum_b q_ab h |dgam_b| |1 - g_b/gsat|^a sign(1 - g_b/gsat)
// q_ab = r, except q_ab = 1 inside blocks of 3 when self_block.  tmp[] is overwritten.
template <int N>
CPFEM_HD void hardening(const double* g_old, int64_t gstride, double* tmp, int tstride, const PointMat& m, int self_block,
                        double* g_new, int64_t ostride) {
  double tot = 0.0;
  for (int b = 0; b < N; ++b) {
    const double x = 1.0 - g_old[b * gstride] / m.tsat;
    const double t = m.h * fabs(tmp[b * tstride]) * pow_pos(fabs(x), m.gss_a) * sgn(x);
    tmp[b * tstride] = t;
    tot += t;
  }
  for (int a = 0; a < N; ++a) {
    double inc = m.r * tot;
    if (self_block) {
      const int b0 = (a / 3) * 3;
      const double blk = tmp[b0 * tstride] + tmp[(b0 + 1) * tstride] + tmp[(b0 + 2) * tstride];
      inc += (1.0 - m.r) * blk;
    }
    g_new[a * ostride] = g_old[a * gstride] + inc;
  }
}

}  // namespace cpfem
