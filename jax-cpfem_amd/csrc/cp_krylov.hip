// Jacobi-preconditioned BiCGSTAB on the block-ELL matrix.  Vectors, dot products and
// scalars are FP64; the matrix values are FP64 or an FP32 copy (mixed form, where the
// values are ~90% of the SpMV's HBM traffic).
//
// Iteration semantics follow jax.scipy.sparse.linalg.bicgstab (JAX 0.4.x), which the
// reference calls at crystal_plasticity_OR_design/solver.py:34-40 with tol = atol = 1e-10,
// maxiter = 10000 and M = x / diag(A):
//   r0 = b - A x0; rhat = p = q = r0; rho = alpha = omega = 1; k = 0
//   while <r,r> > max(tol^2 <b,b>, atol^2) and 0 <= k < maxiter:
//     rho_ = <rhat,r>; beta = rho_/rho*alpha/omega; p = r + beta (p - omega q)
//     phat = M p; q = A phat; alpha = rho_/<rhat,q>; s = r - alpha q
//     shat = M s; t = A shat; omega = <t,s>/<t,t>
//     x += alpha phat + omega shat (alpha phat only if <s,s> < atol2); r = s - omega t (or s)
//     k = -11 on alpha/omega = 0, -10 on rho_ = 0, else k + 1
//
// Design: five kernels per iteration; every dot product is fused into the kernel that
// produces its operand and is finished by the last arriving block (fixed-order sum of
// per-block partials, so results are bitwise reproducible run to run).  All scalars stay on
// the device; each kernel returns at once when the device-side `done` flag is set, so the
// host can enqueue iterations in batches and poll rarely.
#include <algorithm>

#include "cp_internal.hpp"

namespace cpfem {

constexpr int KBS = 256;
constexpr int KMAXB = 512;  // blocks of the BiCGSTAB kernels (grid-stride beyond)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sum of NS values; result valid in thread 0.
template <int NS>
__device__ __forceinline__ void block_sums(double* v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const double s = wave_sum(v[j]);
    if (lane == 0) sh[j * (KBS / 64) + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < KBS / 64; ++i) s += sh[j * (KBS / 64) + i];
      v[j] = s;
    }
  }
  __syncthreads();
}

// Publishes this block's NS partial sums; returns true in every thread of the last block,
// which then holds the totals in tot[] (thread 0 only).
template <int NS>
__device__ __forceinline__ bool finish_sums(double* v, double* partial, int nb, uint32_t* counter, double* tot) {
  __shared__ double sh[NS * (KBS / 64)];
  __shared__ int last;
  block_sums<NS>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NS; ++j) partial[j * nb + blockIdx.x] = v[j];
    __threadfence();
    const uint32_t prev = atomicAdd(counter, 1u);
    last = (prev == (uint32_t)(gridDim.x - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NS];
  const volatile double* vp = partial;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += KBS) s += vp[j * nb + i];
    acc[j] = s;
  }
  block_sums<NS>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NS; ++j) tot[j] = acc[j];
    *counter = 0u;
  }
  return true;
}

// Row r of A x.  Hex meshes have exactly 27 node slots (padding slots hold zeros and point
// at the row's own node), so the common case runs a fixed, fully unrolled slot loop: all
// column indices are loaded first, then all values and x entries, giving 27-way memory-level
// parallelism per lane instead of a dependent index->x chain per slot.
template <typename TM>
__device__ __forceinline__ double spmv_row(const BellView& M, const TM* __restrict__ vals, const double* __restrict__ x,
                                           int64_t r) {
  const int64_t nrow = 3 * M.n_nodes;
  const int64_t i = r / 3;
  if (M.n_slots == 27) {
    int32_t j[27];
#pragma unroll
    for (int s = 0; s < 27; ++s) j[s] = M.col[(int64_t)s * M.n_nodes + i];
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      const TM* v = vals + (3 * (int64_t)s) * nrow + r;
      const int64_t jj = 3 * (int64_t)j[s];
      acc += (double)v[0] * x[jj] + (double)v[nrow] * x[jj + 1] + (double)v[2 * nrow] * x[jj + 2];
    }
    return acc;
  }
  const int ns = M.nslot[i];
  double acc = 0.0;
  for (int s = 0; s < ns; ++s) {
    const int64_t j = M.col[(int64_t)s * M.n_nodes + i];
    const TM* v = vals + (3 * (int64_t)s) * nrow + r;
    acc += (double)v[0] * x[3 * j] + (double)v[nrow] * x[3 * j + 1] + (double)v[2 * nrow] * x[3 * j + 2];
  }
  return acc;
}

__global__ void __launch_bounds__(KBS) k_spmv(BellView M, const double* __restrict__ vals, const double* __restrict__ x,
                                             double* __restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (r >= 3 * M.n_nodes) return;
  y[r] = spmv_row<double>(M, vals, x, r);
}

// r = b - A x (FP64)
__global__ void __launch_bounds__(KBS) k_residual64(BellView M, const double* __restrict__ vals, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i >= 3 * M.n_nodes) return;
  r[i] = b[i] - spmv_row<double>(M, vals, x, i);
}

__global__ void __launch_bounds__(KBS) k_inv_diag(BellView M, const double* __restrict__ vals, const int32_t* __restrict__ dslot,
                                                 double* __restrict__ invd, int precond) {
  const int64_t r = (int64_t)blockIdx.x * KBS + threadIdx.x;
  const int64_t nrow = 3 * M.n_nodes;
  if (r >= nrow) return;
  const int64_t i = r / 3;
  const int c = (int)(r - 3 * i);
  const double d = vals[(3 * (int64_t)dslot[i] + c) * nrow + r];
  invd[r] = precond ? 1.0 / d : 1.0;
}

// ---------------------------------------------------------------------------------
// BiCGSTAB kernels.  Grid: KrylovWork::nb blocks (<= KMAXB) with grid-stride loops.  Each
// reducing kernel writes one partial per block into its slot of w.partial; the kernel that
// consumes the sum adds the nb partials itself, in every block and in the same order, so no
// atomics are needed and every block sees the bitwise-identical scalar.  Scalars that carry
// over to the next iteration live in two parity slots: iteration j writes slot (j & 1) and
// reads slot ((j + 1) & 1), which the host passes as `par`.
enum { PS_RR = 0, PS_RHR, PS_RQ, PS_SS, PS_TS, PS_TT, PS_BB, PS_COUNT };

__device__ __forceinline__ void put_partial(double v, double* partial, int slot, int nb, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < KBS / 64; ++i) s += sh[i];
    partial[slot * nb + blockIdx.x] = s;
  }
}

// Sum of one partial slot, computed by the whole block; returned in every thread.
__device__ __forceinline__ double get_sum(const double* partial, int slot, int nb, double* sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += KBS) s += partial[slot * nb + i];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  s = wave_sum(s);
  if (lane == 0) sh[wid] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < KBS / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

template <typename TM>
__global__ void __launch_bounds__(KBS) k_kinit(BellView M, const TM* __restrict__ vals, const double* __restrict__ b, KVec v,
                                              KrylovWork w) {
  __shared__ double sh[KBS / 64];
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double r = b[i] - spmv_row<TM>(M, vals, v.x, i);
    v.r[i] = r;
    v.rhat[i] = r;
    v.p[i] = r;
    v.q[i] = r;
    rr += r * r;
    bb += b[i] * b[i];
  }
  put_partial(rr, w.partial, PS_RR, w.nb, sh);
  __syncthreads();
  put_partial(rr, w.partial, PS_RHR, w.nb, sh);
  __syncthreads();
  put_partial(bb, w.partial, PS_BB, w.nb, sh);
}

// Single block: atol2 and the "previous iteration" scalars seen by iteration 0.
__global__ void __launch_bounds__(KBS) k_kinit2(KrylovWork w) {
  __shared__ double sh[KBS / 64];
  const double bb = get_sum(w.partial, PS_BB, w.nb, sh);
  if (threadIdx.x == 0) {
    KrylovScalars* sc = w.sc;
    sc->bs = bb;
    sc->atol2 = fmax(sc->tol * sc->tol * bb, sc->atol * sc->atol);
    sc->rho2[1] = sc->alpha2[1] = sc->omega2[1] = 1.0;
    sc->k2[1] = -1;  // so that iteration 0 counts as k = 0
    sc->done = 0;
    sc->k = 0;
  }
}

// Start of iteration: convergence test of the previous one, then p and phat.
__global__ void __launch_bounds__(KBS) k_kp(KVec v, KrylovWork w, int par) {
  __shared__ double sh[KBS / 64];
  KrylovScalars* sc = w.sc;
  if (sc->done) return;
  const int pv = par ^ 1;
  const double rs = get_sum(w.partial, PS_RR, w.nb, sh);
  const double rho_ = get_sum(w.partial, PS_RHR, w.nb, sh);
  const double rho = sc->rho2[pv], alpha = sc->alpha2[pv], omega = sc->omega2[pv];
  int k = (omega == 0.0 || alpha == 0.0) ? -11 : sc->k2[pv] + 1;
  if (rho == 0.0) k = -10;
  const bool go = (rs > sc->atol2) && (k < sc->maxiter) && (k >= 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->k2[par] = k;
    sc->rho2[par] = rho_;
    sc->rs = rs;
    sc->k = k;
    if (!go) sc->done = 1;
  }
  if (!go) return;
  const double beta = rho_ / rho * alpha / omega;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double p = v.r[i] + beta * (v.p[i] - omega * v.q[i]);
    v.p[i] = p;
    v.phat[i] = p * v.invd[i];
  }
}

template <typename TM>
__global__ void __launch_bounds__(KBS) k_kq(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w) {
  __shared__ double sh[KBS / 64];
  if (w.sc->done) return;
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double q = spmv_row<TM>(M, vals, v.phat, i);
    v.q[i] = q;
    d += v.rhat[i] * q;
  }
  put_partial(d, w.partial, PS_RQ, w.nb, sh);
}

__global__ void __launch_bounds__(KBS) k_ks(KVec v, KrylovWork w, int par) {
  __shared__ double sh[KBS / 64];
  KrylovScalars* sc = w.sc;
  if (sc->done) return;
  const double alpha = sc->rho2[par] / get_sum(w.partial, PS_RQ, w.nb, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha2[par] = alpha;
  double d = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double s = v.r[i] - alpha * v.q[i];
    v.s[i] = s;
    v.shat[i] = s * v.invd[i];
    d += s * s;
  }
  put_partial(d, w.partial, PS_SS, w.nb, sh);
}

template <typename TM>
__global__ void __launch_bounds__(KBS) k_kt(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w) {
  __shared__ double sh[KBS / 64];
  if (w.sc->done) return;
  double ts = 0.0, tt = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double t = spmv_row<TM>(M, vals, v.shat, i);
    v.t[i] = t;
    ts += t * v.s[i];
    tt += t * t;
  }
  put_partial(ts, w.partial, PS_TS, w.nb, sh);
  __syncthreads();
  put_partial(tt, w.partial, PS_TT, w.nb, sh);
}

__global__ void __launch_bounds__(KBS) k_kxr(KVec v, KrylovWork w, int par) {
  __shared__ double sh[KBS / 64];
  KrylovScalars* sc = w.sc;
  if (sc->done) return;
  const double ss = get_sum(w.partial, PS_SS, w.nb, sh);
  const double omega = get_sum(w.partial, PS_TS, w.nb, sh) / get_sum(w.partial, PS_TT, w.nb, sh);
  const double alpha = sc->alpha2[par];
  const bool early = ss < sc->atol2;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omega2[par] = omega;
  double rr = 0.0, rhr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    double r;
    if (early) {
      v.x[i] = v.x[i] + alpha * v.phat[i];
      r = v.s[i];
    } else {
      v.x[i] = v.x[i] + (alpha * v.phat[i] + omega * v.shat[i]);
      r = v.s[i] - omega * v.t[i];
    }
    v.r[i] = r;
    rr += r * r;
    rhr += v.rhat[i] * r;
  }
  put_partial(rr, w.partial, PS_RR, w.nb, sh);
  __syncthreads();
  put_partial(rhr, w.partial, PS_RHR, w.nb, sh);
}

// Reliable update of the mixed form: replace the recurrence residual by the true FP64
// residual r = b - A64 x (and its partial sums), without restarting the Krylov space.
__global__ void __launch_bounds__(KBS) k_kreplace(BellView M, const double* __restrict__ vals, const double* __restrict__ b,
                                                 KVec v, KrylovWork w) {
  __shared__ double sh[KBS / 64];
  if (w.sc->done) return;
  double rr = 0.0, rhr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double r = b[i] - spmv_row<double>(M, vals, v.x, i);
    v.r[i] = r;
    rr += r * r;
    rhr += v.rhat[i] * r;
  }
  put_partial(rr, w.partial, PS_RR, w.nb, sh);
  __syncthreads();
  put_partial(rhr, w.partial, PS_RHR, w.nb, sh);
}

// Generic dot product a.b finished on device into *out
__global__ void __launch_bounds__(KBS) k_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* partial, int nb, double* out, uint32_t* counter) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  double d[1] = {i < n ? a[i] * b[i] : 0.0};
  double tot[1];
  if (finish_sums<1>(d, partial, nb, counter, tot) && threadIdx.x == 0) *out = tot[0];
}

__global__ void __launch_bounds__(KBS) k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i < n) y[i] = x[i];
}

__global__ void k_axpby(int64_t n, double alpha, const double* __restrict__ x, double beta, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i < n) y[i] = alpha * x[i] + beta * y[i];
}
__global__ void k_cast_f32(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + KBS - 1) / KBS); }

template <typename TM>
void krylov_init(const BellView& M, const TM* vals, const double* b, KVec v, KrylovWork w, hipStream_t st) {
  hipLaunchKernelGGL(k_kinit<TM>, dim3((unsigned)w.nb), dim3(KBS), 0, st, M, vals, b, v, w);
  hipLaunchKernelGGL(k_kinit2, dim3(1), dim3(KBS), 0, st, w);
}

int krylov_blocks(int64_t n) { return (int)std::min<int64_t>(nblk(n), KMAXB); }

void krylov_replace(const BellView& M, const double* vals64, const double* b, KVec v, KrylovWork w, hipStream_t st) {
  hipLaunchKernelGGL(k_kreplace, dim3((unsigned)w.nb), dim3(KBS), 0, st, M, vals64, b, v, w);
}

template <typename TM>
void krylov_iteration(const BellView& M, const TM* vals, KVec v, KrylovWork w, int par, hipStream_t st, Prof* prof) {
  const unsigned nb = (unsigned)w.nb;
  int sl;
  sl = prof_begin(prof, PC_KVEC, st);
  hipLaunchKernelGGL(k_kp, dim3(nb), dim3(KBS), 0, st, v, w, par);
  prof_end(prof, sl, st);
  sl = prof_begin(prof, PC_SPMV, st);
  hipLaunchKernelGGL(k_kq<TM>, dim3(nb), dim3(KBS), 0, st, M, vals, v, w);
  prof_end(prof, sl, st);
  sl = prof_begin(prof, PC_KVEC, st);
  hipLaunchKernelGGL(k_ks, dim3(nb), dim3(KBS), 0, st, v, w, par);
  prof_end(prof, sl, st);
  sl = prof_begin(prof, PC_SPMV, st);
  hipLaunchKernelGGL(k_kt<TM>, dim3(nb), dim3(KBS), 0, st, M, vals, v, w);
  prof_end(prof, sl, st);
  sl = prof_begin(prof, PC_KVEC, st);
  hipLaunchKernelGGL(k_kxr, dim3(nb), dim3(KBS), 0, st, v, w, par);
  prof_end(prof, sl, st);
}

void launch_spmv(const BellView& M, const double* vals, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_spmv, dim3(nblk(3 * M.n_nodes)), dim3(KBS), 0, st, M, vals, x, y);
}

void launch_inv_diag(const BellView& M, const double* vals, const int32_t* dslot, double* invd, int precond, hipStream_t st) {
  hipLaunchKernelGGL(k_inv_diag, dim3(nblk(3 * M.n_nodes)), dim3(KBS), 0, st, M, vals, dslot, invd, precond);
}

void launch_residual_norm(int64_t n, const double* a, const double* b, double* partial, int nb, double* out, uint32_t* counter,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3((unsigned)nb), dim3(KBS), 0, st, n, a, b, partial, nb, out, counter);
}
void launch_copy(int64_t n, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(KBS), 0, st, n, x, y);
}
void launch_axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(KBS), 0, st, n, alpha, x, beta, y);
}
void launch_cast_f32(int64_t n, const double* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_cast_f32, dim3(nblk(n)), dim3(KBS), 0, st, n, x, y);
}
void launch_residual64(const BellView& M, const double* vals, const double* x, const double* b, double* r, hipStream_t st) {
  hipLaunchKernelGGL(k_residual64, dim3(nblk(3 * M.n_nodes)), dim3(KBS), 0, st, M, vals, x, b, r);
}

template void krylov_init<double>(const BellView&, const double*, const double*, KVec, KrylovWork, hipStream_t);
template void krylov_init<float>(const BellView&, const float*, const double*, KVec, KrylovWork, hipStream_t);
template void krylov_iteration<double>(const BellView&, const double*, KVec, KrylovWork, int, hipStream_t, Prof*);
template void krylov_iteration<float>(const BellView&, const float*, KVec, KrylovWork, int, hipStream_t, Prof*);

}  // namespace cpfem
