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
// Design: three kernels per iteration.  p is never stored: with rh = M r, qh = M q and
// pp = M (p - omega q) kept per row (M is diagonal), q = A phat = A rh + beta A pp and
// t = A shat = A rh - alpha A qh, so the p and s updates ride inside the two SpMV kernels and
// only the x/r update is a kernel of its own.  (Same recurrence as JAX's in exact
// arithmetic; the Jacobi scaling and the SpMV are applied before instead of after the vector
// sums, which moves results by rounding only.)  Every dot product is
// fused into the kernel producing its operand as per-block partial sums, which the consuming
// kernel adds in a fixed order (bitwise reproducible run to run, no atomics).  All scalars
// stay on the device; each kernel returns at once when the device-side `done` flag is set,
// so the host enqueues iterations in batches and polls rarely.  (On the largest systems in
// FP64 a one-operand variant stores phat instead: k_kpq_flat1 below.)
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "cp_internal.hpp"

namespace cpfem {

constexpr int KBS = 256;
constexpr int KMAXB = 512;  // blocks of the BiCGSTAB kernels (grid-stride beyond)

// One DPP step on a double: the value of the lane selected by CTRL (0 where the source is
// out of range or the row is masked off).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_from(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave, returned in every lane.  DPP row shifts 1, 2, 4, 8 leave
// each 16-lane row's sum in its lane 15, row broadcasts 15 and 31 carry those into lane 63;
// the order of the additions is fixed, so every wave that sums the same values gets the
// bitwise-same result.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_from<0x111, 0xf>(v);  // row_shr:1
  v += dpp_from<0x112, 0xf>(v);  // row_shr:2
  v += dpp_from<0x114, 0xf>(v);  // row_shr:4
  v += dpp_from<0x118, 0xf>(v);  // row_shr:8
  v += dpp_from<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_from<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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
// BiCGSTAB kernels.  Vector kernels: KrylovWork::nb blocks (<= KMAXB) of KBS threads with
// grid-stride loops.  SpMV kernels: KrylovWork::nbs blocks of SpG<G>::BS threads (below).  Each
// reducing kernel writes one partial per block into its slot of w.partial; the kernel that
// consumes the sum adds the slot's ps entries itself, in every block and in the same order,
// so every block sees the bitwise-identical scalar.  Scalars that carry over to the next
// iteration live in two parity slots: iteration j writes slot (j & 1) and reads slot
// ((j + 1) & 1), which the host passes as `par`.
// PS_RS and PS_RT (<rhat,s>, <rhat,t>: the one-operand recurrence) share PS_BB's slot
// (<b,b>, read once by k_kinit2) and the last one.
enum { PS_RR = 0, PS_RHR, PS_RQ, PS_SS, PS_TS, PS_TT, PS_BB, PS_RT, PS_COUNT };
constexpr int PS_RS = PS_BB;

// Writes this block's NSL sums into entry `own` (a permutation of the block index over the
// grid; the block index by default) of their slots; a grid smaller than the stride ps also
// clears the entries no block of it owns, so every consumer sums exactly ps values.  sh: NSL
// doubles per wave of the block.
template <int BS, int NSL>
__device__ __forceinline__ void put_partials(const double (&v)[NSL], const int (&slot)[NSL], double* partial, int ps, double* sh,
                                             int own = -1) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NSL; ++j) {
    const double t = wave_sum(v[j]);
    if (lane == 0) sh[j * (BS / 64) + wid] = t;
  }
  __syncthreads();
  if (threadIdx.x < NSL) {
    const int j = threadIdx.x;
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) t += sh[j * (BS / 64) + i];
    partial[slot[j] * ps + (own < 0 ? (int)blockIdx.x : own)] = t;
  }
  for (int i = blockIdx.x + (threadIdx.x + 1) * gridDim.x; i < ps; i += BS * gridDim.x) {
#pragma unroll
    for (int j = 0; j < NSL; ++j) partial[slot[j] * ps + i] = 0.0;
  }
}

// Sums of NSL partial slots, computed by the whole block (all slots' loads in one pass);
// returned in every thread, identical in every block.
template <int BS, int NSL>
__device__ __forceinline__ void get_sums(const double* partial, const int (&slot)[NSL], int ps, double* sh, double (&out)[NSL]) {
  double s[NSL];
#pragma unroll
  for (int j = 0; j < NSL; ++j) s[j] = 0.0;
  // unrolled so that several iterations' loads are in flight together (the flat SpMV form
  // leaves thousands of partials per sum); the additions keep their order
#pragma unroll 4
  for (int i = threadIdx.x; i < ps; i += BS) {
#pragma unroll
    for (int j = 0; j < NSL; ++j) s[j] += partial[slot[j] * ps + i];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NSL; ++j) {
    const double t = wave_sum(s[j]);
    if (lane == 0) sh[j * (BS / 64) + wid] = t;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NSL; ++j) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) t += sh[j * (BS / 64) + i];
    out[j] = t;
  }
  __syncthreads();
}

template <typename TM>
__global__ void __launch_bounds__(KBS) k_kinit(BellView M, const TM* __restrict__ vals, const double* __restrict__ b, KVec v,
                                              KrylovWork w) {
  __shared__ double sh[3 * (KBS / 64)];
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double r = b[i] - spmv_row<TM>(M, vals, v.x, i);
    v.r[i] = r;
    v.rhat[i] = r;
    const double rh = r * v.invd[i];
    if (w.single) {  // phat0 = M p0 = rh; rh itself in the first half of rp's storage
      v.phat[i] = rh;
      reinterpret_cast<double*>(v.rp)[i] = rh;
    } else {
      v.rp[i] = make_double2(rh, 0.0);  // pp = M (p0 - omega0 q0) = 0 with p0 = q0 = r0, omega0 = 1
    }
    rr += r * r;
    bb += b[i] * b[i];
  }
  put_partials<KBS, 3>({rr, rr, bb}, {PS_RR, PS_RHR, PS_BB}, w.partial, w.ps, sh);
}

// Single block: atol2 and the "previous iteration" scalars seen by iteration 0.
__global__ void __launch_bounds__(KBS) k_kinit2(KrylovWork w) {
  __shared__ double sh[2 * (KBS / 64)];
  double sums[2];
  get_sums<KBS, 2>(w.partial, {PS_BB, PS_RR}, w.ps, sh, sums);
  const double bb = sums[0];
  if (threadIdx.x == 0) {
    KrylovScalars* sc = w.sc;
    sc->bs = bb;
    sc->rs_ref2[1] = sums[1];  // read by offer 0
    sc->atol2 = fmax(sc->tol * sc->tol * bb, sc->atol * sc->atol);
    sc->rho2[1] = sc->alpha2[1] = sc->omega2[1] = 1.0;
    sc->k2[1] = -1;  // so that iteration 0 counts as k = 0
    sc->done = 0;
    sc->k = 0;
  }
}

// ---- SpMV of the BiCGSTAB iteration.  A block of 1024 threads takes 4 groups of 64
// consecutive rows (one row per lane; 1 group of 256 threads on small systems, see SpG) and
// splits each row's node slots over 4 waves (the
// wave of part p takes slots p, p+4, ...), so a launch keeps four times as many waves in
// flight as a thread-per-row SpMV while every wave still reads the column-major values
// coalesced.  The four partial row sums meet in LDS and the part-0 wave adds them in fixed
// order (bitwise reproducible).
//
// Latency: the iteration's operand is a combination u + c v of two stored vectors whose
// coefficient c (beta or -alpha) comes from a global reduction the same kernel finishes.
// The SpMV therefore accumulates A u and A v separately and combines them afterwards
// (A (u + c v) = A u + c A v; equal up to rounding), so the matrix stream never waits for the
// reduction: a kernel loads its scalars and the first row's epilogue operands, issues its
// column indices before it tests the device-side `done` flag, streams values and gathers,
// and meanwhile loads the partial sums it reduces.  (Measured on the 32^3 matrix with
// 1024-thread blocks at 4 waves/SIMD: this form beats both a single-accumulator form that
// reduces first and smaller blocks at higher occupancy.)
//
// Work split: the row tiles (128 or 64 rows) are dealt out as contiguous runs, one run per
// block (a block walks its run; the next tile's epilogue operands and column indices are
// loaded before its SpMV starts).  XCD
// locality: the hardware hands workgroups to the eight XCDs round-robin (block b runs on XCD
// b % 8, each XCD with its own L2), so block b takes run xcd_rank(b), which gives every XCD
// one contiguous range of rows: its L2 then holds one slab of the gathered vector and its halo
// instead of lines from the whole vector.  Partial sums are stored at the run index.
constexpr int SP_ROWS = 64;
constexpr int SP_PARTS = 4;
constexpr int SP_NS = (27 + SP_PARTS - 1) / SP_PARTS;  // slots per part on 27-slot meshes
// Row groups per block: 2 (512-thread blocks, 128-row tiles, two blocks per CU) when the
// vector has at least two such tiles per CU, else 1 (256-thread blocks, 64-row tiles) so that
// small systems still spread over the CUs; there, when every 64-row tile can have a block of
// its own at three blocks per CU, the all-slots form (spmv2_27_all) runs instead.  Two independent blocks per CU keep one block's
// loads in flight while the other is in its reductions and epilogue (measured against one
// 1024-thread block per CU: 44.4 vs 45.6 us per iteration at 32^3, 79.8 vs 83.5 us at 40^3).
template <int G>
struct SpG {
  static constexpr int BLOCK_ROWS = SP_ROWS * G;        // rows per tile
  static constexpr int BS = SP_ROWS * SP_PARTS * G;     // threads per block
  static constexpr int RED = (SP_PARTS - 1) * G * SP_ROWS;  // LDS doubles per accumulator
};

constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_rank(int b, int nb) {
  const int x = b % N_XCD, per = nb / N_XCD, rem = nb % N_XCD;
  return x * per + min(x, rem) + b / N_XCD;
}
// Tiles [first, end) of run v of nb over n rows (every run is non-empty when nb <= tiles).
struct TileRun {
  int64_t first, end;
  __device__ TileRun(int v, int nb, int64_t n, int block_rows) {
    const int64_t tiles = (n + block_rows - 1) / block_rows;
    first = (int64_t)v * tiles / nb;
    end = (int64_t)(v + 1) * tiles / nb;
  }
};

template <int G>
struct SpmvLane {
  int lane, part, group;
  __device__ SpmvLane() : lane(threadIdx.x & 63), part((threadIdx.x >> 6) / G), group((threadIdx.x >> 6) % G) {}
  // part and group are wave-uniform: readfirstlane keeps them (and the slot bases derived from
  // them) in scalar registers, which the all-slots form needs to stay within its VGPR budget
  __device__ void make_uniform() {
    part = __builtin_amdgcn_readfirstlane(part);
    group = __builtin_amdgcn_readfirstlane(group);
  }
};

// Column nodes of one part's share of a row on a 27-slot mesh, and the row's slot count
// (boundary nodes have fewer than 27 neighbours; their padding slots hold zeros and are not
// read).
struct SlotCols {
  int32_t j[SP_NS];
  int ns;
  // UNIFORM: the slot bases are kept in scalar registers (part is wave-uniform) and the lane
  // addresses are 32-bit offsets, which saves VGPRs for the all-slots form
  template <bool UNIFORM = false>
  __device__ __forceinline__ void load(const BellView& M, int64_t r, int part_) {
    const int64_t i = r / 3;
    ns = M.nslot[i];
    if constexpr (UNIFORM) {
      const int part = __builtin_amdgcn_readfirstlane(part_);
      const uint32_t ioff = (uint32_t)i * (uint32_t)sizeof(int32_t);
#pragma unroll
      for (int k = 0; k < SP_NS; ++k) {
        const int s = part + SP_PARTS * k;
        const char* cb = reinterpret_cast<const char*>(M.col + (int64_t)(s < 27 ? s : 0) * M.n_nodes);
        j[k] = s < 27 ? *reinterpret_cast<const int32_t*>(cb + ioff) : 0;
      }
    } else {
#pragma unroll
      for (int k = 0; k < SP_NS; ++k) {
        const int s = part_ + SP_PARTS * k;
        j[k] = s < 27 ? M.col[(int64_t)s * M.n_nodes + i] : 0;
      }
    }
  }
};

// a1 += A[r, part's slots] u, a2 += A[r, part's slots] v.
template <typename TM>
__device__ __forceinline__ void spmv2_27(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                         const SlotCols& jc, int64_t r, int part, double& a1, double& a2) {
  const int64_t nrow = 3 * M.n_nodes;
#pragma unroll
  for (int k = 0; k < SP_NS; ++k) {
    const int s = part + SP_PARTS * k;
    if (s < 27 && s < jc.ns) {
      const TM* a = vals + (3 * (int64_t)s) * nrow + r;
      const int64_t jj = 3 * (int64_t)jc.j[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double m = (double)a[c * nrow];
        const double2 x = uv[jj + c];
        a1 += m * x.x;
        a2 += m * x.y;
      }
    }
  }
}

// spmv2_27 with every slot's loads issued at once, for small systems: those whose 64-row
// tiles number at most three per CU, so that every tile has a block of its own and a CU holds
// about one wave per SIMD.  The guarded loop's round trip per slot (each guard compiles to an
// exec-masked block that waits for its loads) is then the kernel's critical path, with
// nothing to overlap it.  (32-bit offsets: such systems have at most 3 x 64 rows per CU.)  A slot past the row's count reads its own in-bounds
// padding position (its column is the row's own node) and the slot index past 27 reads the
// part's first slot; both coefficients are replaced by zero.  Slot bases are wave-uniform and
// lane addresses 32-bit offsets, which keeps the seven slots in flight at ~165 VGPRs (three
// waves per SIMD).  Same products, same order as spmv2_27: the sums are bitwise equal.
template <typename TM>
__device__ __forceinline__ void spmv2_27_all(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                             const SlotCols& jc, int64_t r, int part_, double& a1, double& a2) {
  const int part = __builtin_amdgcn_readfirstlane(part_);
  const int64_t nrow = 3 * M.n_nodes;
  const uint32_t roff = (uint32_t)r * (uint32_t)sizeof(TM);
  const char* __restrict__ ub = reinterpret_cast<const char*>(uv);
  double m[SP_NS][3];
  double2 x[SP_NS][3];
#pragma unroll
  for (int k = 0; k < SP_NS; ++k) {
    const int s = part + SP_PARTS * k;
    const int su = s < 27 ? s : part;  // wave-uniform
    const bool ok = s < jc.ns;         // ns <= 27
    const char* __restrict__ vb = reinterpret_cast<const char*>(vals + 3 * (int64_t)su * nrow);
    const uint32_t goff = (uint32_t)jc.j[k] * (uint32_t)(3 * sizeof(double2));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      m[k][c] = (double)*reinterpret_cast<const TM*>(vb + c * nrow * (int64_t)sizeof(TM) + roff);
      x[k][c] = *reinterpret_cast<const double2*>(ub + goff + c * sizeof(double2));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) m[k][c] = ok ? m[k][c] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < SP_NS; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a1 += m[k][c] * x[k][c].x;
      a2 += m[k][c] * x[k][c].y;
    }
  }
}

template <typename TM, bool ALL>
__device__ __forceinline__ void spmv2_27g(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                          const SlotCols& jc, int64_t r, int part, double& a1, double& a2) {
  if constexpr (ALL)
    spmv2_27_all<TM>(M, vals, uv, jc, r, part, a1, a2);
  else
    spmv2_27<TM>(M, vals, uv, jc, r, part, a1, a2);
}

template <typename TM>
__device__ __forceinline__ void spmv2_any(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                          int64_t r, int part, double& a1, double& a2) {
  const int64_t nrow = 3 * M.n_nodes;
  const int64_t i = r / 3;
  const int ns = M.nslot[i];
  for (int s = part; s < ns; s += SP_PARTS) {
    const int64_t jj = 3 * (int64_t)M.col[(int64_t)s * M.n_nodes + i];
    const TM* a = vals + (3 * (int64_t)s) * nrow + r;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double m = (double)a[c * nrow];
      const double2 x = uv[jj + c];
      a1 += m * x.x;
      a2 += m * x.y;
    }
  }
}

// Part sums of (a1, a2) into the part-0 waves; red: 2 * SpG<G>::RED doubles.  Leaves a
// barrier pending: the caller's next __syncthreads() must precede any reuse of red.
template <int G>
__device__ __forceinline__ void part_sums_write(double a1, double a2, const SpmvLane<G>& t, double* red) {
  if (t.part) {
    const int o = ((t.part - 1) * G + t.group) * SP_ROWS + t.lane;
    red[o] = a1;
    red[SpG<G>::RED + o] = a2;
  }
}
template <int G>
__device__ __forceinline__ void part_sums_read(double& a1, double& a2, const SpmvLane<G>& t, const double* red) {
  if (t.part == 0) {
#pragma unroll
    for (int p = 1; p < SP_PARTS; ++p) {
      const int o = ((p - 1) * G + t.group) * SP_ROWS + t.lane;
      a1 += red[o];
      a2 += red[SpG<G>::RED + o];
    }
  }
}

// One row group: (A u, A v) for rows r0 + 64 group + lane, complete in the part-0 waves.
template <typename TM, int G, bool ALL>
__device__ __forceinline__ void spmv2_group(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                            int64_t r, int64_t n, const SpmvLane<G>& t, double* red, double& a1, double& a2) {
  a1 = a2 = 0.0;
  if (r < n) {
    if (M.n_slots == 27) {
      SlotCols jc;
      jc.load<ALL>(M, r, t.part);
      spmv2_27g<TM, ALL>(M, vals, uv, jc, r, t.part, a1, a2);
    } else {
      spmv2_any<TM>(M, vals, uv, r, t.part, a1, a2);
    }
  }
  part_sums_write(a1, a2, t, red);
  __syncthreads();
  part_sums_read(a1, a2, t, red);
  __syncthreads();
}

// spmv2_group with the row's 27-slot column indices already loaded (jc, valid for r < n).
template <typename TM, int G, bool ALL>
__device__ __forceinline__ void spmv2_group_pre(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                                const SlotCols& jc, int64_t r, int64_t n, const SpmvLane<G>& t, double* red,
                                                double& a1, double& a2) {
  a1 = a2 = 0.0;
  if (r < n) {
    if (M.n_slots == 27)
      spmv2_27g<TM, ALL>(M, vals, uv, jc, r, t.part, a1, a2);
    else
      spmv2_any<TM>(M, vals, uv, r, t.part, a1, a2);
  }
  part_sums_write(a1, a2, t, red);
  __syncthreads();
  part_sums_read(a1, a2, t, red);
  __syncthreads();
}

// Column indices of the row one tile further on, issued early (none past the block's run or
// past the last row; ns = 0 then).
template <bool ALL>
__device__ __forceinline__ void load_next_cols(SlotCols& jn, const BellView& M, int64_t r_next, int64_t n, bool more, int part) {
  jn.ns = 0;
  if (more && M.n_slots == 27 && r_next < n) jn.load<ALL>(M, r_next, part);
}

// Sums of NSL partial slots from registers already loaded (p[j]: this thread's share), with
// the part sums of the first row group riding on the same barrier.
template <int NSL, int G>
__device__ __forceinline__ void finish_first(double (&p)[NSL], double& a1, double& a2, const SpmvLane<G>& t, double* sh,
                                             double* red) {
  constexpr int NW = SpG<G>::BS / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NSL; ++j) {
    const double s = wave_sum(p[j]);
    if (lane == 0) sh[j * NW + wid] = s;
  }
  part_sums_write(a1, a2, t, red);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NSL; ++j) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) s += sh[j * NW + i];
    p[j] = s;
  }
  part_sums_read(a1, a2, t, red);
  __syncthreads();
}

// First row group of an SpMV kernel: column indices issued before the done test.  Returns
// false if the solve is finished (uniform over the grid).
template <typename TM, int G, bool ALL>
__device__ __forceinline__ bool first_group(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                            int64_t r, int64_t n, const SpmvLane<G>& t, int done, double& a1, double& a2) {
  a1 = a2 = 0.0;
  const bool s27 = M.n_slots == 27;
  SlotCols jc;
  if (s27 && r < n) jc.load<ALL>(M, r, t.part);
  if (done) return false;
  if (r < n) {
    if (s27)
      spmv2_27g<TM, ALL>(M, vals, uv, jc, r, t.part, a1, a2);
    else
      spmv2_any<TM>(M, vals, uv, r, t.part, a1, a2);
  }
  return true;
}

// Start of iteration j: q = A phat with phat = rh + beta pp, where beta needs <rhat,r> of
// iteration j-1; also the convergence test of iteration j-1.  Stores phat, q, qh = M q and
// the partials of <rhat,q>.
template <typename TM, int G, bool ALL>
__device__ __forceinline__ void kpq_body(const BellView& M, const TM* __restrict__ vals, const KVec& v, const KrylovWork& w,
                                         int par) {
  constexpr int BS = SpG<G>::BS, BR = SpG<G>::BLOCK_ROWS;
  __shared__ double sh[2 * (BS / 64)];
  __shared__ double red[2 * SpG<G>::RED];
  KrylovScalars* sc = w.sc;
  const int pv = par ^ 1;
  const int done = sc->done;
  const double rho = sc->rho2[pv], alpha = sc->alpha2[pv], omega = sc->omega2[pv], atol2 = sc->atol2;
  const int kprev = sc->k2[pv], maxiter = sc->maxiter;
  const double2* __restrict__ rp = v.rp;
  SpmvLane<G> t;
  if constexpr (ALL) t.make_uniform();
  const int64_t n = w.n;
  const int run = xcd_rank(blockIdx.x, gridDim.x);
  const TileRun tiles(run, gridDim.x, n, BR);
  int64_t tile = tiles.first;
  int64_t r = tile * BR + t.group * SP_ROWS + t.lane;
  double e0 = 0.0, e1 = 0.0;             // rhat, M of the row
  double2 e2 = make_double2(0.0, 0.0);  // (rh, pp) of the row
  if (t.part == 0 && r < n) {
    e0 = v.rhat[r];
    e1 = v.invd[r];
    e2 = rp[r];
  }
  double a1, a2;
  double p[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < w.ps; i += BS) {  // issued before the SpMV: independent of it
    p[0] += w.partial[PS_RR * w.ps + i];
    p[1] += w.partial[PS_RHR * w.ps + i];
  }
  if (!first_group<TM, G, ALL>(M, vals, rp, r, n, t, done, a1, a2)) return;
  SlotCols jn;  // the next tile's columns, in flight during the reductions and the epilogue
  load_next_cols<ALL>(jn, M, r + BR, n, tile + 1 < tiles.end, t.part);
  finish_first<2>(p, a1, a2, t, sh, red);
  const double rs = p[0], rho_ = p[1];
  int k = (omega == 0.0 || alpha == 0.0) ? -11 : kprev + 1;
  if (rho == 0.0) k = -10;
  const bool go = (rs > atol2) && (k < maxiter) && (k >= 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->k2[par] = k;
    sc->rho2[par] = rho_;
    sc->rs = rs;
    sc->k = k;
    if (!go) sc->done = 1;
  }
  if (!go) return;
  const double beta = rho_ / rho * alpha / omega;
  double d = 0.0;
  for (;;) {
    if (t.part == 0 && r < n) {
      const double q = __fma_rn(beta, a2, a1);
      v.phat[r] = __fma_rn(beta, e2.y, e2.x);
      v.q[r] = q;
      v.rq[r] = make_double2(e2.x, q * e1);  // (rh, qh) for k_kst, whole 16-byte pairs
      d += e0 * q;
    }
    if (++tile >= tiles.end) break;  // uniform over the block
    r += BR;
    if (t.part == 0 && r < n) {
      e0 = v.rhat[r];
      e1 = v.invd[r];
      e2 = rp[r];
    }
    const SlotCols jc = jn;
    load_next_cols<ALL>(jn, M, r + BR, n, tile + 1 < tiles.end, t.part);
    spmv2_group_pre<TM, G, ALL>(M, vals, rp, jc, r, n, t, red, a1, a2);
  }
  put_partials<BS, 1>({d}, {PS_RQ}, w.partial, w.ps, sh, run);
}

// t = A shat with shat = rh - alpha qh, alpha = rho / <rhat,q>.  Stores t and the partials
// of <t,s>, <t,t>, <s,s> (s = r - alpha q); the x/r update forms s and shat again.
template <typename TM, int G, bool ALL>
__device__ __forceinline__ void kst_body(const BellView& M, const TM* __restrict__ vals, const KVec& v, const KrylovWork& w,
                                         int par) {
  constexpr int BS = SpG<G>::BS, BR = SpG<G>::BLOCK_ROWS;
  __shared__ double sh[3 * (BS / 64)];
  __shared__ double red[2 * SpG<G>::RED];
  KrylovScalars* sc = w.sc;
  const int done = sc->done;
  const double rho = sc->rho2[par];
  const double2* __restrict__ rq = v.rq;
  SpmvLane<G> t;
  if constexpr (ALL) t.make_uniform();
  const int64_t n = w.n;
  const int run = xcd_rank(blockIdx.x, gridDim.x);
  const TileRun tiles(run, gridDim.x, n, BR);
  int64_t tile = tiles.first;
  int64_t r = tile * BR + t.group * SP_ROWS + t.lane;
  double e0 = 0.0, e1 = 0.0;  // r, q of the row
  if (t.part == 0 && r < n) {
    e0 = v.r[r];
    e1 = v.q[r];
  }
  double a1, a2;
  double p[1] = {0.0};
  for (int i = threadIdx.x; i < w.ps; i += BS) p[0] += w.partial[PS_RQ * w.ps + i];  // before the SpMV
  if (!first_group<TM, G, ALL>(M, vals, rq, r, n, t, done, a1, a2)) return;
  SlotCols jn;  // the next tile's columns, in flight during the reductions and the epilogue
  load_next_cols<ALL>(jn, M, r + BR, n, tile + 1 < tiles.end, t.part);
  finish_first<1>(p, a1, a2, t, sh, red);
  const double alpha = rho / p[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha2[par] = alpha;
  double ts = 0.0, tt = 0.0, ss = 0.0;
  for (;;) {
    if (t.part == 0 && r < n) {
      const double tv = __fma_rn(-alpha, a2, a1);
      const double s = __fma_rn(-alpha, e1, e0);  // k_kxr forms s and shat again
      v.t[r] = tv;
      ts += tv * s;
      tt += tv * tv;
      ss += s * s;
    }
    if (++tile >= tiles.end) break;  // uniform over the block
    r += BR;
    if (t.part == 0 && r < n) {
      e0 = v.r[r];
      e1 = v.q[r];
    }
    const SlotCols jc = jn;
    load_next_cols<ALL>(jn, M, r + BR, n, tile + 1 < tiles.end, t.part);
    spmv2_group_pre<TM, G, ALL>(M, vals, rq, jc, r, n, t, red, a1, a2);
  }
  put_partials<BS, 3>({ts, tt, ss}, {PS_TS, PS_TT, PS_SS}, w.partial, w.ps, sh, run);
}

// The SpMV kernels: 128-row (G = 2) or 64-row (G = 1) tiles with one slot's loads in flight at
// a time (~110 VGPRs: four waves per SIMD); and 64-row tiles with every slot's loads in flight at once
// (spmv2_27_all), three waves per SIMD, for systems whose tiles all fit the device at that
// occupancy (krylov_spmv_form).
template <typename TM, int G>
__global__ void __launch_bounds__(SpG<G>::BS) k_kpq(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  kpq_body<TM, G, false>(M, vals, v, w, par);
}
template <typename TM, int G>
__global__ void __launch_bounds__(SpG<G>::BS) k_kst(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  kst_body<TM, G, false>(M, vals, v, w, par);
}
template <typename TM>
__global__ void __launch_bounds__(SpG<1>::BS) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_kpq_all(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  kpq_body<TM, 1, true>(M, vals, v, w, par);
}
template <typename TM>
__global__ void __launch_bounds__(SpG<1>::BS) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_kst_all(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  kst_body<TM, 1, true>(M, vals, v, w, par);
}

// ---- Flat SpMV form (form 3, 27-slot meshes): one row per thread over a grid of
// ceil(n / FBS) blocks, no tile loop, every slot of the row loaded unconditionally (padding
// slots hold zeros and point at the row's own node), so a lane has its value loads and
// operand-pair gathers in flight together and finished waves are replaced by new blocks
// instead of walking a run of tiles.  Each block writes one partial per dot product (the
// stride ps is at least this grid); the consuming kernel adds all of them in index order,
// issuing the loads before its own SpMV so that they overlap it.  (Finishing the sums in the
// producing kernel instead, with the last block adding the partials, needs an agent-scope
// release fence per block, an L2 write-back on gfx950: 286 us per SpMV at 64^3.)
constexpr int FBS = 256;

// Non-temporal 16-byte load of a (rh, qh) pair whose last use this is (through a clang
// vector type: the builtin takes no HIP vector struct).  Like x and the last reads of q and
// t in the x/r update, it leaves more of the Infinity Cache to the matrix lines the next SpMV
// reads (64^3 FP64 load step 4.34 -> 4.32 s, same bits).
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 nt_load2(const double2* p) {
  const nt_d2 t = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
  return make_double2(t.x, t.y);
}

// First row of this block in the flat forms: form 3 takes row blocks in launch order; form 4
// gives every XCD one contiguous range of row blocks (xcd_rank, as the tile forms' runs), so
// its L2 holds one slab of the gathered vector and its halo.  Partial sums stay at the launch
// index.  rev (the second SpMV kernel): the same row blocks in reverse order, so that each
// SpMV starts on the rows the previous one read last.  The matrix streams twice per
// iteration and from 56^3 up it is larger than the 256 MB Infinity Cache (64^3: FP64 534 MB,
// FP32 267 MB): walked back and forth, each pass finds the other's last rows still in the
// cache (the SpMV pair at 64^3: 133.8 -> 122.0 us per launch in FP64, 82.2 -> 75.8 us with
// FP32 values; at 56^3, form 4: 92.5 -> 80.3 us in FP64).  Where the matrix fits the cache
// (48^3) it changes nothing measurable.
__device__ __forceinline__ int64_t flat_row0(const KrylovWork& w, bool rev = false) {
  const int b = w.sf == 4 ? xcd_rank((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
  return (int64_t)(rev ? (int)gridDim.x - 1 - b : b) * FBS;
}

// Node columns of row r.  With a box numbering (M.box_b > 0) a node off the box's faces
// computes them (i + dz A + dy B + dx, the ascending order its stored slots have), so only
// face nodes load theirs: 27 x 4 bytes per node and the loads' round trip before the operand
// gathers saved.  The node's (ix, iy, iz) come from two floor divisions in FP64: (i + 0.5) / A
// lies at least 0.5 / A from an integer, far beyond the product's rounding for i < 2^30.
__device__ __forceinline__ void flat_cols(const BellView& M, int64_t r, int32_t (&j)[27]) {
  const int64_t i = r / 3;
  if (M.box_b > 0) {  // uniform over the grid
    const int32_t ii = (int32_t)i;
    const int32_t iz = (int32_t)(((double)ii + 0.5) * M.box_inv_a);
    const int32_t rz = ii - iz * M.box_a;
    const int32_t iy = (int32_t)(((double)rz + 0.5) * M.box_inv_b);
    const int32_t ix = rz - iy * M.box_b;
    if ((uint32_t)(ix - 1) < (uint32_t)(M.box_b - 2) && (uint32_t)(iy - 1) < (uint32_t)(M.box_ny - 2) &&
        (uint32_t)(iz - 1) < (uint32_t)(M.box_nz - 2)) {
#pragma unroll
      for (int s = 0; s < 27; ++s) j[s] = ii + (s / 9 - 1) * M.box_a + (s / 3 % 3 - 1) * M.box_b + (s % 3 - 1);
      return;
    }
  }
#pragma unroll
  for (int s = 0; s < 27; ++s) j[s] = M.col[(int64_t)s * M.n_nodes + i];
}

// (A u, A v) of row r for the operand pair uv = (u, v), slots in order.
template <typename TM>
__device__ __forceinline__ void spmv2_flat(const BellView& M, const TM* __restrict__ vals, const double2* __restrict__ uv,
                                           int64_t r, const int32_t (&j)[27], double& a1, double& a2) {
  const int64_t nrow = 3 * M.n_nodes;
  a1 = a2 = 0.0;
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const TM* a = vals + (3 * (int64_t)s) * nrow + r;
    const int64_t jj = 3 * (int64_t)j[s];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double m = (double)a[c * nrow];
      const double2 x = uv[jj + c];
      a1 += m * x.x;
      a2 += m * x.y;
    }
  }
}

// A u of row r for one stored operand u, slots in order.
template <typename TM>
__device__ __forceinline__ double spmv1_flat(const BellView& M, const TM* __restrict__ vals, const double* __restrict__ u,
                                             int64_t r, const int32_t (&j)[27]) {
  const int64_t nrow = 3 * M.n_nodes;
  double a = 0.0;
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const TM* m = vals + (3 * (int64_t)s) * nrow + r;
    const int64_t jj = 3 * (int64_t)j[s];
#pragma unroll
    for (int c = 0; c < 3; ++c) a += (double)m[c * nrow] * u[jj + c];
  }
  return a;
}

// Block sums of NS values in fixed order, returned in every thread; sh: NS * (FBS / 64) doubles.
template <int NS>
__device__ __forceinline__ void flat_block_sums(double (&v)[NS], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const double t = wave_sum(v[j]);
    if (lane == 0) sh[j * (FBS / 64) + wid] = t;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < FBS / 64; ++k) t += sh[j * (FBS / 64) + k];
    v[j] = t;
  }
  __syncthreads();
}

// A grid smaller than the partial stride ps (small systems, where the vector kernels' grid is
// the larger) clears the entries no block of it owns, so that a consumer summing all ps entries
// (the x/r update) sums exactly this grid's partials.
template <int NS>
__device__ __forceinline__ void flat_clear_unowned(const int (&slot)[NS], double* partial, int ps) {
  for (int i = blockIdx.x + (threadIdx.x + 1) * gridDim.x; i < ps; i += FBS * gridDim.x) {
#pragma unroll
    for (int j = 0; j < NS; ++j) partial[slot[j] * ps + i] = 0.0;
  }
}

// Flat form of kpq_body (same recurrence and products).
template <typename TM>
__global__ void __launch_bounds__(FBS) k_kpq_flat(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  __shared__ double sh[2 * (FBS / 64)];
  KrylovScalars* sc = w.sc;
  const int64_t n = w.n;
  const int64_t r = flat_row0(w) + threadIdx.x;
  int32_t j[27];
  const int done = sc->done;
  if (r < n) flat_cols(M, r, j);
  if (done) return;  // uniform over the grid
  const int pv = par ^ 1;
  const double rho = sc->rho2[pv], alpha = sc->alpha2[pv], omega = sc->omega2[pv], atol2 = sc->atol2;
  const int kprev = sc->k2[pv], maxiter = sc->maxiter;
  double p[2] = {0.0, 0.0};  // <r,r>, <rhat,r>: the partials of the x/r update (w.nb blocks)
  for (int i = threadIdx.x; i < w.nb; i += FBS) {
    p[0] += w.partial[PS_RR * w.ps + i];
    p[1] += w.partial[PS_RHR * w.ps + i];
  }
  double a1 = 0.0, a2 = 0.0, e0 = 0.0, e1 = 0.0;
  double2 e2 = make_double2(0.0, 0.0);
  if (r < n) {
    e0 = v.rhat[r];
    e1 = v.invd[r];
    e2 = v.rp[r];
    spmv2_flat<TM>(M, vals, v.rp, r, j, a1, a2);
  }
  flat_block_sums<2>(p, sh);
  const double rs = p[0], rho_ = p[1];
  int k = (omega == 0.0 || alpha == 0.0) ? -11 : kprev + 1;
  if (rho == 0.0) k = -10;
  const bool go = (rs > atol2) && (k < maxiter) && (k >= 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->k2[par] = k;
    sc->rho2[par] = rho_;
    sc->rs = rs;
    sc->k = k;
    if (!go) sc->done = 1;
  }
  if (!go) return;  // uniform over the grid
  const double beta = rho_ / rho * alpha / omega;
  double d[1] = {0.0};
  if (r < n) {
    const double q = __fma_rn(beta, a2, a1);
    v.phat[r] = __fma_rn(beta, e2.y, e2.x);
    v.q[r] = q;
    v.rq[r] = make_double2(e2.x, q * e1);  // (rh, qh) for k_kst_flat, whole 16-byte pairs
    d[0] = e0 * q;
  }
  flat_block_sums<1>(d, sh);
  if (threadIdx.x == 0) w.partial[PS_RQ * w.ps + blockIdx.x] = d[0];
  flat_clear_unowned<1>({PS_RQ}, w.partial, w.ps);
}

// Flat form of kst_body.
template <typename TM>
__global__ void __launch_bounds__(FBS) k_kst_flat(BellView M, const TM* __restrict__ vals, KVec v, KrylovWork w, int par) {
  __shared__ double sh[3 * (FBS / 64)];
  KrylovScalars* sc = w.sc;
  const int64_t n = w.n;
  const int64_t r = flat_row0(w, true) + threadIdx.x;
  int32_t j[27];
  const int done = sc->done;
  if (r < n) flat_cols(M, r, j);
  if (done) return;  // uniform over the grid
  const double rho = sc->rho2[par];
  double p[1] = {0.0};  // <rhat,q>: one partial per block of k_kpq_flat (the same grid)
#pragma unroll 4
  for (int i = threadIdx.x; i < (int)gridDim.x; i += FBS) p[0] += w.partial[PS_RQ * w.ps + i];
  double a1 = 0.0, a2 = 0.0, e0 = 0.0, e1 = 0.0;
  if (r < n) {
    e0 = v.r[r];
    e1 = v.q[r];
    spmv2_flat<TM>(M, vals, v.rq, r, j, a1, a2);
  }
  flat_block_sums<1>(p, sh);
  const double alpha = rho / p[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha2[par] = alpha;
  double d[3] = {0.0, 0.0, 0.0};
  if (r < n) {
    const double tv = __fma_rn(-alpha, a2, a1);
    const double s = __fma_rn(-alpha, e1, e0);  // k_kxr forms s and shat again
    v.t[r] = tv;
    d[0] = s * s;
    d[1] = tv * s;
    d[2] = tv * tv;
  }
  flat_block_sums<3>(d, sh);
  if (threadIdx.x == 0) {
    w.partial[PS_SS * w.ps + blockIdx.x] = d[0];
    w.partial[PS_TS * w.ps + blockIdx.x] = d[1];
    w.partial[PS_TT * w.ps + blockIdx.x] = d[2];
  }
  flat_clear_unowned<3>({PS_SS, PS_TS, PS_TT}, w.partial, w.ps);
}

// ---- One-operand recurrence (FP64 values, flat form).  The x/r update forms the next
// phat = rh + beta (phat - omega qh) itself, with beta from <rhat, r_new> = <rhat,s> -
// omega <rhat,t> (sums the second SpMV kernel adds for it: the improved BiCGSTAB of Yang and
// Brent, equal to JAX's recurrence in exact arithmetic, alpha and the convergence test still
// from the directly summed <rhat,r> and <r,r>), so the first SpMV gathers one 8-byte operand
// per column entry instead of the (rh, pp) pair.  rh is kept in the first half of rp's
// storage.  (The mixed form keeps the pair: its residual replacements change r between the
// x/r update and the first SpMV.)
__global__ void __launch_bounds__(FBS) k_kpq_flat1(BellView M, const double* __restrict__ vals, KVec v, KrylovWork w,
                                                   int par) {
  __shared__ double sh[2 * (FBS / 64)];
  KrylovScalars* sc = w.sc;
  const int64_t n = w.n;
  const int64_t r = flat_row0(w) + threadIdx.x;
  int32_t j[27];
  const int done = sc->done;
  if (r < n) flat_cols(M, r, j);
  if (done) return;  // uniform over the grid
  const int pv = par ^ 1;
  const double rho = sc->rho2[pv], alpha = sc->alpha2[pv], omega = sc->omega2[pv], atol2 = sc->atol2;
  const int kprev = sc->k2[pv], maxiter = sc->maxiter;
  double p[2] = {0.0, 0.0};  // <r,r>, <rhat,r>: the partials of the x/r update (w.nb blocks)
  for (int i = threadIdx.x; i < w.nb; i += FBS) {
    p[0] += w.partial[PS_RR * w.ps + i];
    p[1] += w.partial[PS_RHR * w.ps + i];
  }
  const double* __restrict__ rhv = reinterpret_cast<const double*>(v.rp);
  double a = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
  if (r < n) {
    e0 = v.rhat[r];
    e1 = v.invd[r];
    e2 = __builtin_nontemporal_load(rhv + r);
    a = spmv1_flat<double>(M, vals, v.phat, r, j);
  }
  flat_block_sums<2>(p, sh);
  const double rs = p[0], rho_ = p[1];
  int k = (omega == 0.0 || alpha == 0.0) ? -11 : kprev + 1;
  if (rho == 0.0) k = -10;
  const bool go = (rs > atol2) && (k < maxiter) && (k >= 0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->k2[par] = k;
    sc->rho2[par] = rho_;
    sc->rs = rs;
    sc->k = k;
    if (!go) sc->done = 1;
  }
  if (!go) return;  // uniform over the grid
  double d[1] = {0.0};
  if (r < n) {
    v.q[r] = a;
    v.rq[r] = make_double2(e2, a * e1);  // (rh, qh) for the second SpMV
    d[0] = e0 * a;
  }
  flat_block_sums<1>(d, sh);
  if (threadIdx.x == 0) w.partial[PS_RQ * w.ps + blockIdx.x] = d[0];
  flat_clear_unowned<1>({PS_RQ}, w.partial, w.ps);
}

// k_kst_flat with the partials of <rhat,s> and <rhat,t> as well.  It stores only t: the x/r
// update recomputes s = r - alpha q and shat = rh - alpha qh (same operations, same bits) from
// r, q and the (rh, qh) pair it reads anyway, two vector writes fewer per iteration.
__global__ void __launch_bounds__(FBS) k_kst_flat1(BellView M, const double* __restrict__ vals, KVec v, KrylovWork w,
                                                   int par) {
  __shared__ double sh[5 * (FBS / 64)];
  KrylovScalars* sc = w.sc;
  const int64_t n = w.n;
  const int64_t r = flat_row0(w, true) + threadIdx.x;
  int32_t j[27];
  const int done = sc->done;
  if (r < n) flat_cols(M, r, j);
  if (done) return;  // uniform over the grid
  const double rho = sc->rho2[par];
  double p[1] = {0.0};  // <rhat,q>: one partial per block of k_kpq_flat1 (the same grid)
#pragma unroll 4
  for (int i = threadIdx.x; i < (int)gridDim.x; i += FBS) p[0] += w.partial[PS_RQ * w.ps + i];
  double a1 = 0.0, a2 = 0.0, e0 = 0.0, e1 = 0.0, e3 = 0.0;
  if (r < n) {
    e0 = v.r[r];
    e1 = v.q[r];
    e3 = v.rhat[r];
    spmv2_flat<double>(M, vals, v.rq, r, j, a1, a2);
  }
  flat_block_sums<1>(p, sh);
  const double alpha = rho / p[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->alpha2[par] = alpha;
  double d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (r < n) {
    const double tv = __fma_rn(-alpha, a2, a1);
    const double s = __fma_rn(-alpha, e1, e0);  // k_kxr1 forms s and shat again from r, q, rh, qh
    v.t[r] = tv;
    d[0] = s * s;
    d[1] = tv * s;
    d[2] = tv * tv;
    d[3] = e3 * s;
    d[4] = e3 * tv;
  }
  flat_block_sums<5>(d, sh);
  if (threadIdx.x == 0) {
    w.partial[PS_SS * w.ps + blockIdx.x] = d[0];
    w.partial[PS_TS * w.ps + blockIdx.x] = d[1];
    w.partial[PS_TT * w.ps + blockIdx.x] = d[2];
    w.partial[PS_RS * w.ps + blockIdx.x] = d[3];
    w.partial[PS_RT * w.ps + blockIdx.x] = d[4];
  }
  flat_clear_unowned<5>({PS_SS, PS_TS, PS_TT, PS_RS, PS_RT}, w.partial, w.ps);
}

// k_kxr for the one-operand recurrence: also forms the next phat, and s and shat itself.
// x (read and written only here) and the last reads of q and t are non-temporal, so that they
// displace fewer of the matrix lines the Infinity Cache carries from one SpMV to the next
// (64^3 load step 4.39 -> 4.36 s FP64, 3.38 -> 3.36 s mixed, same bits).
__global__ void __launch_bounds__(KBS) k_kxr1(KVec v, KrylovWork w, int par) {
  __shared__ double sh[5 * (KBS / 64)];
  KrylovScalars* sc = w.sc;
  const int done = sc->done;
  const double alpha = sc->alpha2[par], rho = sc->rho2[par], atol2 = sc->atol2;
  const int64_t i0 = (int64_t)blockIdx.x * KBS + threadIdx.x;
  double x = 0, ph = 0, r0 = 0, q = 0, t = 0, rhat = 0, invd = 0;
  double2 rq = make_double2(0.0, 0.0);
  if (i0 < w.n) {  // issued before the done test (one round trip for both)
    x = __builtin_nontemporal_load(v.x + i0); ph = v.phat[i0]; r0 = v.r[i0]; q = __builtin_nontemporal_load(v.q + i0); t = __builtin_nontemporal_load(v.t + i0);
    rq = nt_load2(v.rq + i0); rhat = v.rhat[i0]; invd = v.invd[i0];
  }
  if (done) return;
  double sums[5];
  get_sums<KBS, 5>(w.partial, {PS_SS, PS_TS, PS_TT, PS_RS, PS_RT}, w.ps, sh, sums);
  const double ss = sums[0];
  const double omega = sums[1] / sums[2];
  const double beta = (sums[3] - omega * sums[4]) / rho * alpha / omega;  // of the next iteration
  const bool early = ss < atol2;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omega2[par] = omega;
  double* __restrict__ rhv = reinterpret_cast<double*>(v.rp);
  double rr = 0.0, rhr = 0.0;
  for (int64_t i = i0; i < w.n; i += (int64_t)w.nb * KBS) {
    if (i != i0) {
      x = __builtin_nontemporal_load(v.x + i); ph = v.phat[i]; r0 = v.r[i]; q = __builtin_nontemporal_load(v.q + i); t = __builtin_nontemporal_load(v.t + i);
      rq = nt_load2(v.rq + i); rhat = v.rhat[i]; invd = v.invd[i];
    }
    const double s = __fma_rn(-alpha, q, r0), shv = __fma_rn(-alpha, rq.y, rq.x), qh = rq.y;
    double r;
    if (early) {
      __builtin_nontemporal_store(x + alpha * ph, v.x + i);
      r = s;
    } else {
      __builtin_nontemporal_store(x + (alpha * ph + omega * shv), v.x + i);
      r = s - omega * t;
    }
    v.r[i] = r;
    const double rh = r * invd;
    rhv[i] = rh;
    v.phat[i] = __fma_rn(beta, __fma_rn(-omega, qh, ph), rh);
    rr += r * r;
    rhr += rhat * r;
  }
  put_partials<KBS, 2>({rr, rhr}, {PS_RR, PS_RHR}, w.partial, w.ps, sh);
}

// omega, x and r updates; stores rh = M r, pp = phat - omega qh and the partials of <r,r>,
// <rhat,r>.  s = r - alpha q and shat = rh - alpha qh are formed here again from r, q and the
// (rh, qh) pair (the second SpMV kernel stores only t; the same operations give the same
// bits).  Scalars, the first row's operands and the partials are loaded together; x and the
// last reads of q and t are non-temporal, as in k_kxr1.
__global__ void __launch_bounds__(KBS) k_kxr(KVec v, KrylovWork w, int par) {
  __shared__ double sh[3 * (KBS / 64)];
  KrylovScalars* sc = w.sc;
  const int done = sc->done;
  const double alpha = sc->alpha2[par], atol2 = sc->atol2;
  const int64_t i0 = (int64_t)blockIdx.x * KBS + threadIdx.x;
  double x = 0, ph = 0, r0 = 0, q = 0, t = 0, rhat = 0, invd = 0;
  double2 rq = make_double2(0.0, 0.0);
  if (i0 < w.n) {  // issued before the done test (one round trip for both)
    x = __builtin_nontemporal_load(v.x + i0); ph = v.phat[i0]; r0 = v.r[i0]; q = __builtin_nontemporal_load(v.q + i0); t = __builtin_nontemporal_load(v.t + i0);
    rq = nt_load2(v.rq + i0); rhat = v.rhat[i0]; invd = v.invd[i0];
  }
  if (done) return;
  double sums[3];
  get_sums<KBS, 3>(w.partial, {PS_SS, PS_TS, PS_TT}, w.ps, sh, sums);
  const double ss = sums[0];
  const double omega = sums[1] / sums[2];
  const bool early = ss < atol2;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->omega2[par] = omega;
  double rr = 0.0, rhr = 0.0;
  for (int64_t i = i0; i < w.n; i += (int64_t)w.nb * KBS) {
    if (i != i0) {
      x = __builtin_nontemporal_load(v.x + i); ph = v.phat[i]; r0 = v.r[i]; q = __builtin_nontemporal_load(v.q + i); t = __builtin_nontemporal_load(v.t + i);
      rq = nt_load2(v.rq + i); rhat = v.rhat[i]; invd = v.invd[i];
    }
    const double s = __fma_rn(-alpha, q, r0), shv = __fma_rn(-alpha, rq.y, rq.x), qh = rq.y;
    double r;
    if (early) {
      __builtin_nontemporal_store(x + alpha * ph, v.x + i);
      r = s;
    } else {
      __builtin_nontemporal_store(x + (alpha * ph + omega * shv), v.x + i);
      r = s - omega * t;
    }
    v.r[i] = r;
    const double rh = r * invd;
    v.rp[i] = make_double2(rh, __fma_rn(-omega, qh, ph));
    rr += r * r;
    rhr += rhat * r;
  }
  put_partials<KBS, 2>({rr, rhr}, {PS_RR, PS_RHR}, w.partial, w.ps, sh);
}

// Reliable update of the mixed form: replace the recurrence residual by the true FP64
// residual r = b - A64 x (with rh and the partial sums), without restarting the Krylov space.
// The host offers a replacement after every batch of iterations; it happens once the
// recurrence residual has fallen by RU_DROP (in <r,r>) since the last one, so the Krylov
// process is perturbed a few times per solve while the drift between the recurrence and the
// true residual stays at the level of the FP32 operator error times the last drop.
constexpr double RU_DROP = 1e-4;

__global__ void __launch_bounds__(KBS) k_kreplace(BellView M, const double* __restrict__ vals, const double* __restrict__ b,
                                                 KVec v, KrylovWork w, int par) {
  __shared__ double sh[2 * (KBS / 64)];
  KrylovScalars* sc = w.sc;
  if (sc->done) return;
  const double ref = sc->rs_ref2[par ^ 1], rs = sc->rs;
  const bool now = rs <= RU_DROP * ref;  // the same in every block
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->rs_ref2[par] = now ? rs : ref;
  if (!now) return;
  double rr = 0.0, rhr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x; i < w.n; i += (int64_t)w.nb * KBS) {
    const double r = b[i] - spmv_row<double>(M, vals, v.x, i);
    v.r[i] = r;
    const double rh = r * v.invd[i];
    v.rp[i].x = rh;
    rr += r * r;
    rhr += v.rhat[i] * r;
  }
  put_partials<KBS, 2>({rr, rhr}, {PS_RR, PS_RHR}, w.partial, w.ps, sh);
}

// Generic dot product a.b into *out: k_dot writes one partial per block, k_dot_finish (one
// block) adds them in index order.  (Finishing in k_dot's last block needs an agent-scope
// fence in every block, an L2 write-back on gfx950: 68 us for one 64^3 vector, against
// 4.6 + 4.7 us for the two kernels.)
__global__ void __launch_bounds__(KBS) k_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* __restrict__ partial) {
  __shared__ double sh[KBS / 64];
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  double d[1] = {i < n ? a[i] * b[i] : 0.0};
  block_sums<1>(d, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = d[0];
}
__global__ void __launch_bounds__(KBS) k_dot_finish(const double* __restrict__ partial, int nb, double* __restrict__ out) {
  __shared__ double sh[KBS / 64];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += KBS) acc[0] += partial[i];
  block_sums<1>(acc, sh);
  if (threadIdx.x == 0) *out = acc[0];
}

__global__ void __launch_bounds__(KBS) k_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i < n) y[i] = x[i];
}

__global__ void __launch_bounds__(KBS) k_fill(int64_t n, double v, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * KBS + threadIdx.x;
  if (i < n) y[i] = v;
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
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cus[dev] == 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  return cus[dev];
}

// SpMV form picked by size: on 27-slot meshes 3 = one row per thread (at least FLAT_MIN_ROWS
// rows) and 4 = one row per thread over XCD-local row ranges (from FLAT_XCD_MIN_ROWS rows);
// otherwise 2 = 128-row tiles (at least two per CU), 0 = 64-row tiles with all slot loads in
// flight (all tiles resident at three 256-thread blocks per CU), 1 = 64-row tiles.  Measured
// on the dual-phase load step (s per load step, FP64 / mixed; form 2 against 3 against 4):
// 32^3 (107,811 rows) 0.384 / 0.335, 0.372 / 0.334, 0.366 / 0.336; 40^3 0.809 / 0.681, 0.808
// / 0.701, 0.767 / 0.650; 48^3 1.557 / 1.184, 1.677 / 1.307, 1.384 / 1.072; 56^3 3.138 /
// 1.989, 3.002 / 1.975, 3.030 / 1.784; 64^3 (823,875 rows) forms 3 and 4: 4.638 / 3.513
// against 4.792 / 3.704.  (The XCD-local ranges keep each XCD's gathers in its own L2 slab;
// at 64^3 the plain launch order streams better.  These are one-direction figures: with the
// second SpMV walking back, form 4 at 56^3 takes 2.724 / 1.720 s.)  Below 100,000 rows the
// tile forms stay:
// 16^3 0.105 / 0.117 with form 0 against 0.130 / 0.130 with form 4, 24^3 0.184 / 0.199
// against 0.198 / 0.197; 28^3 (73,167 rows) 0.302 / 0.279 with form 2 against 0.291 / 0.268.
constexpr int64_t FLAT_MIN_ROWS = 700000;
constexpr int64_t FLAT_XCD_MIN_ROWS = 100000;
int krylov_spmv_form(int64_t n, int n_slots) {
  const int cus = device_cus();
  if (n_slots == 27 && n >= FLAT_MIN_ROWS) return 3;
  if (n_slots == 27 && n >= FLAT_XCD_MIN_ROWS) return 4;
  if ((n + SpG<2>::BLOCK_ROWS - 1) / SpG<2>::BLOCK_ROWS >= 2 * cus) return 2;
  return (n + SpG<1>::BLOCK_ROWS - 1) / SpG<1>::BLOCK_ROWS <= 3 * cus ? 0 : 1;
}

// SpMV blocks: a grid that fits the device in one round, each block walking a run of tiles
// (at the serial forms' register use a CU holds 16 waves: two 512-thread or four 256-thread
// blocks; the all-slots form holds three 256-thread blocks and is only picked by size when
// every 64-row tile gets a block of its own; a grid beyond that runs in two rounds: on the 32^3
// matrix with 1024-thread blocks, 256 blocks took 46.3 us per BiCGSTAB iteration, 384 blocks
// 61.7 us, 422 blocks 50.4 us).
int krylov_spmv_blocks(int64_t n, int form) {
  if (form == 3 || form == 4) return (int)std::max<int64_t>(1, (n + FBS - 1) / FBS);
  const int64_t per_cu = form == 2 ? 2 : form == 0 ? 3 : 4;
  const int64_t rows = form == 2 ? SpG<2>::BLOCK_ROWS : SpG<1>::BLOCK_ROWS;
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + rows - 1) / rows, per_cu * device_cus()));
}

// Whether a form may run on n rows: the all-slots form keeps row and column byte offsets in
// 32 bits, which holds below 2^28 rows.  (Forced on the 64^3 system, 823,875 rows, where its
// grid of three blocks per CU walks runs of 64-row tiles, it takes 153.5 us per SpMV against
// 148 us for the 128-row form.)
bool krylov_form_allowed(int64_t n, int form, int n_slots) {
  if (form == 1 || form == 2) return true;
  if (form == 3 || form == 4) return n_slots == 27 && (n + FBS - 1) / FBS <= INT32_MAX;
  return form == 0 && n < ((int64_t)1 << 28);
}

// Partial-sum stride that holds every form's grid (the form can be set per context).
int krylov_stride_max(int64_t n) {
  return std::max(std::max(krylov_blocks(n), krylov_spmv_blocks(n, 3)), std::max(krylov_spmv_blocks(n, 1), krylov_spmv_blocks(n, 2)));
}

void krylov_replace(const BellView& M, const double* vals64, const double* b, KVec v, KrylovWork w, int offer,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_kreplace, dim3((unsigned)w.nb), dim3(KBS), 0, st, M, vals64, b, v, w, offer & 1);
}

// Launch, timed with kernel-bound events when the profiler samples it.
template <typename F, typename... Args>
static inline void timed_launch(Prof* prof, int cls, F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args) {
  hipEvent_t a, b;
  if (prof_sample(prof, cls, &a, &b))
    hipExtLaunchKernelGGL(kernel, grid, block, 0, st, a, b, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
}

template <typename TM>
void krylov_iteration(const BellView& M, const TM* vals, KVec v, KrylovWork w, int par, hipStream_t st, Prof* prof) {
  const unsigned nb = (unsigned)w.nb, nbs = (unsigned)w.nbs;
  if (w.sf >= 3) {
    if constexpr (std::is_same<TM, double>::value) {
      if (w.single) {
        timed_launch(prof, PC_SPMV, k_kpq_flat1, dim3(nbs), dim3(FBS), st, M, vals, v, w, par);
        timed_launch(prof, PC_SPMV, k_kst_flat1, dim3(nbs), dim3(FBS), st, M, vals, v, w, par);
        timed_launch(prof, PC_KVEC, k_kxr1, dim3(nb), dim3(KBS), st, v, w, par);
        return;
      }
    }
    timed_launch(prof, PC_SPMV, k_kpq_flat<TM>, dim3(nbs), dim3(FBS), st, M, vals, v, w, par);
    timed_launch(prof, PC_SPMV, k_kst_flat<TM>, dim3(nbs), dim3(FBS), st, M, vals, v, w, par);
  } else if (w.sf == 2) {
    timed_launch(prof, PC_SPMV, k_kpq<TM, 2>, dim3(nbs), dim3(SpG<2>::BS), st, M, vals, v, w, par);
    timed_launch(prof, PC_SPMV, k_kst<TM, 2>, dim3(nbs), dim3(SpG<2>::BS), st, M, vals, v, w, par);
  } else if (w.sf == 0) {
    timed_launch(prof, PC_SPMV, k_kpq_all<TM>, dim3(nbs), dim3(SpG<1>::BS), st, M, vals, v, w, par);
    timed_launch(prof, PC_SPMV, k_kst_all<TM>, dim3(nbs), dim3(SpG<1>::BS), st, M, vals, v, w, par);
  } else {
    timed_launch(prof, PC_SPMV, k_kpq<TM, 1>, dim3(nbs), dim3(SpG<1>::BS), st, M, vals, v, w, par);
    timed_launch(prof, PC_SPMV, k_kst<TM, 1>, dim3(nbs), dim3(SpG<1>::BS), st, M, vals, v, w, par);
  }
  timed_launch(prof, PC_KVEC, k_kxr, dim3(nb), dim3(KBS), st, v, w, par);
}

void launch_spmv(const BellView& M, const double* vals, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_spmv, dim3(nblk(3 * M.n_nodes)), dim3(KBS), 0, st, M, vals, x, y);
}

void launch_inv_diag(const BellView& M, const double* vals, const int32_t* dslot, double* invd, int precond, hipStream_t st) {
  hipLaunchKernelGGL(k_inv_diag, dim3(nblk(3 * M.n_nodes)), dim3(KBS), 0, st, M, vals, dslot, invd, precond);
}

void launch_residual_norm(int64_t n, const double* a, const double* b, double* partial, int nb, double* out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_dot, dim3((unsigned)nb), dim3(KBS), 0, st, n, a, b, partial);
  hipLaunchKernelGGL(k_dot_finish, dim3(1), dim3(KBS), 0, st, partial, nb, out);
}
void launch_copy(int64_t n, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(KBS), 0, st, n, x, y);
}
void launch_fill(int64_t n, double v, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(KBS), 0, st, n, v, y);
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
