// Banded direct solve of the Newton system for the drivers that ask for a direct solver
// ('umfpack_solver' in crystal_plasticity_OR_design/solver.py:100-101, used by
// calibration_case4_UQ_polyCrystalSteel_1D_GB.py:77,90 on a 20^3 mesh).
//
// The block-ELL matrix (Dirichlet rows already replaced by unit rows) is copied into LAPACK-
// style band storage under a node ordering that keeps the bandwidth small (the mesh's own
// order or reverse Cuthill-McKee, whichever is narrower), factored in place as A = L U
// without pivoting (right-looking, panels of BAND_NB columns: the diagonal block in one
// wave's registers, then L21 = A21 U11^-1 and U12 = L11^-1 A12 one row / column per thread in
// LDS, then an LDS-tiled trailing update A22 -= L21 U12), and solved with two triangular
// sweeps of two launches per panel.  FP64 iterative refinement against the block-ELL matrix follows in the
// caller.  Without pivoting the factorisation exists when every leading principal minor is
// non-zero, which holds for these stiffness matrices (elastic-plastic tangents close to
// SPD, unit Dirichlet rows); a zero or non-finite pivot is reported instead of being worked
// round, and the caller turns it into an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <vector>

#include "cp_internal.hpp"

namespace cpfem {

constexpr int BAND_NB = 32;     // panel width
constexpr int BAND_TILE = 64;   // trailing-update tile (rows x columns)

// Element (i, j) of the band (valid when -kl <= j - i <= ku).
__device__ __forceinline__ int64_t bidx(int64_t i, int64_t j, int kl, int W) { return i * W + (j - i + kl); }

// Copy of the block-ELL matrix into the band under the node map ip (old -> new).
__global__ void k_band_scatter(BellView M, const double* __restrict__ vals, const int32_t* __restrict__ ip,
                               double* __restrict__ band, int kl, int W) {
  const int64_t n = 3 * M.n_nodes;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t i = r / 3;
  const int c = (int)(r - 3 * i);
  const int64_t R = 3 * (int64_t)ip[i] + c;
  const int ns = M.nslot[i];
  for (int s = 0; s < ns; ++s) {
    const int64_t j = M.col[(int64_t)s * M.n_nodes + i];
    const int64_t C0 = 3 * (int64_t)ip[j];
#pragma unroll
    for (int k = 0; k < 3; ++k) band[bidx(R, C0 + k, kl, W)] = vals[(3 * (int64_t)s + k) * n + r];
  }
}

// Diagonal block of panel k0 (nb <= BAND_NB rows and columns): A11 = L11 U11 in registers
// of one wave, lane p holding row k0+p; the pivot row is broadcast with shuffles.  Entries
// outside the band are zero and stay zero.
__global__ void __launch_bounds__(64) k_band_diag(double* __restrict__ band, int kl, int ku, int W, int64_t k0, int nb,
                                                 int32_t* __restrict__ flag) {
  const int p = threadIdx.x;
  double a[BAND_NB];
#pragma unroll
  for (int q = 0; q < BAND_NB; ++q)
    a[q] = (p < nb && q < nb && q - p <= ku && p - q <= kl) ? band[bidx(k0 + p, k0 + q, kl, W)] : 0.0;
  bool bad = false;
#pragma unroll
  for (int c = 0; c < BAND_NB; ++c) {
    const bool act = c < nb;  // uniform; no early exit, so a[] stays in registers
    const double piv = __shfl(a[c], c);
    bad |= act && !(piv != 0.0 && std::isfinite(piv));
    const double l = act ? a[c] / piv : 0.0;
    if (act && p > c) a[c] = l;
#pragma unroll
    for (int q = c + 1; q < BAND_NB; ++q) {
      const double u = __shfl(a[q], c);
      if (act && p > c) a[q] -= l * u;
    }
  }
  if (p == 0 && bad) flag[0] = 1;
#pragma unroll
  for (int q = 0; q < BAND_NB; ++q)
    if (p < nb && q < nb && q - p <= ku && p - q <= kl) band[bidx(k0 + p, k0 + q, kl, W)] = a[q];
}

// Rows (columns) per block of the L21 (U12) kernels: small blocks spread a panel's ~kl rows
// over many CUs.
constexpr int LR_BS = 64;

// L21 = A21 U11^-1: one thread per row below the full panel k0..k0+BAND_NB-1, U11 and each
// thread's row of panel entries in LDS (register arrays of this size spill).
__global__ void __launch_bounds__(LR_BS) k_band_lrow(double* __restrict__ band, int64_t n, int kl, int ku, int W, int64_t k0) {
  __shared__ double U[BAND_NB][BAND_NB + 1];
  __shared__ double A[BAND_NB][LR_BS + 1];  // A[q][thread]
  for (int e = threadIdx.x; e < BAND_NB * BAND_NB; e += LR_BS) {
    const int p = e / BAND_NB, q = e % BAND_NB;
    U[p][q] = (q >= p && q - p <= ku) ? band[bidx(k0 + p, k0 + q, kl, W)] : 0.0;
  }
  const int64_t i0 = k0 + BAND_NB + (int64_t)blockIdx.x * LR_BS, ilast = std::min<int64_t>(n - 1, k0 + BAND_NB - 1 + kl);
  // rows read and written as contiguous lines: consecutive threads take consecutive columns
  for (int e = threadIdx.x; e < LR_BS * BAND_NB; e += LR_BS) {
    const int r = e / BAND_NB, q = e % BAND_NB;
    const int64_t i = i0 + r;
    A[q][r] = (i <= ilast && i - (k0 + q) <= kl) ? band[bidx(i, k0 + q, kl, W)] : 0.0;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (i0 + t <= ilast) {
    for (int p = 0; p < BAND_NB; ++p) {
      double v = A[p][t];
      for (int q = 0; q < p; ++q) v -= A[q][t] * U[q][p];
      A[p][t] = v / U[p][p];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < LR_BS * BAND_NB; e += LR_BS) {
    const int r = e / BAND_NB, q = e % BAND_NB;
    const int64_t i = i0 + r;
    if (i <= ilast && i - (k0 + q) <= kl) band[bidx(i, k0 + q, kl, W)] = A[q][r];
  }
}

// U12 = L11^-1 A12: one thread per column right of the full panel, L11 (unit lower) and each
// thread's column of panel entries in LDS.
__global__ void __launch_bounds__(LR_BS) k_band_urow(double* __restrict__ band, int64_t n, int kl, int ku, int W, int64_t k0) {
  __shared__ double L[BAND_NB][BAND_NB + 1];
  __shared__ double U[BAND_NB][LR_BS + 1];  // U[p][thread]
  for (int e = threadIdx.x; e < BAND_NB * BAND_NB; e += LR_BS) {
    const int p = e / BAND_NB, q = e % BAND_NB;
    L[p][q] = (q < p && p - q <= kl) ? band[bidx(k0 + p, k0 + q, kl, W)] : 0.0;
  }
  const int t = threadIdx.x;
  const int64_t j = k0 + BAND_NB + (int64_t)blockIdx.x * LR_BS + t;
  const bool on = j < n && j <= k0 + BAND_NB - 1 + ku;
  for (int p = 0; p < BAND_NB; ++p) U[p][t] = (on && j - (k0 + p) <= ku) ? band[bidx(k0 + p, j, kl, W)] : 0.0;
  __syncthreads();
  if (!on) return;
  for (int p = 1; p < BAND_NB; ++p) {
    double v = U[p][t];
    for (int q = 0; q < p; ++q) v -= L[p][q] * U[q][t];
    U[p][t] = v;
  }
  for (int p = 0; p < BAND_NB; ++p)
    if (j - (k0 + p) <= ku) band[bidx(k0 + p, j, kl, W)] = U[p][t];
}

// Trailing update A22 -= L21 U12 over BAND_TILE x BAND_TILE tiles (256 threads, 4 x 4 each).
__global__ void __launch_bounds__(256) k_band_trail(double* __restrict__ band, int64_t n, int kl, int ku, int W, int64_t k0,
                                                   int nb) {
  __shared__ double Ls[BAND_TILE][BAND_NB + 1];
  __shared__ double Us[BAND_NB][BAND_TILE + 1];
  const int64_t base = k0 + nb;
  const int64_t i0 = base + (int64_t)blockIdx.y * BAND_TILE, j0 = base + (int64_t)blockIdx.x * BAND_TILE;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int e = threadIdx.x; e < BAND_TILE * BAND_NB; e += 256) {
    const int a = e / BAND_NB, p = e % BAND_NB;  // L: row a of the tile, panel column p
    const int64_t i = i0 + a, c = k0 + p;
    Ls[a][p] = (p < nb && i < n && i - c <= kl) ? band[bidx(i, c, kl, W)] : 0.0;
    const int pu = e / BAND_TILE, b = e % BAND_TILE;  // U: panel row pu, column b of the tile
    const int64_t rr = k0 + pu, j = j0 + b;
    Us[pu][b] = (pu < nb && j < n && j - rr <= ku) ? band[bidx(rr, j, kl, W)] : 0.0;
  }
  __syncthreads();
  double acc[4][4] = {};
#pragma unroll 8
  for (int p = 0; p < BAND_NB; ++p) {
    double lv[4], uv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) lv[a] = Ls[ty + 16 * a][p];
#pragma unroll
    for (int b = 0; b < 4; ++b) uv[b] = Us[p][tx + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += lv[a] * uv[b];
  }
  const int64_t iend = std::min<int64_t>(k0 + nb - 1 + kl, n - 1), jend = std::min<int64_t>(k0 + nb - 1 + ku, n - 1);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t i = i0 + ty + 16 * a;
    if (i > iend) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + tx + 16 * b;
      if (j > jend || j - i > ku || i - j > kl) continue;
      band[bidx(i, j, kl, W)] -= acc[a][b];
    }
  }
}

// Triangular sweeps, two launches per panel: one wave solves the panel rows (lane p holds
// row k0+p with its panel row of L or U preloaded; rows are finalised in order and broadcast
// by shuffles), then a grid over the rows the panel reaches subtracts the panel's
// contribution.
__global__ void __launch_bounds__(64) k_band_fwd_diag(const double* __restrict__ band, double* __restrict__ y, int kl, int W,
                                                     int64_t k0, int nb) {
  const int p = threadIdx.x;
  double lr[BAND_NB];
#pragma unroll
  for (int q = 0; q < BAND_NB; ++q) lr[q] = (q < p && p < nb && p - q <= kl) ? band[bidx(k0 + p, k0 + q, kl, W)] : 0.0;
  double v = p < nb ? y[k0 + p] : 0.0;
#pragma unroll
  for (int q = 0; q < BAND_NB; ++q) {
    const double vq = __shfl(v, q);  // row k0+q is final here
    v -= lr[q] * vq;
  }
  if (p < nb) y[k0 + p] = v;
}

// 8 lanes per row, 4 consecutive panel columns each (a row's 32 entries are one contiguous
// 256-byte run), reduced with shuffles; 32 rows per 256-thread block, so a panel's rows spread
// over many CUs.
constexpr int UPD_LANES = 8;
__global__ void __launch_bounds__(256) k_band_fwd_upd(const double* __restrict__ band, double* __restrict__ y, int64_t n, int kl,
                                                     int W, int64_t k0, int nb) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t i = k0 + nb + g / UPD_LANES;
  const int sub = (int)(g % UPD_LANES);
  const bool on = i < n && i <= k0 + nb - 1 + kl;
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < BAND_NB / UPD_LANES; ++c) {
    const int p = sub * (BAND_NB / UPD_LANES) + c;
    if (on && p < nb && i - (k0 + p) <= kl) acc += band[bidx(i, k0 + p, kl, W)] * y[k0 + p];
  }
#pragma unroll
  for (int off = UPD_LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, UPD_LANES);
  if (on && sub == 0) y[i] -= acc;
}

__global__ void __launch_bounds__(64) k_band_bwd_diag(const double* __restrict__ band, double* __restrict__ y, int kl, int ku,
                                                     int W, int64_t k0, int nb) {
  const int p = threadIdx.x;
  double ur[BAND_NB];
#pragma unroll
  for (int q = 0; q < BAND_NB; ++q) ur[q] = (q >= p && q < nb && q - p <= ku) ? band[bidx(k0 + p, k0 + q, kl, W)] : 0.0;
  double v = p < nb ? y[k0 + p] : 0.0;
#pragma unroll
  for (int q = BAND_NB - 1; q >= 0; --q) {
    if (p == q && p < nb) v /= ur[q];
    const double vq = __shfl(v, q);  // row k0+q is final here
    if (p < q) v -= ur[q] * vq;
  }
  if (p < nb) y[k0 + p] = v;
}

__global__ void __launch_bounds__(256) k_band_bwd_upd(const double* __restrict__ band, double* __restrict__ y, int kl, int ku,
                                                     int W, int64_t k0, int nb) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t i = std::max<int64_t>(0, k0 - ku) + g / UPD_LANES;
  const int sub = (int)(g % UPD_LANES);
  const bool on = i < k0;
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < BAND_NB / UPD_LANES; ++c) {
    const int p = sub * (BAND_NB / UPD_LANES) + c;
    if (on && p < nb && (k0 + p) - i <= ku) acc += band[bidx(i, k0 + p, kl, W)] * y[k0 + p];
  }
#pragma unroll
  for (int off = UPD_LANES / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, UPD_LANES);
  if (on && sub == 0) y[i] -= acc;
}

// out[new order] = in[old order] (dir = 0) or out[old] = in[new] (dir = 1), dof-wise.
__global__ void k_band_perm(int64_t nn, const int32_t* __restrict__ ip, const double* __restrict__ in,
                            double* __restrict__ out, int dir) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= 3 * nn) return;
  const int64_t i = r / 3;
  const int64_t R = 3 * (int64_t)ip[i] + (r - 3 * i);
  if (dir == 0)
    out[R] = in[r];
  else
    out[r] = in[R];
}

// ---------------------------------------------------------------------------------- host

// Largest |ip[i] - ip[j]| over the node graph.
static int node_bandwidth(int64_t nn, const int32_t* col, const int32_t* nslot, const std::vector<int32_t>& ip) {
  int64_t bw = 0;
  for (int64_t i = 0; i < nn; ++i)
    for (int s = 0; s < nslot[i]; ++s) {
      const int64_t j = col[(size_t)s * nn + i];
      bw = std::max<int64_t>(bw, std::llabs((int64_t)ip[i] - ip[j]));
    }
  return (int)bw;
}

// Breadth-first level structure from `start` over unvisited nodes; returns the last level.
static std::vector<int32_t> bfs_last_level(int64_t nn, const int32_t* col, const int32_t* nslot, int32_t start,
                                           const std::vector<char>& taken, int* depth) {
  std::vector<int32_t> lev(nn, -1);
  std::vector<int32_t> cur{start}, last{start};
  lev[start] = 0;
  int d = 0;
  while (!cur.empty()) {
    last = cur;
    std::vector<int32_t> nxt;
    for (int32_t u : cur)
      for (int s = 0; s < nslot[u]; ++s) {
        const int32_t v = col[(size_t)s * nn + u];
        if (lev[v] < 0 && !taken[v]) {
          lev[v] = d + 1;
          nxt.push_back(v);
        }
      }
    cur.swap(nxt);
    if (!cur.empty()) ++d;
  }
  *depth = d;
  return last;
}

int band_ordering(int64_t nn, const int32_t* col, const int32_t* nslot, int32_t* ip_out, int32_t* bw_out) {
  // candidate 1: the mesh's own order
  std::vector<int32_t> ident(nn);
  for (int64_t i = 0; i < nn; ++i) ident[i] = (int32_t)i;
  const int bw0 = node_bandwidth(nn, col, nslot, ident);
  // candidate 2: reverse Cuthill-McKee from a pseudo-peripheral node of each component
  std::vector<int32_t> deg(nn);
  for (int64_t i = 0; i < nn; ++i) deg[i] = nslot[i];
  std::vector<char> taken(nn, 0);
  std::vector<int32_t> order;
  order.reserve(nn);
  while ((int64_t)order.size() < nn) {
    int32_t start = -1;
    for (int64_t i = 0; i < nn; ++i)
      if (!taken[i] && (start < 0 || deg[i] < deg[start])) start = (int32_t)i;
    int depth = 0, prev = -1;
    for (int it = 0; it < 4 && depth > prev; ++it) {  // pseudo-peripheral node (George-Liu)
      prev = depth;
      std::vector<int32_t> last = bfs_last_level(nn, col, nslot, start, taken, &depth);
      if (depth <= prev) break;
      int32_t best = last[0];
      for (int32_t v : last)
        if (deg[v] < deg[best]) best = v;
      start = best;
    }
    std::deque<int32_t> q{start};
    taken[start] = 1;
    while (!q.empty()) {
      const int32_t u = q.front();
      q.pop_front();
      order.push_back(u);
      std::vector<int32_t> nb;
      for (int s = 0; s < nslot[u]; ++s) {
        const int32_t v = col[(size_t)s * nn + u];
        if (!taken[v]) {
          taken[v] = 1;
          nb.push_back(v);
        }
      }
      std::sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); });
      for (int32_t v : nb) q.push_back(v);
    }
  }
  std::vector<int32_t> ip(nn);
  for (int64_t k = 0; k < nn; ++k) ip[order[nn - 1 - k]] = (int32_t)k;
  const int bw1 = node_bandwidth(nn, col, nslot, ip);
  const std::vector<int32_t>& best = bw1 < bw0 ? ip : ident;
  std::copy(best.begin(), best.end(), ip_out);
  *bw_out = std::min(bw0, bw1);
  return 0;
}

void band_assemble(const BandLU& B, const BellView& M, const double* vals, hipStream_t st) {
  (void)hipMemsetAsync(B.band, 0, (size_t)B.n * B.W * sizeof(double), st);
  hipLaunchKernelGGL(k_band_scatter, dim3((unsigned)((B.n + 255) / 256)), dim3(256), 0, st, M, vals, B.ip, B.band, B.kl, B.W);
}

void band_factor(const BandLU& B, hipStream_t st) {
  (void)hipMemsetAsync(B.flag, 0, sizeof(int32_t), st);
  for (int64_t k0 = 0; k0 < B.n; k0 += BAND_NB) {
    const int nb = (int)std::min<int64_t>(BAND_NB, B.n - k0);
    hipLaunchKernelGGL(k_band_diag, dim3(1), dim3(64), 0, st, B.band, B.kl, B.ku, B.W, k0, nb, B.flag);
    const int64_t rest = B.n - (k0 + nb);
    if (rest <= 0) break;  // only the last panel is narrower than BAND_NB
    // rows k0+nb .. k0+nb-1+kl and columns k0+nb .. k0+nb-1+ku reach the panel
    const int64_t ncol = std::min<int64_t>(rest, B.ku), nrow = std::min<int64_t>(rest, B.kl);
    hipLaunchKernelGGL(k_band_lrow, dim3((unsigned)((nrow + LR_BS - 1) / LR_BS)), dim3(LR_BS), 0, st, B.band, B.n, B.kl, B.ku,
                       B.W, k0);
    hipLaunchKernelGGL(k_band_urow, dim3((unsigned)((ncol + LR_BS - 1) / LR_BS)), dim3(LR_BS), 0, st, B.band, B.n, B.kl, B.ku,
                       B.W, k0);
    const dim3 grid((unsigned)((ncol + BAND_TILE - 1) / BAND_TILE), (unsigned)((nrow + BAND_TILE - 1) / BAND_TILE));
    hipLaunchKernelGGL(k_band_trail, grid, dim3(256), 0, st, B.band, B.n, B.kl, B.ku, B.W, k0, nb);
  }
}

void band_solve(const BandLU& B, const double* b, double* x, hipStream_t st) {
  const int64_t nn = B.n / 3;
  const unsigned g = (unsigned)((B.n + 255) / 256);
  hipLaunchKernelGGL(k_band_perm, dim3(g), dim3(256), 0, st, nn, B.ip, b, B.y, 0);
  for (int64_t k0 = 0; k0 < B.n; k0 += BAND_NB) {
    const int nb = (int)std::min<int64_t>(BAND_NB, B.n - k0);
    hipLaunchKernelGGL(k_band_fwd_diag, dim3(1), dim3(64), 0, st, B.band, B.y, B.kl, B.W, k0, nb);
    const int64_t rows = std::min<int64_t>(B.kl, B.n - (k0 + nb));
    if (rows > 0)
      hipLaunchKernelGGL(k_band_fwd_upd, dim3((unsigned)((rows * UPD_LANES + 255) / 256)), dim3(256), 0, st, B.band, B.y, B.n,
                         B.kl, B.W, k0, nb);
  }
  for (int64_t k0 = ((B.n - 1) / BAND_NB) * BAND_NB; k0 >= 0; k0 -= BAND_NB) {
    const int nb = (int)std::min<int64_t>(BAND_NB, B.n - k0);
    hipLaunchKernelGGL(k_band_bwd_diag, dim3(1), dim3(64), 0, st, B.band, B.y, B.kl, B.ku, B.W, k0, nb);
    const int64_t rows = std::min<int64_t>(B.ku, k0);
    if (rows > 0)
      hipLaunchKernelGGL(k_band_bwd_upd, dim3((unsigned)((rows * UPD_LANES + 255) / 256)), dim3(256), 0, st, B.band, B.y, B.kl,
                         B.ku, B.W, k0, nb);
  }
  hipLaunchKernelGGL(k_band_perm, dim3(g), dim3(256), 0, st, nn, B.ip, B.y, x, 1);
}

}  // namespace cpfem
