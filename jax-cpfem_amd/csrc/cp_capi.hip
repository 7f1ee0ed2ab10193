// C-ABI layer of libcpfem_hip.so: context construction on the host (hex8 geometry, block-ELL
// pattern, element-to-block and node-to-element maps) and the entry points of cpfem.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpfem.h"
#include "cp_internal.hpp"

namespace cpfem {

static hipEvent_t prof_event(Prof* p) {
  if (!p->pool.empty()) {
    hipEvent_t e = p->pool.back();
    p->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Launches of a class between two timed ones: the Krylov classes (tens of thousands of short
// launches per load step) are sampled; every launch of the point and assembly classes (a few
// hundred of milliseconds-long launches, several kernels sharing a class) is timed.
static int prof_every(const Prof* p, int cls) {
  return (cls == PC_SPMV || cls == PC_KVEC || cls == PC_KREPLACE) ? p->sample : 1;
}

int prof_begin(Prof* p, int cls, hipStream_t st) {
  if (!p || !p->on) return -1;
  const int64_t k = p->launches[cls]++;
  if (k % prof_every(p, cls)) return -1;
  if (p->pend.size() >= 8192) prof_flush(p);
  Prof::Pending e{cls, prof_event(p), prof_event(p)};
  (void)hipEventRecord(e.a, st);
  p->pend.push_back(e);
  return (int)p->pend.size() - 1;
}

bool prof_sample(Prof* p, int cls, hipEvent_t* start, hipEvent_t* stop) {
  if (!p || !p->on) return false;
  const int64_t k = p->launches[cls]++;
  if (k % prof_every(p, cls)) return false;
  if (p->pend.size() >= 8192) prof_flush(p);
  Prof::Pending e{cls, prof_event(p), prof_event(p)};
  p->pend.push_back(e);
  *start = e.a;
  *stop = e.b;
  return true;
}

void prof_end(Prof* p, int slot, hipStream_t st) {
  if (!p || slot < 0) return;
  (void)hipEventRecord(p->pend[slot].b, st);
}

int prof_flush(Prof* p) {
  for (auto& e : p->pend) {
    float ms = 0.f;
    if (hipEventSynchronize(e.b) == hipSuccess && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      p->ms[e.cls] += ms;
      p->timed[e.cls] += 1;
    }
    p->pool.push_back(e.a);
    p->pool.push_back(e.b);
  }
  p->pend.clear();
  return 0;
}

}  // namespace cpfem

using namespace cpfem;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                            \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return fail(CPFEM_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
  return CPFEM_OK;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

struct cpfem_ctx {
  int64_t n_nodes = 0, n_cells = 0, n_qp = 0, n_dofs = 0;
  int n_slip = 0, n_slots = 0, uniform = 0;
  SlipTable<12> tab12{};
  SlipTable<24> tab24{};
  std::vector<int32_t> h_nslot, h_col, h_dslot;
  std::vector<double> h_sg, h_jxw;  // full per-cell geometry (host copy)
  // device: mesh and maps
  int32_t* d_cells = nullptr;
  double* d_sg = nullptr;
  double* d_jxw = nullptr;
  int32_t *d_col = nullptr, *d_nslot = nullptr, *d_dslot = nullptr, *d_tslot = nullptr;
  int32_t *d_cptr = nullptr, *d_clist = nullptr;
  int32_t *d_nptr = nullptr, *d_nlist = nullptr;
  uint8_t* d_isbc = nullptr;
  uint8_t* d_nobc = nullptr;  // all-zero flags (raw residual / matrix)
  double* d_ubar = nullptr;
  // device: element work buffers
  double *d_wP = nullptr, *d_wA = nullptr, *d_KE = nullptr, *d_RE = nullptr, *d_sig = nullptr;
  double* d_gi = nullptr;     // 1/g_old table of the Newton point kernel (allocated on first use)
  double* d_sconv = nullptr;  // converged stresses [6][n_qp] for the tangent kernel (Jacobian assemblies)
  int32_t* d_status = nullptr;
  unsigned long long* d_counts = nullptr;  // [0] local Newton iterations, [1] residual evaluations, [2] parked points
  // parking of the assembly point kernel's slow lanes (allocated on first use)
  int32_t *d_park_idx = nullptr, *d_park_it = nullptr;
  double* d_park_s = nullptr;
  uint32_t* d_park_ctr = nullptr;
  int park_lanes = 0;  // 0: no parking
  // reductions
  int nb = 0;
  double* d_partial = nullptr;
  double* d_scalar = nullptr;
  // Krylov
  KrylovScalars* d_ksc = nullptr;
  double* d_kpartial = nullptr;  // [PS_COUNT][krylov blocks]
  double* kv64[12] = {};  // r rhat (spare) q phat (spare) (spare) t invd (spare) rp rq; rp, rq hold 2 n
  float* d_mat32 = nullptr;
  double *d_r64 = nullptr, *d_tmp64 = nullptr;
  // pinned host staging
  KrylovScalars* h_ksc = nullptr;
  double* h_dbl = nullptr;
  int32_t* h_int = nullptr;
  unsigned long long* h_cnt = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // banded direct solver (allocated on first use)
  BandLU band;
  int32_t* d_ip = nullptr;
  int32_t* d_bflag = nullptr;
  Prof prof;
  int spmv_form = -1;  // BiCGSTAB SpMV form: -1 = picked by size (krylov_spmv_form), else forced
  int64_t box[3] = {0, 0, 0};  // box numbering (nx, ny, nz) of the pattern, or zeros
  int box_cols = 1;            // the flat SpMV computes the box numbering's columns (0: loads them)
  double* scratch = nullptr;  // device address of the fixed workspace array on this context's device

  int krylov_form() const { return spmv_form >= 0 ? spmv_form : krylov_spmv_form(n_dofs, n_slots); }

  Geom geom() const {
    Geom G;
    G.cells = d_cells;
    G.sg = d_sg;
    G.jxw = d_jxw;
    G.n_cells = n_cells;
    G.n_qp = n_qp;
    G.uniform = uniform;
    return G;
  }
  BellView bell() const {
    BellView M;
    M.col = d_col;
    M.nslot = d_nslot;
    M.n_nodes = n_nodes;
    M.n_slots = n_slots;
    const bool bx = box_cols && box[0] > 0;
    M.box_b = bx ? (int32_t)box[0] : 0;
    M.box_a = bx ? (int32_t)(box[0] * box[1]) : 0;
    M.box_ny = bx ? (int32_t)box[1] : 0;
    M.box_nz = bx ? (int32_t)box[2] : 0;
    M.box_inv_b = bx ? 1.0 / (double)M.box_b : 0.0;
    M.box_inv_a = bx ? 1.0 / (double)M.box_a : 0.0;
    return M;
  }
  int64_t mat_len() const { return 3LL * n_slots * n_dofs; }
};

// ------------------------------------------------------------------------- hex8 geometry
// Trilinear shape functions on [0,1]^3 with VTK node order and 2x2x2 Gauss points,
// first reference coordinate slowest (the element jax-fem builds through basix).
static const double HEX_NODES[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

static void hex8_ref_grads(double dN[8][8][3]) {
  const double g[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
  int q = 0;
  for (int a0 = 0; a0 < 2; ++a0)
    for (int a1 = 0; a1 < 2; ++a1)
      for (int a2 = 0; a2 < 2; ++a2, ++q) {
        const double xi[3] = {g[a0], g[a1], g[a2]};
        for (int a = 0; a < 8; ++a) {
          double f[3], df[3];
          for (int d = 0; d < 3; ++d) {
            const bool one = HEX_NODES[a][d] == 1.0;
            f[d] = one ? xi[d] : 1.0 - xi[d];
            df[d] = one ? 1.0 : -1.0;
          }
          dN[q][a][0] = df[0] * f[1] * f[2];
          dN[q][a][1] = f[0] * df[1] * f[2];
          dN[q][a][2] = f[0] * f[1] * df[2];
        }
      }
}

static int build_geometry(cpfem_ctx* c, const double* pts, const int32_t* cells) {
  double dN[8][8][3];
  hex8_ref_grads(dN);
  const int64_t nc = c->n_cells;
  c->h_sg.assign((size_t)nc * 8 * 8 * 3, 0.0);
  c->h_jxw.assign((size_t)nc * 8, 0.0);
  for (int64_t e = 0; e < nc; ++e) {
    double X[8][3];
    for (int a = 0; a < 8; ++a) {
      const int64_t nd = cells[e * 8 + a];
      if (nd < 0 || nd >= c->n_nodes) return fail(CPFEM_EBADARG, "cell node index out of range");
      for (int d = 0; d < 3; ++d) X[a][d] = pts[nd * 3 + d];
    }
    for (int q = 0; q < 8; ++q) {
      double J[9] = {0};
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) J[3 * i + j] += X[a][i] * dN[q][a][j];
      const double det = det3(J);
      if (!(det > 0.0)) return fail(CPFEM_EBADARG, "non-positive element Jacobian (check hex8 node order)");
      double Ji[9];
      inv3(J, det, Ji);
      for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
          for (int j = 0; j < 3; ++j) s += dN[q][a][j] * Ji[3 * j + i];
          c->h_sg[((e * 8 + q) * 8 + a) * 3 + i] = s;
        }
      c->h_jxw[e * 8 + q] = det * 0.125;
    }
  }
  // identical cells (structured box meshes) share one table
  double smax = 0.0;
  for (int i = 0; i < 192; ++i) smax = std::max(smax, std::fabs(c->h_sg[i]));
  bool uni = true;
  for (int64_t e = 1; e < nc && uni; ++e) {
    for (int i = 0; i < 192; ++i)
      if (std::fabs(c->h_sg[e * 192 + i] - c->h_sg[i]) > 1e-13 * smax) { uni = false; break; }
    for (int q = 0; q < 8 && uni; ++q)
      if (std::fabs(c->h_jxw[e * 8 + q] - c->h_jxw[q]) > 1e-13 * std::fabs(c->h_jxw[q])) uni = false;
  }
  c->uniform = uni ? 1 : 0;
  // device layout: [24][8] (uniform) or [24][n_qp]
  const int64_t Q = c->n_qp;
  std::vector<double> sgd, jd;
  if (uni) {
    sgd.resize(24 * 8);
    for (int q = 0; q < 8; ++q)
      for (int k = 0; k < 24; ++k) sgd[k * 8 + q] = c->h_sg[q * 24 + k];
    jd.assign(c->h_jxw.begin(), c->h_jxw.begin() + 8);
  } else {
    sgd.resize((size_t)24 * Q);
    for (int64_t qp = 0; qp < Q; ++qp)
      for (int k = 0; k < 24; ++k) sgd[k * Q + qp] = c->h_sg[qp * 24 + k];
    jd = c->h_jxw;
  }
  int rc;
  if ((rc = dalloc(&c->d_sg, sgd.size())) || (rc = dalloc(&c->d_jxw, jd.size()))) return rc;
  HIPCHK(hipMemcpy(c->d_sg, sgd.data(), sgd.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_jxw, jd.data(), jd.size() * sizeof(double), hipMemcpyHostToDevice));
  return CPFEM_OK;
}

// ------------------------------------------------------------------------- pattern / maps
// Box numbering of a block-ELL pattern (host): the strides B = nx and A = nx ny are read off
// the first 27-slot row (its slots 16 and 22 hold the neighbours at +B and +A), then every
// node off the box's faces must have exactly the 27 columns i + dz A + dy B + dx, ascending.
// dims = (nx, ny, nz), or zeros if the pattern is not so numbered.
static void box_numbering(int64_t nn, const int32_t* col, const int32_t* nslot, int64_t dims[3]) {
  dims[0] = dims[1] = dims[2] = 0;
  int ms = 0;
  for (int64_t i = 0; i < nn; ++i) ms = std::max(ms, (int)nslot[i]);
  if (ms != 27) return;
  int64_t i0 = 0;
  while (i0 < nn && nslot[i0] != 27) ++i0;
  if (i0 == nn) return;
  const int64_t B = (int64_t)col[16 * nn + i0] - i0, A = (int64_t)col[22 * nn + i0] - i0;
  if (B < 3 || A < 3 * B || A % B != 0 || nn % A != 0) return;
  const int64_t nx = B, ny = A / B, nz = nn / A;
  if (nz < 3) return;
  for (int64_t iz = 1; iz + 1 < nz; ++iz)
    for (int64_t iy = 1; iy + 1 < ny; ++iy)
      for (int64_t ix = 1; ix + 1 < nx; ++ix) {
        const int64_t i = ix + B * iy + A * iz;
        if (nslot[i] != 27) return;
        for (int s = 0; s < 27; ++s)
          if ((int64_t)col[(int64_t)s * nn + i] != i + (s / 9 - 1) * A + (s / 3 % 3 - 1) * B + (s % 3 - 1)) return;
      }
  dims[0] = nx;
  dims[1] = ny;
  dims[2] = nz;
}

static int build_maps(cpfem_ctx* c, const int32_t* cells) {
  const int64_t nn = c->n_nodes, nc = c->n_cells;
  // node -> (cell, local) list, ascending cell
  std::vector<int32_t> nptr(nn + 1, 0), nlist((size_t)nc * 8);
  for (int64_t e = 0; e < nc; ++e)
    for (int a = 0; a < 8; ++a) nptr[cells[e * 8 + a] + 1]++;
  for (int64_t i = 0; i < nn; ++i) nptr[i + 1] += nptr[i];
  {
    std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t e = 0; e < nc; ++e)
      for (int a = 0; a < 8; ++a) nlist[fill[cells[e * 8 + a]]++] = (int32_t)(e * 8 + a);
  }
  // neighbour sets
  std::vector<std::vector<int32_t>> nb(nn);
  int maxs = 0;
  for (int64_t i = 0; i < nn; ++i) {
    auto& v = nb[i];
    for (int32_t k = nptr[i]; k < nptr[i + 1]; ++k) {
      const int64_t e = nlist[k] / 8;
      for (int b = 0; b < 8; ++b) v.push_back(cells[e * 8 + b]);
    }
    if (v.empty()) v.push_back((int32_t)i);  // isolated node: keep a diagonal block
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    maxs = std::max(maxs, (int)v.size());
  }
  c->n_slots = maxs;
  c->h_nslot.assign(nn, 0);
  c->h_col.assign((size_t)maxs * nn, 0);
  c->h_dslot.assign(nn, 0);
  for (int64_t i = 0; i < nn; ++i) {
    const auto& v = nb[i];
    c->h_nslot[i] = (int32_t)v.size();
    for (int s = 0; s < maxs; ++s) c->h_col[(size_t)s * nn + i] = s < (int)v.size() ? v[s] : (int32_t)i;
    c->h_dslot[i] = (int32_t)(std::lower_bound(v.begin(), v.end(), (int32_t)i) - v.begin());
  }
  box_numbering(nn, c->h_col.data(), c->h_nslot.data(), c->box);
  // transpose map: slot s of row j points at node i; tslot gives the slot of node j in row i
  std::vector<int32_t> tslot((size_t)maxs * nn, 0);
  for (int64_t j = 0; j < nn; ++j)
    for (int s = 0; s < (int)nb[j].size(); ++s) {
      const int32_t i = nb[j][s];
      const auto& v = nb[i];
      tslot[(size_t)s * nn + j] = (int32_t)(std::lower_bound(v.begin(), v.end(), (int32_t)j) - v.begin());
    }
  // (node i, slot s) -> element blocks (cell*64 + a*8 + b), ascending cell
  const int64_t nkey = (int64_t)maxs * nn;
  std::vector<int32_t> cptr(nkey + 1, 0), clist((size_t)nc * 64);
  auto slot_of = [&](int64_t i, int32_t j) {
    const auto& v = nb[i];
    return (int)(std::lower_bound(v.begin(), v.end(), j) - v.begin());
  };
  std::vector<int32_t> keys((size_t)nc * 64);
  for (int64_t e = 0; e < nc; ++e)
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        const int64_t i = cells[e * 8 + a];
        const int s = slot_of(i, cells[e * 8 + b]);
        const int64_t key = (int64_t)s * nn + i;
        keys[e * 64 + a * 8 + b] = (int32_t)key;
        cptr[key + 1]++;
      }
  for (int64_t k = 0; k < nkey; ++k) cptr[k + 1] += cptr[k];
  {
    std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
    for (int64_t t = 0; t < nc * 64; ++t) clist[fill[keys[t]]++] = (int32_t)t;
  }
  int rc;
  if ((rc = dalloc(&c->d_nptr, nptr.size())) || (rc = dalloc(&c->d_nlist, nlist.size())) ||
      (rc = dalloc(&c->d_cptr, cptr.size())) || (rc = dalloc(&c->d_clist, clist.size())) ||
      (rc = dalloc(&c->d_col, c->h_col.size())) || (rc = dalloc(&c->d_nslot, c->h_nslot.size())) ||
      (rc = dalloc(&c->d_dslot, c->h_dslot.size())) || (rc = dalloc(&c->d_tslot, tslot.size())))
    return rc;
  HIPCHK(hipMemcpy(c->d_tslot, tslot.data(), tslot.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_nptr, nptr.data(), nptr.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_nlist, nlist.data(), nlist.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_cptr, cptr.data(), cptr.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_clist, clist.data(), clist.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_col, c->h_col.data(), c->h_col.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_nslot, c->h_nslot.data(), c->h_nslot.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_dslot, c->h_dslot.data(), c->h_dslot.size() * 4, hipMemcpyHostToDevice));
  return CPFEM_OK;
}

// ------------------------------------------------------------------------- helpers
static int check_state(const cpfem_ctx* c, const CpfemState* s, const CpfemParams* p) {
  if (!s || !s->Fp_inv || !s->g || !s->slip || !s->rot) return fail(CPFEM_EBADARG, "state pointers must be set");
  if (!p) return fail(CPFEM_EBADARG, "params must be set");
  for (int k = 0; k < CPFEM_NPARAM; ++k) {
    if (!p->ptr[k]) return fail(CPFEM_EBADARG, "material parameter pointer " + std::to_string(k) + " is null");
    if (p->stride[k] != 0 && p->stride[k] != 1) return fail(CPFEM_EBADARG, "parameter stride must be 0 or 1");
  }
  if (!(p->tol > 0.0) || p->max_sub_step < 1 || p->max_iter < 1) return fail(CPFEM_EBADARG, "bad local solver settings");
  (void)c;
  return CPFEM_OK;
}

static StateView state_view(const CpfemState* s, int64_t ld) {
  StateView v;
  v.Fpi = s->Fp_inv;
  v.g = s->g;
  v.slip = s->slip;
  v.rot = s->rot;
  v.ld = ld;
  return v;
}

static MatView mat_view(const CpfemParams* p) {
  MatView m;
  for (int k = 0; k < MP_COUNT; ++k) {
    m.ptr[k] = p->ptr[k];
    m.stride[k] = p->stride[k];
  }
  m.ao = p->gamma0;
  m.dt = p->dt;
  return m;
}

static SolveCfg solve_cfg(const CpfemParams* p) {
  SolveCfg s;
  s.tol = p->tol;
  s.max_sub_step = p->max_sub_step;
  s.max_iter = p->max_iter;
  return s;
}

static int ensure_matrix_buffers(cpfem_ctx* c) {
  int rc;
  if (!c->d_wA && (rc = dalloc(&c->d_wA, (size_t)81 * c->n_qp))) return rc;
  if (!c->d_sconv && (rc = dalloc(&c->d_sconv, (size_t)6 * c->n_qp))) return rc;
  if (!c->d_KE && (rc = dalloc(&c->d_KE, (size_t)576 * c->n_cells))) return rc;
  return CPFEM_OK;
}

// 1/g_old table of the Newton point kernel ([n_slip][n_qp]), allocated on first use
static int ensure_point_scratch(cpfem_ctx* c) {
  int rc;
  if (!c->d_gi && (rc = dalloc(&c->d_gi, (size_t)c->n_slip * c->n_qp))) return rc;
  return CPFEM_OK;
}

static int ensure_park(cpfem_ctx* c) {
  int rc;
  if (c->d_park_ctr) return CPFEM_OK;
  if ((rc = dalloc(&c->d_park_idx, (size_t)c->n_qp)) || (rc = dalloc(&c->d_park_it, (size_t)c->n_qp)) ||
      (rc = dalloc(&c->d_park_s, (size_t)6 * c->n_qp)) || (rc = dalloc(&c->d_park_ctr, 2)))
    return rc;
  return CPFEM_OK;
}

static int ensure_krylov(cpfem_ctx* c, bool mixed) {
  int rc;
  const int64_t n = c->n_dofs;
  if (!c->d_ksc) {
    if ((rc = dalloc(&c->d_ksc, 1)) || (rc = dalloc(&c->d_kpartial, (size_t)8 * krylov_stride_max(n)))) return rc;
    for (int i = 0; i < 12; ++i)
      if ((rc = dalloc(&c->kv64[i], i >= 10 ? 2 * n : n))) return rc;
    if ((rc = dalloc(&c->d_r64, n)) || (rc = dalloc(&c->d_tmp64, n))) return rc;
    HIPCHK(hipMemset(c->d_ksc, 0, sizeof(KrylovScalars)));
  }
  if (mixed && !c->d_mat32 && (rc = dalloc(&c->d_mat32, c->mat_len()))) return rc;
  return CPFEM_OK;
}

// |a . b| finished on the device, copied to the host (synchronising)
static int dot_host(cpfem_ctx* c, hipStream_t st, const double* a, const double* b, double* out) {
  launch_residual_norm(c->n_dofs, a, b, c->d_partial, c->nb, c->d_scalar, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_dbl, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *out = c->h_dbl[0];
  return CPFEM_OK;
}

static KVec kvec(double** v, double* x) {
  KVec k;
  k.x = x;
  k.r = v[0];
  k.rhat = v[1];

  k.q = v[3];
  k.phat = v[4];
  k.t = v[7];
  k.invd = v[8];
  k.rp = reinterpret_cast<double2*>(v[10]);
  k.rq = reinterpret_cast<double2*>(v[11]);
  return k;
}

// Fixed workspace of cpfem_solve_noalloc: a file-scope device array (each device holds its own
// copy until the process ends).  Host code never passes the array's name to a kernel: its
// device address is looked up once per device (scratch_address) and used from then on.
// Layout: Krylov scalars, partial sums [PS_SLOTS][NOALLOC_PARTIALS], then 11 n doubles of
// vectors for n <= CPFEM_NOALLOC_MAX_DOFS.
constexpr int NOALLOC_SCALARS = 32;  // doubles reserved for KrylovScalars (256 B)
constexpr int PS_SLOTS = 8;          // partial-sum slots (PS_COUNT)
constexpr int NOALLOC_PARTIALS = 2048;
constexpr size_t NOALLOC_DOUBLES = (size_t)NOALLOC_SCALARS + (size_t)PS_SLOTS * NOALLOC_PARTIALS + 11ull * CPFEM_NOALLOC_MAX_DOFS;
static_assert(sizeof(KrylovScalars) <= NOALLOC_SCALARS * sizeof(double), "Krylov scalars exceed their slot");
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

static std::mutex g_scratch_mu;
static double* g_scratch_addr[64] = {};

// Device address of `scratch` on the current device, looked up on the first call per device.
static int scratch_address(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(CPFEM_EHIP, "device index out of range for the fixed workspace");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CPFEM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_addr[dev];
  return CPFEM_OK;
}

// Krylov workspace of one solve: the context's own buffers (cpfem_solve) or the library's
// fixed device array (cpfem_solve_noalloc).
struct KBufs {
  KrylovScalars* sc;
  double* partial;  // [PS_COUNT][krylov_stride_max(n)]
  double* v[12];    // as cpfem_ctx::kv64
  double* tmp64;    // n
  double* r64;      // n (mixed form only)
  float* mat32;     // FP32 copy of the matrix values (mixed form only)
  Prof* prof;       // kernel-class timing, or nullptr
};

static KBufs ctx_bufs(cpfem_ctx* c) {
  KBufs kb{};
  kb.sc = c->d_ksc;
  kb.partial = c->d_kpartial;
  for (int i = 0; i < 12; ++i) kb.v[i] = c->kv64[i];
  kb.tmp64 = c->d_tmp64;
  kb.r64 = c->d_r64;
  kb.mat32 = c->d_mat32;
  kb.prof = &c->prof;
  return kb;
}

// One BiCGSTAB run (JAX semantics) with the Jacobi vector already in kb.v[8]; returns the
// iteration counter (negative: breakdown code).
// With vals64 != nullptr (mixed form) a replacement of the recurrence residual by the true
// FP64 residual is offered after every batch (reliable updates, k_kreplace).
template <typename TM>
static int run_bicgstab(cpfem_ctx* c, hipStream_t st, const KBufs& kb, const TM* vals, const double* b, double* x, double tol,
                        double atol, int maxiter, int32_t* k_out, const double* vals64 = nullptr) {
  const BellView M = c->bell();
  KrylovScalars h{};
  h.tol = tol;
  h.atol = atol;
  h.maxiter = maxiter;
  *c->h_ksc = h;
  HIPCHK(hipMemcpyAsync(kb.sc, c->h_ksc, sizeof(KrylovScalars), hipMemcpyHostToDevice, st));
  KrylovWork w;
  w.n = c->n_dofs;
  w.nb = krylov_blocks(c->n_dofs);
  w.sf = c->krylov_form();
  w.nbs = krylov_spmv_blocks(c->n_dofs, w.sf);
  w.ps = std::max(w.nb, w.nbs);
  w.single = (w.sf >= 3 && std::is_same<TM, double>::value) ? 1 : 0;
  w.partial = kb.partial;
  w.sc = kb.sc;
  KVec v = kvec(const_cast<double**>(kb.v), x);
  int sl = prof_begin(kb.prof, PC_SPMV, st);
  krylov_init<TM>(M, vals, b, v, w, st);
  prof_end(kb.prof, sl, st);
  HIPCHK(hipGetLastError());
  // Batches of iterations; the done flag of batch j is read while batch j+1 runs.
  const int CH = 16;
  int issued = 0, batch = 0;
  bool done = false;
  while (!done) {
    for (int i = 0; i < CH; ++i) krylov_iteration<TM>(M, vals, v, w, i & 1, st, kb.prof);
    if (vals64) {
      const int sr = prof_begin(kb.prof, PC_KREPLACE, st);
      krylov_replace(M, vals64, b, v, w, batch, st);
      prof_end(kb.prof, sr, st);
    }
    issued += CH;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&c->h_int[2 * (batch & 1)], &kb.sc->k, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev[batch & 1], st));
    if (batch > 0) {
      const int pb = (batch - 1) & 1;
      HIPCHK(hipEventSynchronize(c->ev[pb]));
      if (c->h_int[2 * pb + 1] != 0 || issued > maxiter + 2 * CH) done = true;
    }
    ++batch;
    if (issued > maxiter + 4 * CH) break;
  }
  HIPCHK(hipMemcpyAsync(c->h_ksc, kb.sc, sizeof(KrylovScalars), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *k_out = c->h_ksc->k;
  return CPFEM_OK;
}

// Jacobi-BiCGSTAB of cpfem_solve on the workspace kb (arguments already checked).
static int solve_impl(cpfem_ctx* c, hipStream_t st, const KBufs& kb, const double* mat, const double* b, double* x, double tol,
                      double atol, int32_t maxiter, int32_t precond, int32_t precision, CpfemSolveInfo* info) {
  CpfemSolveInfo inf{};
  double bb;
  int rc;
  if ((rc = dot_host(c, st, b, b, &bb))) return rc;
  inf.b_norm = std::sqrt(bb);
  launch_inv_diag(c->bell(), mat, c->d_dslot, kb.v[8], precond, st);
  if (precision == CPFEM_FP64) {
    int32_t k = 0;
    if ((rc = run_bicgstab<double>(c, st, kb, mat, b, x, tol, atol, maxiter, &k))) return rc;
    inf.iterations = k;
    inf.outer = 1;
  } else {
    // One BiCGSTAB whose SpMVs read an FP32 copy of the matrix (vectors and scalars FP64),
    // with the recurrence residual replaced by the FP64 residual b - A x whenever it has
    // fallen 100x since the last replacement (offered after every batch); a restart from the
    // current x follows only if the true residual is still above max(tol |b|, atol) when the
    // iteration stops.
    const BellView M = c->bell();
    launch_cast_f32(c->mat_len(), mat, kb.mat32, st);
    const double thr2 = std::max(tol * tol * bb, atol * atol);
    int total = 0;
    int outer = 0;
    for (; outer < 8; ++outer) {
      int32_t k = 0;
      if ((rc = run_bicgstab<float>(c, st, kb, kb.mat32, b, x, tol, atol, maxiter - total, &k, mat))) return rc;
      total += (k > 0 ? k : 0);
      launch_residual64(M, mat, x, b, kb.r64, st);
      double rr;
      if ((rc = dot_host(c, st, kb.r64, kb.r64, &rr))) return rc;
      if (!(rr > thr2) || total >= maxiter || k <= 0) { ++outer; break; }
    }
    inf.iterations = total;
    inf.outer = outer;
  }
  // |A x - b| < 0.1 check of jax_solve (solver.py:43-46)
  launch_residual64(c->bell(), mat, x, b, kb.tmp64, st);
  double err2;
  if ((rc = dot_host(c, st, kb.tmp64, kb.tmp64, &err2))) return rc;
  inf.residual = std::sqrt(err2);
  if (info) *info = inf;
  if (!(inf.residual < 0.1)) {
    launch_fill(c->n_dofs, std::nan(""), x, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    return fail(CPFEM_EKRYLOV, "JAX linear solver failed to converge with err = " + std::to_string(inf.residual));
  }
  return CPFEM_OK;
}

// ------------------------------------------------------------------------- C ABI
extern "C" {

const char* cpfem_last_error(void) { return g_err.c_str(); }
int cpfem_version(void) { return 1; }

int cpfem_create(const CpfemMesh* mesh, const CpfemSlip* slip, cpfem_ctx** out) {
  if (!mesh || !slip || !out || !mesh->points || !mesh->cells) return fail(CPFEM_EBADARG, "null argument");
  if (mesh->n_nodes <= 0 || mesh->n_cells <= 0) return fail(CPFEM_EBADARG, "empty mesh");
  if (mesh->n_nodes > (1LL << 30) || mesh->n_cells * 64 > (1LL << 31) - 1) return fail(CPFEM_EBADARG, "mesh too large for int32 maps");
  if (slip->n_slip != 12 && slip->n_slip != 24) return fail(CPFEM_EBADARG, "n_slip must be 12 or 24");
  if (!slip->directions || !slip->normals) return fail(CPFEM_EBADARG, "slip table pointers are null");
  cpfem_ctx* c = new cpfem_ctx();
  c->n_nodes = mesh->n_nodes;
  c->n_cells = mesh->n_cells;
  c->n_qp = 8 * mesh->n_cells;
  c->n_dofs = 3 * mesh->n_nodes;
  c->n_slip = slip->n_slip;
  for (int a = 0; a < slip->n_slip; ++a)
    for (int d = 0; d < 3; ++d) {
      if (slip->n_slip == 12) {
        c->tab12.d[a][d] = slip->directions[3 * a + d];
        c->tab12.n[a][d] = slip->normals[3 * a + d];
      } else {
        c->tab24.d[a][d] = slip->directions[3 * a + d];
        c->tab24.n[a][d] = slip->normals[3 * a + d];
      }
    }
  int rc;
  if ((rc = dalloc(&c->d_cells, (size_t)8 * c->n_cells))) { cpfem_destroy(c); return rc; }
  if (hipMemcpy(c->d_cells, mesh->cells, (size_t)32 * c->n_cells, hipMemcpyHostToDevice) != hipSuccess) {
    cpfem_destroy(c);
    return fail(CPFEM_EHIP, "hipMemcpy cells");
  }
  if ((rc = build_geometry(c, mesh->points, mesh->cells)) || (rc = build_maps(c, mesh->cells))) {
    const std::string e = g_err;
    cpfem_destroy(c);
    g_err = e;
    return rc;
  }
  if ((rc = scratch_address(&c->scratch))) {
    const std::string e = g_err;
    cpfem_destroy(c);
    g_err = e;
    return rc;
  }
  c->nb = (int)((c->n_dofs + 255) / 256);
  const int nbq = (int)((c->n_qp + 255) / 256);
  c->nb = std::max(c->nb, 1);
  if ((rc = dalloc(&c->d_isbc, c->n_dofs)) || (rc = dalloc(&c->d_nobc, c->n_dofs)) || (rc = dalloc(&c->d_ubar, c->n_dofs)) ||
      (rc = dalloc(&c->d_wP, (size_t)9 * c->n_qp)) || (rc = dalloc(&c->d_RE, (size_t)24 * c->n_cells)) ||
      (rc = dalloc(&c->d_sig, (size_t)9 * c->n_qp)) || (rc = dalloc(&c->d_status, 1)) || (rc = dalloc(&c->d_counts, 3)) ||
      (rc = dalloc(&c->d_partial, (size_t)4 * std::max(c->nb, nbq))) || (rc = dalloc(&c->d_scalar, 4))) {
    const std::string e = g_err;
    cpfem_destroy(c);
    g_err = e;
    return rc;
  }
  if (hipMemset(c->d_isbc, 0, c->n_dofs) != hipSuccess || hipMemset(c->d_nobc, 0, c->n_dofs) != hipSuccess || hipMemset(c->d_ubar, 0, 8 * c->n_dofs) != hipSuccess ||
      hipMemset(c->d_status, 0, 4) != hipSuccess ||
      hipMemset(c->d_counts, 0, 24) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ksc, sizeof(KrylovScalars)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_dbl, 8 * sizeof(double)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_int, 8 * sizeof(int32_t)) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, 3 * sizeof(unsigned long long)) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming) != hipSuccess) {
    cpfem_destroy(c);
    return fail(CPFEM_EHIP, "context allocation failed");
  }
  *out = c;
  return CPFEM_OK;
}

void cpfem_destroy(cpfem_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  dfree(c->band.band); dfree(c->band.y); dfree(c->d_ip); dfree(c->d_bflag);
  dfree(c->d_cells); dfree(c->d_sg); dfree(c->d_jxw);
  dfree(c->d_col); dfree(c->d_nslot); dfree(c->d_dslot); dfree(c->d_tslot);
  dfree(c->d_cptr); dfree(c->d_clist); dfree(c->d_nptr); dfree(c->d_nlist);
  dfree(c->d_isbc); dfree(c->d_nobc); dfree(c->d_ubar);
  dfree(c->d_gi); dfree(c->d_sconv);
  dfree(c->d_wP); dfree(c->d_wA); dfree(c->d_KE); dfree(c->d_RE); dfree(c->d_sig); dfree(c->d_status); dfree(c->d_counts);
  dfree(c->d_park_idx); dfree(c->d_park_it); dfree(c->d_park_s); dfree(c->d_park_ctr);
  dfree(c->d_partial); dfree(c->d_scalar);
  dfree(c->d_ksc);
  dfree(c->d_kpartial);
  for (int i = 0; i < 12; ++i) dfree(c->kv64[i]);
  dfree(c->d_mat32); dfree(c->d_r64); dfree(c->d_tmp64);
  if (c->h_ksc) (void)hipHostFree(c->h_ksc);
  if (c->h_dbl) (void)hipHostFree(c->h_dbl);
  if (c->h_int) (void)hipHostFree(c->h_int);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  for (int i = 0; i < 2; ++i)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  prof_flush(&c->prof);
  for (auto e : c->prof.pool) (void)hipEventDestroy(e);
  delete c;
}

int cpfem_info(const cpfem_ctx* c, int64_t* info) {
  if (!c || !info) return fail(CPFEM_EBADARG, "null argument");
  info[0] = c->n_nodes;
  info[1] = c->n_cells;
  info[2] = c->n_qp;
  info[3] = c->n_dofs;
  info[4] = c->n_slots;
  info[5] = c->uniform;
  info[6] = c->mat_len();
  info[7] = c->n_slip;
  return CPFEM_OK;
}

int cpfem_pattern(const cpfem_ctx* c, int32_t* nslot, int32_t* col) {
  if (!c || !nslot || !col) return fail(CPFEM_EBADARG, "null argument");
  std::memcpy(nslot, c->h_nslot.data(), c->h_nslot.size() * 4);
  std::memcpy(col, c->h_col.data(), c->h_col.size() * 4);
  return CPFEM_OK;
}

int cpfem_geometry(const cpfem_ctx* c, double* sg, double* jxw) {
  if (!c || !sg || !jxw) return fail(CPFEM_EBADARG, "null argument");
  std::memcpy(sg, c->h_sg.data(), c->h_sg.size() * 8);
  std::memcpy(jxw, c->h_jxw.data(), c->h_jxw.size() * 8);
  return CPFEM_OK;
}

int cpfem_set_dirichlet(cpfem_ctx* c, void* stream, int64_t n, const int64_t* dofs, const double* vals) {
  if (!c || (n > 0 && (!dofs || !vals))) return fail(CPFEM_EBADARG, "null argument");
  std::vector<uint8_t> isbc(c->n_dofs, 0);
  std::vector<double> ubar(c->n_dofs, 0.0);
  for (int64_t k = 0; k < n; ++k) {
    if (dofs[k] < 0 || dofs[k] >= c->n_dofs) return fail(CPFEM_EBADARG, "Dirichlet dof out of range");
    isbc[dofs[k]] = 1;
    ubar[dofs[k]] = vals[k];
  }
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(c->d_isbc, isbc.data(), c->n_dofs, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_ubar, ubar.data(), 8 * c->n_dofs, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return CPFEM_OK;
}

int cpfem_local(cpfem_ctx* c, void* stream, int64_t n_qp, const double* u_grad, const double* dofs, const CpfemState* state,
                const CpfemParams* params, double* P, double* A, double* sigma, const CpfemStateOut* out, int32_t* iters) {
  if (!c) return fail(CPFEM_EBADARG, "null context");
  int rc;
  if ((rc = check_state(c, state, params))) return rc;
  if (!u_grad) {
    if (!dofs) return fail(CPFEM_EBADARG, "either u_grad or dofs is required");
    if (n_qp != c->n_qp) return fail(CPFEM_EBADARG, "n_qp must equal 8 * n_cells when u_grad comes from dofs");
  }
  if (n_qp <= 0) return CPFEM_OK;
  StateOut so{nullptr, nullptr, nullptr};
  if (out) {
    if (!out->Fp_inv || !out->g || !out->slip) return fail(CPFEM_EBADARG, "state_out pointers must all be set");
    so.Fpi = out->Fp_inv;
    so.g = out->g;
    so.slip = out->slip;
  }
  hipStream_t st = (hipStream_t)stream;
  const Geom G = c->geom();
  const StateView sv = state_view(state, n_qp);
  const MatView mv = mat_view(params);
  const SolveCfg cfg = solve_cfg(params);
  const int sl = prof_begin(&c->prof, PC_LOCAL, st);
  if (c->n_slip == 12)
    launch_local<12>(n_qp, u_grad, G, dofs, sv, mv, c->tab12, cfg, params->self_block, P, A, sigma, so, iters, c->d_status, st);
  else
    launch_local<24>(n_qp, u_grad, G, dofs, sv, mv, c->tab24, cfg, params->self_block, P, A, sigma, so, iters, c->d_status, st);
  prof_end(&c->prof, sl, st);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_assemble(cpfem_ctx* c, void* stream, const double* dofs, const CpfemState* state, const CpfemParams* params,
                   double* res, double* mat, int32_t* iters, int32_t flags) {
  if (!c || !dofs || !res) return fail(CPFEM_EBADARG, "null argument");
  int rc;
  if ((rc = check_state(c, state, params))) return rc;
  if (mat && (rc = ensure_matrix_buffers(c))) return rc;
  if ((rc = ensure_point_scratch(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const Geom G = c->geom();
  const StateView sv = state_view(state, c->n_qp);
  const MatView mv = mat_view(params);
  const SolveCfg cfg = solve_cfg(params);
  double* wA = mat ? c->d_wA : nullptr;
  double* sout = mat ? c->d_sconv : nullptr;  // converged stresses for the tangent kernel
  Prof* pf = &c->prof;
  PointPark pk{c->d_park_idx, c->d_park_s, c->d_park_it, c->d_park_ctr, c->n_qp, c->park_lanes};
  if (pk.lanes > 0) {
    if ((rc = ensure_park(c))) return rc;
    pk = PointPark{c->d_park_idx, c->d_park_s, c->d_park_it, c->d_park_ctr, c->n_qp, c->park_lanes};
    launch_zero_words(pk.ctr, 2, st);
  }
  int sl = prof_begin(pf, PC_ELEMENT_QP, st);
  if (c->n_slip == 12)
    launch_element_qp<12>(G, dofs, sv, mv, c->tab12, cfg, c->d_wP, sout, iters, c->d_status, c->d_counts, pk, c->d_gi, st);
  else
    launch_element_qp<24>(G, dofs, sv, mv, c->tab24, cfg, c->d_wP, sout, iters, c->d_status, c->d_counts, pk, c->d_gi, st);
  prof_end(pf, sl, st);
  if (pk.lanes > 0) {
    sl = prof_begin(pf, PC_ELEMENT_PARK, st);
    if (c->n_slip == 12)
      launch_element_qp_park<12>(G, dofs, sv, mv, c->tab12, cfg, c->d_wP, sout, iters, c->d_status, c->d_counts, pk, st);
    else
      launch_element_qp_park<24>(G, dofs, sv, mv, c->tab24, cfg, c->d_wP, sout, iters, c->d_status, c->d_counts, pk, st);
    prof_end(pf, sl, st);
  }
  if (mat) {
    sl = prof_begin(pf, PC_TANGENT, st);
    if (c->n_slip == 12)
      launch_point_tangent<12>(G, dofs, sv, mv, c->tab12, sout, wA, st);
    else
      launch_point_tangent<24>(G, dofs, sv, mv, c->tab24, sout, wA, st);
    prof_end(pf, sl, st);
  }
  sl = prof_begin(pf, PC_ELEMENT_K, st);
  launch_element_K(G, c->d_wP, wA, c->d_KE, c->d_RE, st);
  prof_end(pf, sl, st);
  const uint8_t* isbc = (flags & CPFEM_APPLY_BC) ? c->d_isbc : c->d_nobc;
  sl = prof_begin(pf, PC_GATHER_RES, st);
  launch_gather_residual(c->n_nodes, c->d_RE, c->d_nptr, c->d_nlist, isbc, c->d_ubar, dofs, res, st);
  prof_end(pf, sl, st);
  if (mat) {
    sl = prof_begin(pf, PC_GATHER_MAT, st);
    launch_gather_matrix(c->bell(), c->d_KE, c->d_cptr, c->d_clist, isbc, mat, st);
    prof_end(pf, sl, st);
  }
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_cell_stress(cpfem_ctx* c, void* stream, const double* dofs, const CpfemState* state, const CpfemParams* params,
                      double* sigma_cells) {
  if (!c || !dofs || !sigma_cells) return fail(CPFEM_EBADARG, "null argument");
  int rc;
  if ((rc = cpfem_local(c, stream, c->n_qp, nullptr, dofs, state, params, nullptr, nullptr, c->d_sig, nullptr, nullptr)))
    return rc;
  launch_cell_average(c->geom(), c->d_sig, sigma_cells, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_update_state(cpfem_ctx* c, void* stream, const double* dofs, const CpfemState* state, const CpfemParams* params,
                       const CpfemStateOut* out) {
  if (!c || !dofs || !out) return fail(CPFEM_EBADARG, "null argument");
  int rc;
  if ((rc = check_state(c, state, params))) return rc;
  if (out->Fp_inv == state->Fp_inv || out->g == state->g || out->slip == state->slip)
    return fail(CPFEM_EBADARG, "state_out must not alias state");
  return cpfem_local(c, stream, c->n_qp, nullptr, dofs, state, params, nullptr, nullptr, nullptr, out, nullptr);
}

int cpfem_band_order(int64_t n_nodes, const int32_t* col, const int32_t* nslot, int32_t* ip, int32_t* bandwidth) {
  if (n_nodes <= 0 || !col || !nslot || !ip || !bandwidth) return fail(CPFEM_EBADARG, "null argument");
  return band_ordering(n_nodes, col, nslot, ip, bandwidth);
}

int cpfem_box_numbering(int64_t n_nodes, const int32_t* col, const int32_t* nslot, int64_t* dims) {
  if (n_nodes <= 0 || !col || !nslot || !dims) return fail(CPFEM_EBADARG, "null argument");
  box_numbering(n_nodes, col, nslot, dims);
  return CPFEM_OK;
}

// Band storage is capped (BAND_MAX_BYTES); larger systems are refused with a message that
// names the Krylov solver.
static constexpr double BAND_MAX_BYTES = 64.0 * (1ull << 30);

static int ensure_band(cpfem_ctx* c) {
  if (c->band.band) return CPFEM_OK;
  std::vector<int32_t> ip(c->n_nodes);
  int32_t bw = 0;
  band_ordering(c->n_nodes, c->h_col.data(), c->h_nslot.data(), ip.data(), &bw);
  BandLU B;
  B.n = c->n_dofs;
  B.kl = B.ku = 3 * bw + 2;
  B.W = B.kl + B.ku + 1;
  const double bytes = (double)B.n * B.W * sizeof(double);
  if (bytes > BAND_MAX_BYTES)
    return fail(CPFEM_EBADARG, "direct solve: band storage of " + std::to_string((long long)(bytes / (1 << 20))) +
                                   " MiB exceeds the limit; use the Krylov solver ('jax_solver')");
  int rc;
  if ((rc = dalloc(&c->d_ip, (size_t)c->n_nodes)) || (rc = dalloc(&c->d_bflag, 1)) || (rc = dalloc(&B.y, (size_t)B.n)) ||
      (rc = dalloc(&B.band, (size_t)B.n * B.W)))
    return rc;
  HIPCHK(hipMemcpy(c->d_ip, ip.data(), ip.size() * 4, hipMemcpyHostToDevice));
  B.ip = c->d_ip;
  B.flag = c->d_bflag;
  c->band = B;
  return CPFEM_OK;
}

int cpfem_solve_direct(cpfem_ctx* c, void* stream, const double* mat, const double* b, double* x, int32_t refine,
                       CpfemSolveInfo* info) {
  if (!c || !mat || !b || !x) return fail(CPFEM_EBADARG, "null argument");
  if (refine < 0 || refine > 10) return fail(CPFEM_EBADARG, "refine must be in 0..10");
  int rc;
  if ((rc = ensure_krylov(c, false)) || (rc = ensure_band(c))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const BandLU& B = c->band;
  CpfemSolveInfo inf{};
  double bb;
  if ((rc = dot_host(c, st, b, b, &bb))) return rc;
  inf.b_norm = std::sqrt(bb);
  band_assemble(B, c->bell(), mat, st);
  band_factor(B, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->h_int, c->d_bflag, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (c->h_int[0] != 0) return fail(CPFEM_EDIRECT, "direct solve: zero or non-finite pivot in the unpivoted band LU");
  band_solve(B, b, x, st);
  // x += A^-1 (b - A x), with the residual in FP64 on the block-ELL matrix
  for (int k = 0; k < refine; ++k) {
    launch_residual64(c->bell(), mat, x, b, c->d_r64, st);
    band_solve(B, c->d_r64, c->d_tmp64, st);
    launch_axpby(c->n_dofs, 1.0, c->d_tmp64, 1.0, x, st);
  }
  HIPCHK(hipGetLastError());
  launch_residual64(c->bell(), mat, x, b, c->d_tmp64, st);
  double err2;
  if ((rc = dot_host(c, st, c->d_tmp64, c->d_tmp64, &err2))) return rc;
  inf.residual = std::sqrt(err2);
  inf.iterations = refine;
  inf.outer = 1;
  if (info) *info = inf;
  if (!std::isfinite(inf.residual)) return fail(CPFEM_EDIRECT, "direct solve: non-finite solution");
  return CPFEM_OK;
}

int cpfem_spmv(cpfem_ctx* c, void* stream, const double* mat, const double* x, double* y) {
  if (!c || !mat || !x || !y) return fail(CPFEM_EBADARG, "null argument");
  launch_spmv(c->bell(), mat, x, y, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_norm(cpfem_ctx* c, void* stream, const double* v, double* out, int32_t* point_status) {
  if (!c || !v || !out) return fail(CPFEM_EBADARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  // status word: read and cleared in stream order, behind every point kernel queued so far
  HIPCHK(hipMemcpyAsync(&c->h_int[4], c->d_status, 4, hipMemcpyDeviceToHost, st));
  launch_zero_words(reinterpret_cast<uint32_t*>(c->d_status), 1, st);
  double s;
  int rc;
  if ((rc = dot_host(c, st, v, v, &s))) return rc;  // synchronises st
  *out = std::sqrt(s);
  const int32_t bits = c->h_int[4];
  if (point_status) *point_status = bits;
  const std::string flags = std::string((bits & PT_CAPPED) ? " [local Newton iteration cap reached]" : "") +
                            ((bits & PT_NONFINITE) ? " [non-finite local state]" : "");
  if (!std::isfinite(*out)) return fail(CPFEM_ENONFINITE, "non-finite residual norm" + flags);
  if (bits & PT_CAPPED)
    return fail(CPFEM_ELOCAL, "some quadrature points hit the local Newton iteration cap" + flags);
  if (bits & PT_NONFINITE) return fail(CPFEM_ENONFINITE, "non-finite local state at some quadrature points" + flags);
  return CPFEM_OK;
}

int cpfem_set_point_park(cpfem_ctx* c, int32_t lanes) {
  if (!c) return fail(CPFEM_EBADARG, "null context");
  if (lanes < 0 || lanes > 63) return fail(CPFEM_EBADARG, "park lanes must be in 0..63");
  c->park_lanes = lanes;
  return CPFEM_OK;
}

int cpfem_krylov_config(const cpfem_ctx* c, int32_t* out) {
  if (!c || !out) return fail(CPFEM_EBADARG, "null argument");
  out[0] = c->krylov_form();
  out[1] = krylov_spmv_blocks(c->n_dofs, out[0]);
  out[2] = krylov_blocks(c->n_dofs);
  return CPFEM_OK;
}

int cpfem_set_spmv_form(cpfem_ctx* c, int32_t form) {
  if (!c) return fail(CPFEM_EBADARG, "null context");
  if (form < -1 || form > 4) return fail(CPFEM_EBADARG, "SpMV form must be -1 (by size), 0, 1, 2, 3 or 4");
  if (form >= 0 && !krylov_form_allowed(c->n_dofs, form, c->n_slots))
    return fail(CPFEM_EBADARG, form >= 3 ? "the one-row-per-thread SpMV forms need a 27-slot (hex8 box-like) mesh"
                                         : "the all-slots SpMV form needs fewer than 2^28 rows");
  c->spmv_form = form;
  return CPFEM_OK;
}

int cpfem_set_box_columns(cpfem_ctx* c, int32_t enable) {
  if (!c) return fail(CPFEM_EBADARG, "null context");
  if (enable != 0 && enable != 1) return fail(CPFEM_EBADARG, "enable must be 0 or 1");
  c->box_cols = enable;
  return CPFEM_OK;
}

int cpfem_point_counts(cpfem_ctx* c, void* stream, int32_t reset, int64_t* out) {
  if (!c || !out) return fail(CPFEM_EBADARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(c->h_cnt, c->d_counts, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  if (reset) launch_zero_words(reinterpret_cast<uint32_t*>(c->d_counts), 6, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  out[0] = (int64_t)c->h_cnt[0];
  out[1] = (int64_t)c->h_cnt[1];
  out[2] = (int64_t)c->h_cnt[2];
  return CPFEM_OK;
}

int cpfem_axpby(cpfem_ctx* c, void* stream, double alpha, const double* x, double beta, double* y) {
  if (!c || !x || !y) return fail(CPFEM_EBADARG, "null argument");
  launch_axpby(c->n_dofs, alpha, x, beta, y, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_solve(cpfem_ctx* c, void* stream, const double* mat, const double* b, double* x, double tol, double atol,
                int32_t maxiter, int32_t precond, int32_t precision, CpfemSolveInfo* info) {
  if (!c || !mat || !b || !x) return fail(CPFEM_EBADARG, "null argument");
  if (precision != CPFEM_FP64 && precision != CPFEM_MIXED) return fail(CPFEM_EBADARG, "unknown precision");
  if (maxiter < 0) return fail(CPFEM_EBADARG, "maxiter must be >= 0");
  int rc;
  if ((rc = ensure_krylov(c, precision == CPFEM_MIXED))) return rc;
  return solve_impl(c, (hipStream_t)stream, ctx_bufs(c), mat, b, x, tol, atol, maxiter, precond, precision, info);
}

int cpfem_solve_noalloc(cpfem_ctx* c, void* stream, const double* mat, const double* b, double* x, double tol, double atol,
                        int32_t maxiter, int32_t precond, CpfemSolveInfo* info) {
  if (!c || !mat || !b || !x) return fail(CPFEM_EBADARG, "null argument");
  if (maxiter < 0) return fail(CPFEM_EBADARG, "maxiter must be >= 0");
  if (c->n_dofs > CPFEM_NOALLOC_MAX_DOFS)
    return fail(CPFEM_EBADARG, "cpfem_solve_noalloc: " + std::to_string((long long)c->n_dofs) +
                                   " dofs exceed its fixed workspace of " + std::to_string(CPFEM_NOALLOC_MAX_DOFS) +
                                   " dofs; use cpfem_solve");
  if (krylov_stride_max(c->n_dofs) > NOALLOC_PARTIALS)
    return fail(CPFEM_EBADARG, "cpfem_solve_noalloc: the Krylov grid exceeds the fixed workspace's partial sums");
  if (!c->scratch) return fail(CPFEM_EHIP, "cpfem_solve_noalloc: the fixed workspace was not resolved on this device");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = c->n_dofs;
  KBufs kb{};
  double* p = c->scratch;
  kb.sc = reinterpret_cast<KrylovScalars*>(p);
  p += NOALLOC_SCALARS;
  kb.partial = p;
  p += (size_t)PS_SLOTS * NOALLOC_PARTIALS;
  double* vec0 = p;
  // the vectors this solve uses (kv64 slots 2, 5, 6 and 9 are spare); rp and rq hold 2 n each
  for (int i : {0, 1, 3, 4, 7, 8}) { kb.v[i] = p; p += n; }
  kb.v[10] = p; p += 2 * n;
  kb.v[11] = p; p += 2 * n;
  kb.tmp64 = p; p += n;
  // the array keeps the previous call's contents: the vectors start from zeros
  launch_zero_words(reinterpret_cast<uint32_t*>(vec0), 2 * (int64_t)(p - vec0), st);
  kb.prof = nullptr;
  return solve_impl(c, st, kb, mat, b, x, tol, atol, maxiter, precond, CPFEM_FP64, info);
}

int cpfem_to_csr(cpfem_ctx* c, void* stream, const double* mat, int64_t* indptr, int32_t* indices, double* data) {
  if (!c || !mat || !indptr || !indices || !data) return fail(CPFEM_EBADARG, "null argument");
  std::vector<double> h(c->mat_len());
  HIPCHK(hipMemcpyAsync(h.data(), mat, h.size() * 8, hipMemcpyDeviceToHost, (hipStream_t)stream));
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  const int64_t nr = c->n_dofs, nn = c->n_nodes;
  int64_t nz = 0;
  indptr[0] = 0;
  for (int64_t r = 0; r < nr; ++r) {
    const int64_t i = r / 3;
    for (int s = 0; s < c->h_nslot[i]; ++s) {
      const int64_t j = c->h_col[(size_t)s * nn + i];
      for (int k = 0; k < 3; ++k) {
        indices[nz] = (int32_t)(3 * j + k);
        data[nz] = h[(size_t)(3 * s + k) * nr + r];
        ++nz;
      }
    }
    indptr[r + 1] = nz;
  }
  return CPFEM_OK;
}

int cpfem_local_vjp(cpfem_ctx* c, void* stream, const double* dofs, const CpfemState* state, const CpfemParams* params,
                    const CpfemVjp* vjp) {
  if (!c || !dofs || !vjp) return fail(CPFEM_EBADARG, "null argument");
  int rc;
  if ((rc = check_state(c, state, params))) return rc;
  VjpArgs va;
  va.Pbar = vjp->Pbar;
  va.sbar_cell = vjp->sbar_cell;
  va.Fbar = vjp->Fbar;
  va.gbar = vjp->gbar;
  va.slipbar = vjp->slipbar;
  va.Hbar = vjp->Hbar;
  va.Fpibar = vjp->Fpibar;
  va.gbar_old = vjp->gbar_old;
  va.pbar = vjp->pbar;
  va.Rbar = vjp->Rbar;
  va.mask = vjp->mask;
  va.accumulate = vjp->accumulate;
  hipStream_t st = (hipStream_t)stream;
  const Geom G = c->geom();
  const StateView sv = state_view(state, c->n_qp);
  const MatView mv = mat_view(params);
  const SolveCfg cfg = solve_cfg(params);
  if (c->n_slip == 12)
    launch_local_vjp<12>(G, dofs, sv, mv, c->tab12, cfg, params->self_block, va, st);
  else
    launch_local_vjp<24>(G, dofs, sv, mv, c->tab24, cfg, params->self_block, va, st);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_ugrad_transpose(cpfem_ctx* c, void* stream, const double* Hbar, double* out) {
  if (!c || !Hbar || !out) return fail(CPFEM_EBADARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  launch_element_K(c->geom(), Hbar, nullptr, c->d_KE, c->d_RE, st);
  launch_gather_residual(c->n_nodes, c->d_RE, c->d_nptr, c->d_nlist, c->d_nobc, c->d_ubar, Hbar, out, st);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_grad_field(cpfem_ctx* c, void* stream, const double* lam, double scale, double* out) {
  if (!c || !lam || !out) return fail(CPFEM_EBADARG, "null argument");
  launch_grad_field(c->geom(), lam, c->d_isbc, scale, out, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_transpose(cpfem_ctx* c, void* stream, const double* mat, double* matT) {
  if (!c || !mat || !matT || mat == matT) return fail(CPFEM_EBADARG, "bad transpose arguments");
  launch_transpose(c->bell(), c->d_tslot, mat, matT, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_profile(cpfem_ctx* c, int32_t enable, int32_t sample) {
  if (!c || sample < 1) return fail(CPFEM_EBADARG, "bad profile arguments");
  prof_flush(&c->prof);
  for (int k = 0; k < PC_COUNT; ++k) {
    c->prof.launches[k] = 0;
    c->prof.timed[k] = 0;
    c->prof.ms[k] = 0.0;
  }
  c->prof.on = enable != 0;
  c->prof.sample = sample;
  return CPFEM_OK;
}

int cpfem_profile_read(cpfem_ctx* c, int32_t n, double* ms, int64_t* launches, int64_t* timed) {
  if (!c || !ms || !launches || !timed || n < 1) return fail(CPFEM_EBADARG, "null argument");
  prof_flush(&c->prof);
  for (int k = 0; k < n && k < PC_COUNT; ++k) {
    ms[k] = c->prof.ms[k];
    launches[k] = c->prof.launches[k];
    timed[k] = c->prof.timed[k];
  }
  return CPFEM_OK;
}

int cpfem_copy(cpfem_ctx* c, void* stream, const double* src, double* dst, int64_t n) {
  if (!c || !src || !dst || n < 0) return fail(CPFEM_EBADARG, "bad copy arguments");
  hipStream_t st = (hipStream_t)stream;
  const int sl = prof_begin(&c->prof, PC_COPY, st);
  launch_copy(n, src, dst, st);
  prof_end(&c->prof, sl, st);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

int cpfem_bc_x0(cpfem_ctx* c, void* stream, const double* dofs, double* x0) {
  if (!c || !dofs || !x0) return fail(CPFEM_EBADARG, "null argument");
  launch_bc_x0(c->n_dofs, c->d_isbc, c->d_ubar, dofs, x0, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return CPFEM_OK;
}

}  // extern "C"
