// Internal types shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "cp_point.hpp"

namespace cpfem {

enum { MP_XM = 0, MP_GSS_A, MP_H, MP_TSAT, MP_R, MP_C11, MP_C12, MP_C44, MP_COUNT };

// Mesh geometry on the device
struct Geom {
  const int32_t* cells;  // (n_cells, 8)
  const double* sg;      // uniform ? [24][8] : [24][n_qp]  (dN_a/dX_d, comp = 3a+d)
  const double* jxw;     // uniform ? [8] : [n_qp]
  int64_t n_cells;
  int64_t n_qp;
  int uniform;
};

// Per-point history, SoA with leading dimension ld (= n_qp)
struct StateView {
  const double* Fpi;
  const double* g;
  const double* slip;
  const double* rot;
  int64_t ld;
};
struct StateOut {
  double* Fpi;
  double* g;
  double* slip;
};

// Per-point material parameters; value of point q is ptr[k][q * stride[k]]
struct MatView {
  const double* ptr[MP_COUNT];
  int64_t stride[MP_COUNT];
  double ao;
  double dt;
};

struct SolveCfg {
  double tol;
  int max_sub_step;
  int max_iter;
};

// Block-ELL matrix pattern (3x3 node blocks, values [(3 s + k)][3 i + c])
struct BellView {
  const int32_t* col;    // [n_slots][n_nodes]
  const int32_t* nslot;  // [n_nodes]
  int64_t n_nodes;
  int n_slots;
  // Box numbering (box_b > 0): the nodes of a box of box_b x (box_a / box_b) x box_nz nodes are
  // numbered i = ix + box_b iy + box_a iz, and every node off the box's faces has the 27
  // columns i + dz box_a + dy box_b + dx (dx, dy, dz in {-1, 0, 1}), in that (ascending)
  // order, which the one-row-per-thread SpMV computes instead of loading.
  int32_t box_b, box_a, box_ny, box_nz;
  double box_inv_b, box_inv_a;  // 1 / box_b, 1 / box_a
};

// Parking of the assembly point kernel's slow lanes.  The lanes of a wave iterate their local
// Newton in lockstep; once at most `lanes` of them are still iterating, those park their
// point (index, stress iterate, iteration count) in this queue and the wave moves on.  A
// second kernel then finishes the parked points densely, one queue entry per lane.
// lanes = 0: no parking.
struct PointPark {
  int32_t* idx;   // [cap] parked point index
  double* s;      // [6][cap] its stress iterate (Voigt)
  int32_t* it;    // [cap] iterations done
  uint32_t* ctr;  // [0] parked points, [1] unused
  int64_t cap;
  int lanes;
};

// The assembly's point kernels: launch_element_qp runs the local Newton and writes JxW P and,
// if sout ([6][n_qp]) is set, the converged stress, keeping the 1/g_old table in gi
// ([N][n_qp] device scratch); launch_point_tangent forms the JxW-weighted consistent tangent
// wA ([81][n_qp]) from that stress.
template <int N>
void launch_element_qp(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                       const SolveCfg& cfg, double* wP, double* sout, int32_t* iters, int32_t* status,
                       unsigned long long* counts, const PointPark& park, double* gi, hipStream_t st);
// The second kernel over the parked points (grid sized to the device, chunks taken atomically).
template <int N>
void launch_element_qp_park(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                            const SolveCfg& cfg, double* wP, double* sout, int32_t* iters, int32_t* status,
                            unsigned long long* counts, const PointPark& park, hipStream_t st);
template <int N>
void launch_point_tangent(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                          const double* sin, double* wA, hipStream_t st);
template <int N>
void launch_local(int64_t n_qp, const double* Hd, const Geom& G, const double* u, const StateView& sv, const MatView& mv,
                  const SlipTable<N>& tab, const SolveCfg& cfg, int self_block, double* P, double* A, double* sig,
                  const StateOut& so, int32_t* iters, int32_t* status, hipStream_t st);
void launch_element_K(const Geom& G, const double* wP, const double* wA, double* KE, double* RE, hipStream_t st);
void launch_gather_matrix(const BellView& M, const double* KE, const int32_t* cptr, const int32_t* clist, const uint8_t* is_bc,
                          double* vals, hipStream_t st);
void launch_gather_residual(int64_t n_nodes, const double* RE, const int32_t* nptr, const int32_t* nlist, const uint8_t* is_bc,
                            const double* ubar, const double* u, double* res, hipStream_t st);
void launch_cell_average(const Geom& G, const double* sig, double* out, hipStream_t st);
void launch_bc_x0(int64_t n, const uint8_t* is_bc, const double* ubar, const double* u, double* x0, hipStream_t st);
// Zero n 32-bit words with a kernel (stream-ordered and safe inside captured graphs).
void launch_zero_words(uint32_t* p, int64_t n, hipStream_t st);

// ---------------------------------------------------------------- adjoint
struct VjpArgs {
  const double* Pbar;       // [9][Q] or null
  const double* sbar_cell;  // [9][n_cells] or null (cotangent of the cell-averaged Cauchy stress)
  const double* Fbar;       // [9][Q] or null (cotangent of Fp_inv_new)
  const double* gbar;       // [N][Q] or null (cotangent of g_new)
  const double* slipbar;    // [N][Q] or null (cotangent of slip_new)
  double* Hbar;             // [9][Q]
  double* Fpibar;           // [9][Q]
  double* gbar_old;         // [N][Q]
  double* pbar;             // [8][Q] (xm, gss_a, h, t_sat, r, C11, C12, C44)
  double* Rbar;             // [9][Q]
  int mask;
  int accumulate;
};

template <int N>
void launch_local_vjp(const Geom& G, const double* u, const StateView& sv, const MatView& mv, const SlipTable<N>& tab,
                      const SolveCfg& cfg, int self_block, const VjpArgs& va, hipStream_t st);
void launch_grad_field(const Geom& G, const double* lam, const uint8_t* is_bc, double scale_by_w, double* out, hipStream_t st);
void launch_transpose(const BellView& M, const int32_t* tslot, const double* A, double* At, hipStream_t st);

// ---------------------------------------------------------------- profiling
// Optional per-kernel-class timing with HIP events recorded on the launch stream around every
// `sample`-th launch of the Krylov classes and every launch of the others; all are counted.
enum ProfClass {
  PC_ELEMENT_QP = 0,  // constitutive update (local Newton, P) of the assembly (k_element_res)
  PC_ELEMENT_K,       // element stiffness / residual (k_element_K, k_element_R)
  PC_GATHER_MAT,      // block-ELL gather (k_gather_matrix)
  PC_GATHER_RES,      // residual gather (k_gather_residual)
  PC_SPMV,            // Krylov kernels containing the SpMV (k_kq, k_kt, k_kinit)
  PC_KVEC,            // Krylov vector kernels (k_kp, k_ks, k_kxr)
  PC_LOCAL,           // point kernel for stress / history (k_local)
  PC_COPY,            // plain copy (bandwidth calibration)
  PC_KREPLACE,        // residual replacement of the mixed form (k_kreplace)
  PC_ELEMENT_PARK,    // parked points of the assembly's point kernel (k_element_qp_park)
  PC_TANGENT,         // consistent tangent from the converged stresses (k_point_tangent)
  PC_COUNT
};

struct Prof {
  bool on = false;
  int sample = 1;
  int64_t launches[PC_COUNT] = {};
  int64_t timed[PC_COUNT] = {};
  double ms[PC_COUNT] = {};
  struct Pending { int cls; hipEvent_t a, b; };
  std::vector<Pending> pend;   // recorded, not yet read
  std::vector<hipEvent_t> pool;  // free events
};

// Returns the event pair slot to record around the launch, or -1 (not sampled).
int prof_begin(Prof* p, int cls, hipStream_t st);
void prof_end(Prof* p, int slot, hipStream_t st);
// For launches through hipExtLaunchKernelGGL, whose start/stop events time the kernel itself
// (without the dispatch gap that stream-recorded events include): true and the event pair
// if this launch of class cls is sampled.
bool prof_sample(Prof* p, int cls, hipEvent_t* start, hipEvent_t* stop);
int prof_flush(Prof* p);  // synchronises pending events into ms[]

// ---------------------------------------------------------------- Krylov
// Scalars of one BiCGSTAB run, resident in device memory.  Per-iteration values live in
// two parity slots (iteration j writes [j & 1], reads [(j + 1) & 1]).
struct KrylovScalars {
  double rho2[2], alpha2[2], omega2[2];
  double rs;     // <r, r> at the last convergence test
  double rs_ref2[2];  // <r, r> at the last replacement of the recurrence residual (mixed form),
                      // parity slots over the replacement offers (offer j writes [j & 1])
  double bs;     // <b, b>
  double atol2;
  double tol, atol;
  int32_t k2[2];
  int32_t k;     // iteration counter at the last test (negative: breakdown code)
  int32_t done;  // read by the host together with k
  int32_t maxiter;
  int32_t pad;
};

struct KrylovWork {
  int64_t n;
  int nb;            // blocks of the vector kernels (one row per thread, grid-stride)
  int nbs;           // blocks of the SpMV kernels (one or two 64-row groups x 4 slot parts per block)
  int sf;            // SpMV form (krylov_spmv_form): 4 = one row per thread over XCD-local row
                     // ranges, 3 = one row per thread, 2 = 128-row tiles, 1 = 64-row tiles,
                     // 0 = 64-row tiles with every slot's loads in flight at once
  int ps;            // partial-sum stride: max(nb, nbs); unused entries hold zeros
  int single;        // one-operand recurrence (FP64 flat forms): phat formed by the x/r update
  double* partial;   // [slot][ps]
  KrylovScalars* sc;
};

// BiCGSTAB vectors.  Besides JAX's x, r, rhat, q, phat = M p, t the iteration keeps
// rh = M r, qh = M q and pp = M (p - omega q), so that the next phat = rh + beta pp and
// shat = rh - alpha qh can be formed inside the SpMV kernels; s = r - alpha q and shat are
// not stored (the x/r update forms them again from r, q and the (rh, qh) pair).  They are
// stored as interleaved pairs rp = (rh, pp) and rq = (rh, qh), so that each SpMV gathers its
// two operand vectors with one 16-byte load per column entry.  rh is written into rp (x/r
// update, init, residual replacement); the first SpMV kernel writes rq whole, with its own
// row's rh, when it forms qh.
struct KVec {
  double *x, *r, *rhat, *q, *phat, *t;
  double2 *rp, *rq;
  const double* invd;
};

// Banded LU of the Newton matrix (direct-solve option): n dofs, half bandwidths kl = ku,
// row-major band of width W = kl + ku + 1 under the node map ip (old -> new), work vector y,
// pivot flag.
struct BandLU {
  int64_t n = 0;
  int kl = 0, ku = 0, W = 0;
  double* band = nullptr;
  const int32_t* ip = nullptr;
  double* y = nullptr;
  int32_t* flag = nullptr;
};
int band_ordering(int64_t n_nodes, const int32_t* col, const int32_t* nslot, int32_t* ip, int32_t* bandwidth);
void band_assemble(const BandLU& B, const BellView& M, const double* vals, hipStream_t st);
void band_factor(const BandLU& B, hipStream_t st);
void band_solve(const BandLU& B, const double* b, double* x, hipStream_t st);

template <typename TM>
void krylov_init(const BellView& M, const TM* vals, const double* b, KVec v, KrylovWork w, hipStream_t st);
template <typename TM>
void krylov_iteration(const BellView& M, const TM* vals, KVec v, KrylovWork w, int par, hipStream_t st, Prof* prof);
int krylov_blocks(int64_t n);
int krylov_spmv_blocks(int64_t n, int form);
int krylov_spmv_form(int64_t n, int n_slots);
bool krylov_form_allowed(int64_t n, int form, int n_slots);
int krylov_stride_max(int64_t n);
void krylov_replace(const BellView& M, const double* vals64, const double* b, KVec v, KrylovWork w, int offer, hipStream_t st);
void launch_spmv(const BellView& M, const double* vals, const double* x, double* y, hipStream_t st);
void launch_inv_diag(const BellView& M, const double* vals, const int32_t* dslot, double* invd, int precond, hipStream_t st);

void launch_residual_norm(int64_t n, const double* a, const double* b, double* partial, int nb, double* out,
                          hipStream_t st);
void launch_axpby(int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t st);
void launch_cast_f32(int64_t n, const double* x, float* y, hipStream_t st);
void launch_residual64(const BellView& M, const double* vals, const double* x, const double* b, double* r, hipStream_t st);
void launch_copy(int64_t n, const double* x, double* y, hipStream_t st);
// y[i] = v for i < n
void launch_fill(int64_t n, double v, double* y, hipStream_t st);

}  // namespace cpfem
