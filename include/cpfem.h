/*
 * cpfem.h — C ABI of the MI355X (gfx950) crystal-plasticity hot path.
 *
 * One shared library, libcpfem_hip.so.  Plain pointers and sizes only; every array
 * argument documented "device" is a HIP device pointer (for instance a torch tensor's
 * data_ptr()), "host" arrays are ordinary CPU memory.  Every entry point returns 0 on
 * success or a CPFEM_E* code; cpfem_last_error() gives the message.  No exception crosses
 * the ABI.  A context is not re-entrant: use one context per stream.
 *
 * Layouts
 *   quadrature point  q = 8 * cell + local_q (2x2x2 Gauss points, first coordinate slowest)
 *   per-point arrays  structure-of-arrays [component][n_qp]  (e.g. Fp_inv[9][n_qp], row-major 3x3)
 *   dof vector        3 * node + component   (the reference's flattened sol_list)
 *   matrix            block-ELL of 3x3 node blocks: value (s, k) of scalar row r = 3*i + c is
 *                     mat[(3*s + k) * n_dofs + r], column 3*col[s][i] + k, s < nslot[i]
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   cpfem_local          tensor_map / update_int_vars_map, singlecrystal_copper/models_copper.py:263-269
 *                        (local Newton :204-249, IFT tangent :251-259, P :155-162, update :164-169)
 *   cpfem_assemble       problem.newton_update + apply_bc_vec + get_A
 *                        (crystal_plasticity_OR_design/solver.py:387-400, :119-133, :279-304)
 *   cpfem_cell_stress    compute_avg_stress, models_copper.py:297-319
 *   cpfem_update_state   update_int_vars_gp, models_copper.py:273-282
 *   cpfem_solve          jax_solve (Jacobi-BiCGSTAB, tol = atol = 1e-10), solver.py:19-48
 *   cpfem_set_dirichlet  fe.update_Dirichlet_boundary_conditions + node/vec/vals lists
 *                        (singlecrystal_copper.py:186; solver.py:119-153)
 */
#ifndef CPFEM_ABI_H_
#define CPFEM_ABI_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CPFEM_OK = 0,
  CPFEM_EBADARG = 1,    /* invalid argument */
  CPFEM_EHIP = 2,       /* HIP runtime error */
  CPFEM_ELOCAL = 3,     /* some point hit the local iteration cap */
  CPFEM_EKRYLOV = 4,    /* |A x - b| >= 0.1 after the Krylov solve (solver.py:45) */
  CPFEM_ENONFINITE = 5, /* NaN / Inf in a residual */
  CPFEM_ENOMEM = 6,
  CPFEM_EDIRECT = 7     /* zero / non-finite pivot or non-finite result in the direct solve */
};

enum { CPFEM_FP64 = 0, CPFEM_MIXED = 1 };

typedef struct cpfem_ctx cpfem_ctx;

typedef struct CpfemMesh {
  int64_t n_nodes;
  int64_t n_cells;
  const double* points; /* host (n_nodes, 3) */
  const int32_t* cells; /* host (n_cells, 8), VTK hexahedron node order */
} CpfemMesh;

typedef struct CpfemSlip {
  int32_t n_slip;           /* 12 or 24 */
  const double* directions; /* host (n_slip, 3), normalised slip directions d */
  const double* normals;    /* host (n_slip, 3), normalised plane normals n */
} CpfemSlip;

typedef struct CpfemState { /* device, SoA, leading dimension n_qp */
  const double* Fp_inv;     /* [9][n_qp]      inverse plastic deformation gradient */
  const double* g;          /* [n_slip][n_qp] slip resistance */
  const double* slip;       /* [n_slip][n_qp] accumulated slip */
  const double* rot;        /* [9][n_qp]      crystal rotation R (lab = R crystal) */
} CpfemState;

typedef struct CpfemStateOut { /* device, same layout as CpfemState */
  double* Fp_inv;
  double* g;
  double* slip;
} CpfemStateOut;

enum { CPFEM_XM = 0, CPFEM_GSS_A, CPFEM_H, CPFEM_TSAT, CPFEM_R, CPFEM_C11, CPFEM_C12, CPFEM_C44, CPFEM_NPARAM };

typedef struct CpfemParams {
  const double* ptr[CPFEM_NPARAM]; /* device; value of point q is ptr[k][q * stride[k]] */
  int64_t stride[CPFEM_NPARAM];    /* 1 = one value per point, 0 = one value for all */
  double gamma0;                   /* reference slip rate, `ao` = 1e-3 (models_copper.py:147) */
  double dt;                       /* load-step time increment */
  double tol;                      /* local Newton tolerance on |r| (1e-8, models_copper.py:212) */
  int32_t max_sub_step;            /* backtracking trials (5, models_copper.py:231) */
  int32_t max_iter;                /* local iteration cap; the reference has none */
  int32_t self_block;              /* 1: q = 1 in blocks of 3 (models_copper.py:71-76);
                                      0: q = r everywhere (calibration/case1.py:175) */
} CpfemParams;

typedef struct CpfemSolveInfo {
  int32_t iterations; /* BiCGSTAB counter (negative = breakdown code -10/-11) */
  int32_t outer;      /* refinement sweeps (mixed precision) */
  double residual;    /* |A x - b| after the solve */
  double b_norm;
} CpfemSolveInfo;

const char* cpfem_last_error(void);
int cpfem_version(void);

/* Build the context: shape gradients, JxW, block-ELL pattern, element-to-block maps. */
int cpfem_create(const CpfemMesh* mesh, const CpfemSlip* slip, cpfem_ctx** out);
void cpfem_destroy(cpfem_ctx* ctx);
/* info[0..7] = n_nodes, n_cells, n_qp, n_dofs, n_slots, uniform_cells, matrix_len, n_slip */
int cpfem_info(const cpfem_ctx* ctx, int64_t* info);
/* host copies of the pattern: nslot (n_nodes), col (n_slots * n_nodes) */
int cpfem_pattern(const cpfem_ctx* ctx, int32_t* nslot, int32_t* col);
/* shape gradients (n_cells, 8, 8, 3) and JxW (n_cells, 8), host copies */
int cpfem_geometry(const cpfem_ctx* ctx, double* shape_grads, double* jxw);

/* Dirichlet dofs (host, unique) and values; later entries of the reference lists must
   already have overwritten earlier ones on shared dofs. */
int cpfem_set_dirichlet(cpfem_ctx* ctx, void* stream, int64_t n, const int64_t* dofs, const double* vals);

/* x0 of the incremental solve (solver.py:221-227): ubar - dofs on Dirichlet dofs, 0 elsewhere. */
int cpfem_bc_x0(cpfem_ctx* ctx, void* stream, const double* dofs, double* x0);

/* Point-level maps.  u_grad [9][n_qp] (device) or, if u_grad is NULL, the gradient of the
   nodal dofs (device, n_dofs) with n_qp = 8 * n_cells.  Outputs may be NULL:
   P [9][n_qp], A = dP/du_grad [81][n_qp] ((ij),(kl) row-major), sigma (Cauchy) [9][n_qp],
   out (new history), iters [n_qp]. */
int cpfem_local(cpfem_ctx* ctx, void* stream, int64_t n_qp, const double* u_grad, const double* dofs,
                const CpfemState* state, const CpfemParams* params, double* P, double* A, double* sigma,
                const CpfemStateOut* out, int32_t* iters);

enum { CPFEM_APPLY_BC = 1 };

/* Residual (n_dofs) and, if mat != NULL, the block-ELL Jacobian.  With CPFEM_APPLY_BC in
   flags the Dirichlet rows follow apply_bc_vec (res = u - ubar) and the row elimination of
   get_A (zero row, unit diagonal); without it both are the raw assembled quantities.
   iters (n_qp) optional. */
int cpfem_assemble(cpfem_ctx* ctx, void* stream, const double* dofs, const CpfemState* state,
                   const CpfemParams* params, double* res, double* mat, int32_t* iters, int32_t flags);

/* Cell-averaged Cauchy stress (n_cells, 3, 3), device. */
int cpfem_cell_stress(cpfem_ctx* ctx, void* stream, const double* dofs, const CpfemState* state,
                      const CpfemParams* params, double* sigma_cells);

/* History update at the converged dofs (out must not alias state). */
int cpfem_update_state(cpfem_ctx* ctx, void* stream, const double* dofs, const CpfemState* state,
                       const CpfemParams* params, const CpfemStateOut* out);

/* Jacobi-BiCGSTAB on the block-ELL matrix.  x holds x0 on entry.  precision CPFEM_FP64:
   the reference iteration in FP64; CPFEM_MIXED: FP32 inner BiCGSTAB with FP64 iterative
   refinement to the same stopping threshold.  Returns CPFEM_EKRYLOV (x set to NaN) when
   |A x - b| >= 0.1 as the reference asserts. */
int cpfem_solve(cpfem_ctx* ctx, void* stream, const double* mat, const double* b, double* x, double tol,
                double atol, int32_t maxiter, int32_t precond, int32_t precision, CpfemSolveInfo* info);

/* cpfem_solve in FP64 without any device allocation.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  The Krylov workspace (scalars, partial sums,
   11 n_dofs doubles of vectors) lives in a fixed device array of the library, about 17 MiB,
   which each device that the library runs on holds until the process ends; the array's
   device address is looked up when a context is created on that device.  Limits: n_dofs <=
   CPFEM_NOALLOC_MAX_DOFS (larger systems return CPFEM_EBADARG and must use cpfem_solve);
   FP64 only.  All calls on one device share the array: calls that overlap on one device
   (other streams, other threads, other contexts) race and must be serialised by the caller.
   Arguments, results and errors are those of cpfem_solve with precision CPFEM_FP64; the
   context must already exist (cpfem_create allocates). */
#define CPFEM_NOALLOC_MAX_DOFS 196608
int cpfem_solve_noalloc(cpfem_ctx* ctx, void* stream, const double* mat, const double* b, double* x, double tol,
                        double atol, int32_t maxiter, int32_t precond, CpfemSolveInfo* info);

/* Direct solve of A x = b (replaces umfpack_solve, crystal_plasticity_OR_design/solver.py:50-60,
   selected by solver_options 'umfpack_solver' at :100-101): LU of the band form of A under a
   bandwidth-reducing node order (unpivoted; a zero pivot returns CPFEM_EDIRECT), then `refine`
   (0..10) steps of FP64 iterative refinement.  x need not hold a guess.  The band storage,
   n_dofs x (4 * (3 * node bandwidth + 2) + 1) doubles under 64 GiB, is allocated on first use
   and kept with the context.  info->residual = |A x - b|, info->iterations = refine. */
int cpfem_solve_direct(cpfem_ctx* ctx, void* stream, const double* mat, const double* b, double* x, int32_t refine,
                       CpfemSolveInfo* info);

/* Node order of the direct solve, computed on the host (no device needed): the mesh's own
   order or reverse Cuthill-McKee, whichever has the smaller node bandwidth.  col/nslot: the
   block-ELL pattern ([n_slots][n_nodes] node columns, slot counts, as cpfem_pattern returns);
   ip[old node] = new node; *bandwidth = max |ip[i] - ip[j]| over coupled nodes. */
int cpfem_band_order(int64_t n_nodes, const int32_t* col, const int32_t* nslot, int32_t* ip, int32_t* bandwidth);

/* Box numbering of a block-ELL pattern, detected on the host (no device needed; col/nslot as
   cpfem_pattern returns them): dims = (nx, ny, nz) nodes if the nodes of an nx x ny x nz box
   are numbered i = ix + nx iy + nx ny iz and every node off the box's faces has exactly the
   27 columns i + dz nx ny + dy nx + dx (dx, dy, dz in {-1, 0, 1}); (0, 0, 0) otherwise.  A
   context whose pattern has it lets the one-row-per-thread SpMV compute those nodes' columns
   instead of loading them.  No counterpart in the reference (a property of the mesh's node
   numbering; the reference's box meshes have it). */
int cpfem_box_numbering(int64_t n_nodes, const int32_t* col, const int32_t* nslot, int64_t* dims);

/* y = A x on the block-ELL matrix (device). */
int cpfem_spmv(cpfem_ctx* ctx, void* stream, const double* mat, const double* x, double* y);

/* |v|_2 of a device vector of n_dofs entries, returned on the host (synchronises);
   *point_status (optional) receives the OR of point status bits (1: a point hit the local
   iteration cap, 2: a point's local state became non-finite) set by the point kernels queued
   on `stream` since the last call, which reads and clears them in stream order.  Returns
   CPFEM_ENONFINITE if the norm is NaN / Inf or a point's state became non-finite, else
   CPFEM_ELOCAL if a point hit the iteration cap (*out and *point_status are set in both
   cases).  The reference has no such exits: its residual check asserts on NaN
   (crystal_plasticity_OR_design/solver.py:416-417) and its local Newton has no cap
   (models_copper.py:247), so a point that does not converge never returns there. */
int cpfem_norm(cpfem_ctx* ctx, void* stream, const double* v, double* out, int32_t* point_status);

/* Launch configuration of the Jacobi-BiCGSTAB on this context (no counterpart in the
   reference: diagnostics of the device solver): out[0] = SpMV form (4: one row per thread
   over XCD-local row ranges, 3: one row per thread, 2: 128-row tiles, 1: 64-row tiles, 0:
   64-row tiles with every slot's loads in flight), out[1] = SpMV blocks, out[2] =
   vector-kernel blocks. */
int cpfem_krylov_config(const cpfem_ctx* ctx, int32_t* out);
/* Forces the SpMV form of this context's BiCGSTAB (-1: picked by system size, the default).
   For tests and measurements that run two forms on one system; form 0 is refused
   (CPFEM_EBADARG) on systems of 2^28 rows or more, forms 3 and 4 on meshes whose rows have
   other than 27 node slots.  No counterpart in the reference. */
int cpfem_set_spmv_form(cpfem_ctx* ctx, int32_t form);
/* Computed node columns of the one-row-per-thread SpMV on a box-numbered pattern (see
   cpfem_box_numbering): 1 = on (the default), 0 = every column loaded.  The products are
   bitwise the same either way; for tests and measurements.  No counterpart in the
   reference. */
int cpfem_set_box_columns(cpfem_ctx* ctx, int32_t enable);

/* Work counters of the assembly's point kernel, summed over its launches since the last
   reset: out[0] = local Newton iterations, out[1] = residual evaluations, out[2] = parked
   points (synchronises; reset != 0 clears them in stream order).  Diagnostics with no
   counterpart in the reference: the measured inputs of the point kernel's algorithmic flop
   count. */
int cpfem_point_counts(cpfem_ctx* ctx, void* stream, int32_t reset, int64_t* out);

/* Parking of the assembly point kernel's slow lanes (0..63, default 0 = none): the lanes of a
   wave iterate their local Newton in lockstep, and once at most `lanes` of them are still
   iterating those park their point in a queue that a second kernel finishes, one point per lane.
   Every output is bitwise the same for any setting; only the time changes.  No counterpart
   in the reference (whose vmapped while_loop runs every batch to its slowest point,
   models_copper.py:247). */
int cpfem_set_point_park(cpfem_ctx* ctx, int32_t lanes);

/* y = alpha x + beta y on n_dofs entries (device). */
int cpfem_axpby(cpfem_ctx* ctx, void* stream, double alpha, const double* x, double beta, double* y);

/* CSR export of the block-ELL matrix (host outputs: indptr n_dofs+1, indices/data nnz). */
int cpfem_to_csr(cpfem_ctx* ctx, void* stream, const double* mat, int64_t* indptr, int32_t* indices, double* data);

/* ---- adjoint (reverse mode) ----------------------------------------------------------
   Replaces the backward pass JAX derives through the custom_jvp local solver
   (models_copper.py:251-259) and the parameter VJP of implicit_vjp (solver.py:801-853). */
typedef struct CpfemVjp {
  const double* Pbar;      /* [9][n_qp] cotangent of P, or NULL */
  const double* sbar_cell; /* [9][n_cells] cotangent of the cell-averaged Cauchy stress, or NULL */
  const double* Fbar;      /* [9][n_qp] cotangent of Fp_inv_new, or NULL */
  const double* gbar;      /* [n_slip][n_qp] cotangent of g_new, or NULL */
  const double* slipbar;   /* [n_slip][n_qp] cotangent of slip_new, or NULL */
  double* Hbar;            /* [9][n_qp]      out: cotangent of u_grad          (mask 1) */
  double* Fpibar;          /* [9][n_qp]      out: cotangent of Fp_inv_old      (mask 2) */
  double* gbar_old;        /* [n_slip][n_qp] out: cotangent of g_old           (mask 4) */
  double* pbar;            /* [8][n_qp]      out: cotangent of xm, gss_a, h, t_sat, r, C11, C12, C44 (mask 8) */
  double* Rbar;            /* [9][n_qp]      out: cotangent of R               (mask 16) */
  int32_t mask;
  int32_t accumulate;      /* 1: add into the outputs */
} CpfemVjp;

/* Per-point vector-Jacobian product through the converged local Newton (implicit-function
   theorem; the slip cotangent passes through slip_old unchanged and is not an output). */
int cpfem_local_vjp(cpfem_ctx* ctx, void* stream, const double* dofs, const CpfemState* state,
                    const CpfemParams* params, const CpfemVjp* vjp);
/* out (n_dofs) = B^T Hbar: the transpose of the dofs -> u_grad map (no Dirichlet handling). */
int cpfem_ugrad_transpose(cpfem_ctx* ctx, void* stream, const double* Hbar, double* out);
/* out [9][n_qp] = scale * JxW_q * grad(lambda)_q with lambda set to zero on Dirichlet dofs. */
int cpfem_grad_field(cpfem_ctx* ctx, void* stream, const double* lambda, double scale, double* out);
/* matT = A^T in the same block-ELL layout (for the adjoint solve A^T lambda = v). */
int cpfem_transpose(cpfem_ctx* ctx, void* stream, const double* mat, double* matT);

/* Kernel-class timing with HIP events on the launch stream (of the Krylov classes 4, 5 and 8
   every `sample`-th launch is timed, of the others every launch; all are counted); enable
   resets the counters.  Classes (index): 0 point update (local Newton, P), 1 element
   stiffness/residual, 2 matrix gather, 3 residual gather, 4 Krylov SpMV kernels, 5 Krylov
   vector kernels, 6 point stress/history, 7 copy, 8 residual replacement of the mixed form,
   9 parked points of the point kernel, 10 consistent tangent of the assembly.  No counterpart
   in the reference. */
int cpfem_profile(cpfem_ctx* ctx, int32_t enable, int32_t sample);
int cpfem_profile_read(cpfem_ctx* ctx, int32_t n, double* ms, int64_t* launches, int64_t* timed);

/* Plain device copy of n doubles (bandwidth / counter calibration). */
int cpfem_copy(cpfem_ctx* ctx, void* stream, const double* src, double* dst, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* CPFEM_ABI_H_ */
