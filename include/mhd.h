/*
 * mhd.h -- C ABI of libmhd, the MI355X-native backend for GridapMHD's hot path:
 * H1-HDiv residual / Newton-Jacobian assembly and the block-preconditioned
 * Newton-FGMRES solver.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success, a
 * positive value for a solver outcome (not converged, breakdown), a negative
 * value for an error; mhd_last_error() gives the text.  Calls are synchronous
 * unless they take a stream.  One context is single-threaded.
 *
 * Host buffers are owned by the caller; the context owns all device memory.
 * Functions with the _d suffix take device pointers (hipDeviceptr_t) and an
 * optional hipStream_t (NULL = the context's stream); they return after the
 * work is enqueued, not after it completes.
 */
#ifndef MHD_H
#define MHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Local dof layout of one hexahedral cell (129 dofs):
 *   u   [0, 81)    c*27 + a, Q2 node a = i + 3j + 9k at (i/2, j/2, k/2)
 *   p   [81, 85)   discontinuous P1: 1-x-y-z, x, y, z (reference coordinates)
 *   j   [85, 121)  Raviart-Thomas order 1: 24 face dofs (faces z-, z+, y-, y+,
 *                  x-, x+; 4 each at the face's 2x2 Gauss points), 12 interior
 *   phi [121, 129) discontinuous Q1 at vertex v = i + 2j + 4k
 * Local vertex v = i + 2j + 4k of the trilinear geometry. */
#define MHD_NDOF_CELL 129

typedef struct mhd_ctx mhd_ctx;

/* Geometry and thin-wall faces.  Replaces the model / triangulations of
 * setup_geometry! and the thin-wall measures of retrieve_bcs_params
 * (src/geometry.jl:6-36, src/weakforms.jl:153-162). */
typedef struct {
  int32_t ncells;
  int32_t nfaces_tw;            /* thin-wall boundary faces */
  const double *cell_xyz;       /* [ncells][8][3] */
  const int32_t *cell_verts;    /* [ncells][8] topological vertex ids (periodic images identified); NULL
                                   disables the vertex-star smoother of mhd_newton_solve */
  int32_t nverts;
  int32_t pad_;
  const int32_t *tw_face_cell;  /* [nfaces_tw] owning cell */
  const int8_t *tw_face_lface;  /* [nfaces_tw] local face 0..5 (z-, z+, y-, y+, x-, x+) */
  const double *tw_tau;         /* [nfaces_tw] penalty tau */
  const double *tw_cw;          /* [nfaces_tw] wall conductance ratio cw */
  const double *tw_jw;          /* [nfaces_tw] prescribed jw */
} mhd_mesh;

/* Cell dof tables of the multifield space.  Replaces setup_fe_spaces
 * (src/fespaces.jl:13-46).  block_style selects the multifield numbering
 * (src/fespaces.jl:4-9): 0 = Consecutive [u; p; j; phi] (the :julia solver),
 * 1 = Badia2024 blocks [u; j], p, phi, i.e. free ids ordered u, j, p, phi.
 * mhd_newton_solve needs block_style 1 and the p / phi dofs of every cell
 * consecutive (cell-major L2 spaces). */
typedef struct {
  int32_t nfree[4];             /* free dofs of u, p, j, phi */
  int32_t ndir;                 /* Dirichlet values */
  int32_t block_style;
  const int32_t *cell_dofs;     /* [ncells][129]: >= 0 free id, < 0 -> Dirichlet value -(id+1) */
  const int8_t *rt_sign;        /* [ncells][6] face orientation signs, NULL = all +1 */
  const double *dir_vals;       /* [ndir] */
} mhd_spaces;

/* Fluid parameters.  Replaces params[:fluid] as read by retrieve_fluid_params
 * (src/weakforms.jl:73-83; defaults src/parameters.jl:710-719). */
typedef struct {
  double alpha, beta, gamma, sigma;
  double B[3];
  double f0[3], fx[9];          /* f(x) = f0 + fx x (row-major) */
  double g0[3], gx[9];          /* g(x) = g0 + gx x */
  double zeta_u, zeta_j;        /* augmented-Lagrangian weights */
  int32_t convection;           /* 0 :none, 1 :picard, 2 :newton */
  int32_t pad_;
} mhd_params;

/* Solver options.  Replaces default_solver_params(::Val{:badia2024}) and the
 * solver built by Badia2024Solver (src/parameters.jl:259-272,
 * src/Solvers/badia2024.jl:2-48). */
typedef struct {
  int32_t max_newton;           /* :niter (20) */
  int32_t restart;              /* FGMRES Krylov dimension m (:niter_ls, 15) */
  int32_t max_krylov;           /* FGMRES iteration cap per Newton step */
  int32_t precision;            /* 0 FP64; 1 mixed: FP32 inner GMRES + FP32 smoother blocks inside FP64
                                   FGMRES refinement; 2 as 1 with BF16 smoother blocks; 3 as 1 with FP16
                                   smoother blocks scaled by a power of two per patch */
  double rtol, atol;            /* Newton: ||R|| <= max(atol, rtol ||R0||) */
  double ls_rtol;               /* linear: ||r|| <= max(ls_rtol ||b||, max(atol, rtol ||R0||) / 2) */
  int32_t inner_iters;          /* flexible-GMRES steps of the inexact [u,j] block solve */
  int32_t verbose;
} mhd_solver_opts;

typedef struct {
  int32_t newton_iters;
  int32_t krylov_iters;         /* total over the Newton steps */
  int32_t converged;
  int32_t flags;
  double residual0, residual;   /* first and last nonlinear residual norm */
  double t_residual, t_jacobian, t_setup, t_solve;  /* seconds, summed */
  double newton_residuals[64];
  int32_t krylov_per_step[64];
} mhd_solve_info;

/* Interior faces of the skeleton stabilisation
 *   a_Lambda(u, v) = int_Lambda (1/2) mu h^2 [grad u] : [grad v]
 * (src/weakforms.jl:685-688), Lambda the interior faces of the fluid domain,
 * h = sqrt(face area) (the cell size of the skeleton triangulation,
 * src/geometry.jl:150-158), mu per face (the sum over the
 * params[:bcs][:stabilization] entries that hold it, src/weakforms.jl:168-177).
 * Face k joins cell face_cell[2k] through its local face face_lface[2k] with
 * cell face_cell[2k+1] through face_lface[2k+1]; face_perm[9k + q] is the 3x3
 * face Gauss point (lower tangential axis fastest) of the second cell that
 * coincides with point q of the first. */
typedef struct {
  int32_t nfaces;
  int32_t pad_;
  const int32_t *face_cell;     /* [nfaces][2] */
  const int8_t *face_lface;     /* [nfaces][2] local faces 0..5 */
  const int8_t *face_perm;      /* [nfaces][9] */
  const double *mu;             /* [nfaces] */
} mhd_skeleton;

/* Create a context on `device`: copies inputs to HBM, builds the CSR/CSC
 * pattern and the cell-to-nonzero map.  Replaces _fe_operator +
 * SparseMatrixAssembler (src/main.jl:218-224). */
int mhd_create(mhd_ctx **ctx, const mhd_mesh *mesh, const mhd_spaces *spaces,
               const mhd_params *params, int32_t device);
/* mhd_create with the skeleton stabilisation in the residual and Jacobian
 * (the params_Λ loop of src/weakforms.jl:227-229, 248-250).  The pattern gains
 * the velocity couplings of the two cells of every face (same component);
 * the face matrices (geometry only, a_Lambda is linear) are computed once
 * here and kept in HBM (54 x 54 doubles per face). */
int mhd_create_stabilized(mhd_ctx **ctx, const mhd_mesh *mesh, const mhd_spaces *spaces,
                          const mhd_params *params, const mhd_skeleton *skeleton, int32_t device);
void mhd_destroy(mhd_ctx *ctx);
const char *mhd_last_error(const mhd_ctx *ctx);

/* Update fluid parameters / Dirichlet values (e.g. continuation, adjoint FD). */
int mhd_set_params(mhd_ctx *ctx, const mhd_params *params);
int mhd_set_dirichlet(mhd_ctx *ctx, const double *dir_vals);

/* Jacobian scatter.  2 (default) = element matrices stored densely per cell,
 * then every CSR row summed in on-chip memory and written once (coalesced,
 * bitwise reproducible); 1 = one launch per colour class of cells that share
 * no dof, plain read-modify-write (reproducible); 0 = one launch, FP64 atomic
 * adds (summation order not reproducible).  Mode 2 holds ncells * 14913
 * doubles of device memory from the first Jacobian on. */
int mhd_set_assembly_mode(mhd_ctx *ctx, int32_t mode);

/* Sizes of the system. */
int mhd_sizes(mhd_ctx *ctx, int64_t *nfree, int64_t *nnz);

/* Jacobian sparsity in Gridap's SparseMatrixCSC{Float64,Int64} layout:
 * 1-based colptr[n+1], rowval[nnz], rows sorted in every column. */
int mhd_pattern(mhd_ctx *ctx, int64_t *colptr, int64_t *rowval);

/* residual(op, xh) (src/weakforms.jl:211-232 + thin wall 721-726) on host buffers. */
int mhd_residual(mhd_ctx *ctx, const double *x_free, double *r_free);

/* jacobian(op, xh) (src/weakforms.jl:234-253 + 728-733): nzval in CSC order of mhd_pattern. */
int mhd_jacobian(mhd_ctx *ctx, const double *x_free, double *nzval);

/* Per-cell integrals of a state, out[6 * ncells]: (int u.u, int grad u : grad u,
 * int j.j, int (div j)^2, int p, |K|) -- the post-processing norms of _hunt
 * (src/Applications/hunt.jl:247-260) summed by the caller. */
int mhd_cell_integrals(mhd_ctx *ctx, const double *x_free, double *out);

/* Device-pointer variants (inputs resident in HBM).  The Jacobian values are
 * written in the library's CSR order (row-major, columns sorted); the pattern
 * is structurally symmetric, so CSR(A) and CSC(A) share index arrays and the
 * CSC values are CSR(A^T).  mhd_jacobian_d with d_nzval_csr = NULL assembles
 * into the context's own storage (and keeps d_x as its state): the Jacobian
 * that mhd_precond_setup and mhd_spmv_d use. */
int mhd_residual_d(mhd_ctx *ctx, const double *d_x, double *d_r, void *stream);
int mhd_jacobian_d(mhd_ctx *ctx, const double *d_x, double *d_nzval_csr, void *stream);
int mhd_csr_arrays(mhd_ctx *ctx, const int64_t **d_rowptr, const int32_t **d_colind);

/* Allocation-free residual: mhd_residual_d with the element residuals kept in
 * a fixed device workspace of the library (128 MiB per GPU, held until the
 * process ends) instead of context memory allocated on first use.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  It calls no
 * hipMalloc, hipFree or hipMemset.  At most 130,055 cells (16,777,216 / 129
 * element-residual doubles); larger meshes are rejected with an error.  Every
 * call on one GPU shares that GPU's workspace, so calls that overlap on one
 * GPU (any context, any stream) race.  A context made by mhd_create_stabilized
 * adds its skeleton term through face-product memory allocated at creation. */
int mhd_residual_d_noalloc(mhd_ctx *ctx, const double *d_x, double *d_r, void *stream);

/* y = A x with the Jacobian last assembled into the context (device pointers). */
int mhd_spmv_d(mhd_ctx *ctx, const double *d_x, double *d_y, void *stream);

/* Newton + FGMRES + Badia2024 block preconditioner.  Replaces
 * solve!(xh, Badia2024Solver(op, params), op) (src/main.jl:150, 275).
 * x_free is the initial guess on entry, the solution on return. */
int mhd_newton_solve(mhd_ctx *ctx, double *x_free, const mhd_solver_opts *opts, mhd_solve_info *info);
void mhd_default_solver_opts(mhd_solver_opts *opts);

/* Adjoint / sensitivities (no counterpart in the reference, which has no
 * adjoint; they serve the differentiation configuration).
 * mhd_dresidual_dparam: out = dR/dtheta at x for theta = gamma (param 0: the
 * Lorentz term -int (j x B).v) or cw (param 1: int_G tau (s.n) n.(grad j).n
 * over every thin-wall face).  dR/dHa = (2 Ha / Re) dR/dgamma in the :cfd scaling.
 * mhd_adjoint_solve: lambda = A(x)^-T rhs with Badia2024's upper
 * block-triangular preconditioner built on A^T (which has A's saddle-point
 * block structure; an attached V-cycle hierarchy works on the transposed
 * Jacobian of every level).  Then
 * dJ/dtheta = dJ/dtheta|_explicit - lambda . dR/dtheta.
 * mhd_linear_solve: sol = A(x)^-1 rhs with the forward preconditioner; the
 * tangent (forward-sensitivity) form of the same gradient,
 * dJ/dtheta = dJ/dx . A^-1 (-dR/dtheta), one solve per parameter. */
int mhd_dresidual_dparam(mhd_ctx *ctx, const double *x_free, int32_t param, double *out);
int mhd_adjoint_solve(mhd_ctx *ctx, const double *x_state, const double *rhs, double *lambda,
                      const mhd_solver_opts *opts, mhd_solve_info *info);
int mhd_linear_solve(mhd_ctx *ctx, const double *x_state, const double *rhs, double *sol,
                     const mhd_solver_opts *opts, mhd_solve_info *info);

/* Attach a coarser level to the [u,j]-block solve (geometric coarse-grid
 * correction; the x,y semi-coarsened hierarchy of the reference's GMG,
 * src/Meshers/hunt_mesher.jl:140-155, src/Solvers/gmg.jl:323-355).  `coarse` is a
 * context on the coarse mesh with the same parameters and block numbering;
 * P (CSR, nrows = nfree[u]+nfree[j] of ctx, ncols = the same of coarse) maps
 * coarse [u,j] values to fine ones; R (CSR, coarse nfree[u] rows x fine
 * nfree[u] columns) gives the fine velocity at the coarse u nodes, the state
 * the coarse Jacobian is assembled at.  ctx takes ownership of coarse
 * (destroyed with ctx; do not destroy it separately).  The [u,j] smoother of
 * mhd_newton_solve then becomes a V-cycle whose vertex-star Schwarz
 * pre/post-smoothing is damped by smoother_weight (0, 1] (the patches overlap
 * up to 8-fold; 0.175 is a good value, 0.25 can diverge).  On the coarsest
 * level coarse_iters Schwarz-GMRES steps are taken, or, with coarse_iters = 0,
 * the level is solved exactly with a dense FP64 inverse (at most 32768 [u,j]
 * dofs there).  coarse_step 0 adds the coarse correction with a full step;
 * 1 scales it by the minimal-residual step (robust where the rediscretised
 * coarse level misrepresents the fine one, e.g. high Ha on coarse meshes). */
int mhd_attach_coarse(mhd_ctx *ctx, mhd_ctx *coarse, int64_t nrows, int64_t ncols, const int64_t *P_rowptr,
                      const int32_t *P_col, const double *P_val, const int64_t *R_rowptr, const int32_t *R_col,
                      const double *R_val, int32_t coarse_iters, double smoother_weight, int32_t coarse_step);

/* Preconditioner set-up for the Jacobian last assembled into the context (by
 * mhd_jacobian, a solve, or mhd_jacobian_d with d_nzval_csr = NULL; an error
 * before any): the numerical_setup! of the block solver
 * (src/Solvers/badia2024.jl:20-31) -- Schur scalings, vertex-star patch
 * extraction and batched inversion, coarse levels at the injected state.
 * opts->precision selects the storage of the patch inverses (0 FP64, 1 FP32,
 * 2 BF16, 3 FP16 with a per-patch power-of-two scale).  A singular patch is an error. */
int mhd_precond_setup(mhd_ctx *ctx, const mhd_solver_opts *opts);

/* y += sum_P R_P^T A_P^-1 R_P r on the [u,j] block (device pointers, length
 * nfree[u] + nfree[j]; y is accumulated into).  The smoother of the [u,j]
 * block solve, in the precision of the last mhd_precond_setup: r and y are
 * double for precision 0 and float (the inner-solve vectors) for 1 to 3.
 * The patch products of a dof are summed in FP64 in a fixed order (the same
 * result in every run) through a workspace of the context, so calls that
 * overlap each other or a solve of the same context on another stream race. */
int mhd_smoother_apply_d(mhd_ctx *ctx, const void *d_r, void *d_y, void *stream);
/* The vertex-patch inversions of mhd_precond_setup (batched Gauss-Jordan)
 * pivot on the diagonal while |a_kk| >= tau max_{i>k} |a_ik| and invert a patch
 * whose diagonal pivot fails that test again with partial pivoting (the
 * largest |a_ik|).  tau in [0, 1]: 1 = every patch with partial pivoting,
 * 0 = diagonal pivots unless one is zero.  Default 0.03; applies to the
 * attached coarse levels too.  No counterpart in the reference (its patch
 * solves are library LU factorisations). */
int mhd_set_pivot_threshold(mhd_ctx *ctx, double tau);

/* Patch count, padded patch dimension, sum over patches of size^2, and the
 * patches of the last set-up inverted with partial pivoting. */
int mhd_smoother_info(mhd_ctx *ctx, int64_t *npatch, int64_t *patch_dim, int64_t *patch_entries, int64_t *npivoted);

/* Operation counts of the solver work since the last reset, summed over this
 * context and its coarse levels.  No counterpart in the reference: a
 * measurement aid that lets other hardware be priced on exactly the work of a
 * solve (the CPU Newton-step baseline).  out[MHD_NOPS]:
 *   0  nonzeros streamed by SpMVs with FP64 values (A, the A11 / A1x blocks, P, P^T)
 *   1  nonzeros streamed by SpMVs with FP32 values
 *   2  patch-inverse entries applied by the smoother in FP64
 *   3  ... in FP32
 *   4  ... in BF16
 *   5  entries of the dense coarsest inverse applied
 *   6  bytes of Krylov vector work (dot products, Gram-Schmidt updates,
 *      linear combinations, minimal-residual coarse steps)
 *   7  cells of the coarse-level Jacobian assemblies (preconditioner set-up)
 *   8  sum of dim^3 over the inverted patches (set-up)
 *   9  sum of n^3 over the dense coarsest inversions (set-up; an inverse whose
 *      block Gauss-Jordan check fails and is redone by the pivoted LU counts twice)
 *   10 outer FGMRES iterations
 * reset != 0 zeroes the counts after reading them. */
#define MHD_NOPS 11
int mhd_op_counts(mhd_ctx *ctx, double *out, int reset);

#ifdef __cplusplus
}
#endif
#endif
