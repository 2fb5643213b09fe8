// Internal declarations of libmhd (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mhd.h"

namespace mhd {

constexpr int NDOF = 129;
constexpr int NU = 81, NP = 4, NJ = 36, NPHI = 8;
constexpr int OFF_P = 81, OFF_J = 85, OFF_PHI = 121;
constexpr int NQ = 27;          // 3x3x3 Gauss points per cell (quadrature degree 5)
constexpr int NQF = 9;          // 3x3 Gauss points per face
// Touched entries of one cell in the canonical row-major enumeration:
// u rows (81) x [u p j] (121), p rows (4) x u (81), j rows (36) x [u j phi] (125), phi rows (8) x j (36).
constexpr int ROWS_U_LEN = 121, ROWS_P_LEN = 81, ROWS_J_LEN = 125, ROWS_PHI_LEN = 36;
constexpr int E_U = 0;
constexpr int E_P = E_U + NU * ROWS_U_LEN;          // 9801
constexpr int E_J = E_P + NP * ROWS_P_LEN;          // 10125
constexpr int E_PHI = E_J + NJ * ROWS_J_LEN;        // 14625
constexpr int NENT = E_PHI + NPHI * ROWS_PHI_LEN;   // 14913
constexpr uint16_t RANK_NONE = 0xFFFF;

// Reference tables at the cell quadrature points (double precision).
struct RefTables {
  double w[NQ];                 // tensor Gauss weights on [0,1]^3
  double q2[NQ][27];            // Q2 scalar shape values
  double dq2[NQ][27][3];        // reference gradients
  double p1[NQ][4];             // dP1
  double q1[NQ][8];             // dQ1 / geometry shape values
  double dq1[NQ][8][3];         // geometry shape derivatives
  double rt[NQ][36][3];         // RT reference field
  double rtdiv[NQ][36];         // reference divergence
};

// Reference tables at the face quadrature points of each local face.
struct FaceTables {
  double w[NQF];
  double q1[6][NQF][8];
  double dq1[6][NQF][8][3];
  double rt[6][NQF][36][3];
  double drt[6][NQF][36][3][3];  // d rt_l / d xi_m  ([l][m])
  double dq2[6][NQF][27][3];     // Q2 reference gradients (skeleton stabilisation)
};

// Skeleton stabilisation: the velocity nodes of a face are the 27 of its first
// cell (slots 0..26) and those of the second cell not shared with the first
// (slot 27 + local node); SKEL_NONE marks the unused slots.
constexpr int SKEL_SLOTS = 54;
constexpr int SKEL_DOFS = 3 * SKEL_SLOTS;   // [component][slot]
constexpr int32_t SKEL_NONE = INT32_MIN;

void build_ref_tables(RefTables *T);
void build_face_tables(FaceTables *F);

// Pattern (host): CSR with int64 row pointers and int32 sorted column indices.
// skel_dofs ([nskel][SKEL_DOFS], may be null) adds the velocity couplings of
// the skeleton faces: every slot dof of a component with every other of it.
void build_pattern(int32_t ncells, const int32_t *cell_dofs, int64_t n,
                   std::vector<int64_t> &rowptr, std::vector<int32_t> &colind,
                   int32_t nskel = 0, const int32_t *skel_dofs = nullptr);

}  // namespace mhd

// Device-side mirror of the physics parameters.
struct MhdPhys {
  double alpha, beta, gamma, sigma;
  double B[3];
  double f0[3], fx[9], g0[3], gx[9];
  double zeta_u, zeta_j;
  int convection;
};

struct mhd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  std::string err;
  int32_t ncells = 0, nfaces = 0, ndir = 0;
  int64_t n = 0, nnz = 0;
  int32_t nfree[4] = {0, 0, 0, 0};
  MhdPhys phys{};
  // device arrays
  double *d_xyz = nullptr;        // [ncells][24]
  int32_t *d_dofs = nullptr;      // [ncells][129]
  int8_t *d_sign = nullptr;       // [ncells][6]
  double *d_dir = nullptr;        // [max(ndir,1)]
  int32_t *d_fcell = nullptr;     // thin-wall faces
  int8_t *d_flf = nullptr;
  double *d_ftau = nullptr, *d_fcw = nullptr, *d_fjw = nullptr;
  int64_t *d_rowptr = nullptr;    // [n+1]
  int32_t *d_colind = nullptr;    // [nnz]
  uint16_t *d_rank = nullptr;     // [ncells][NENT] rank of each touched entry inside its row
  double *d_val = nullptr;        // [nnz] CSR values of the last Jacobian assembled into the context
  bool jac_valid = false;         // d_val holds A(d_x): what mhd_precond_setup factors
  double pivot_tau = 0.03;        // threshold of a diagonal pivot in the patch inversions
  int64_t *d_tperm = nullptr;     // [nnz] CSR position of the transposed entry (CSC export)
  mhd::RefTables *d_ref = nullptr;
  mhd::FaceTables *d_fref = nullptr;
  // scratch
  double *d_x = nullptr, *d_r = nullptr;
  // host copies needed by the API and the solver set-up
  std::vector<int64_t> h_rowptr;
  std::vector<int32_t> h_colind;
  std::vector<int32_t> h_dofs;     // [ncells][129]
  std::vector<int32_t> h_verts;    // [ncells][8] (empty: no vertex-star smoother)
  std::vector<double> h_xyz;       // [ncells][24]
  int32_t nverts = 0;
  int32_t block_style = 0;
  // Jacobian scatter: 0 atomic, 1 colour classes, 2 element buffer + row assembly (default)
  int32_t assembly_mode = 2;
  bool dup_dofs = false;              // some cell holds a dof twice (forces mode 0)
  // cell colouring: no two cells of a class share a dof
  std::vector<int32_t> color_ptr;     // [ncolors + 1]
  int32_t *d_color_cells = nullptr;   // cells sorted by colour
  // row assembly: the (cell * NDOF + local row) pairs holding each free dof
  int32_t *d_crow = nullptr;          // [n + 1]
  int32_t *d_centry = nullptr;        // [crow[n]], increasing cell order per row
  int32_t max_row = 0;                // longest CSR row
  double *d_ebuf = nullptr;           // [ncells][NENT] element buffer (allocated on first use)
  double *d_rbuf = nullptr;           // [ncells][NDOF] element residuals (allocated on first use)
  double *d_scratch = nullptr;        // this device's allocation-free residual workspace (not owned)
  // skeleton stabilisation (mhd_create_stabilized)
  int32_t nskel = 0;
  int32_t *d_skel_dofs = nullptr;     // [nskel][SKEL_DOFS] slot dofs (free id, -(k+1) Dirichlet, SKEL_NONE)
  double *d_skel_K = nullptr;         // [nskel][54][54] face matrices of one velocity component
  int32_t *d_srow = nullptr;          // [n + 1] skeleton entries of each row
  int32_t *d_sentry = nullptr;        // face * SKEL_DOFS + component * 54 + slot, increasing face order per row
  int64_t nsentry = 0;                // entries over all rows
  uint16_t *d_skel_rank = nullptr;    // [nsentry][54] rank of each slot dof of the entry's face in the entry's row
  double *d_skel_res = nullptr;       // [nskel][SKEL_DOFS] the face products K_f x_c of the last residual
  // solver workspace (allocated lazily)
  struct Solver *solver = nullptr;
  // coarser level of the [u,j] block solve (owned), see mhd_attach_coarse
  struct CoarseLink *coarse = nullptr;
};

// Link from a level to the next coarser one: prolongation P (n1 x n1c CSR),
// its transpose, the u-state interpolation and V-cycle temporaries.
struct CoarseLink {
  mhd_ctx *ctx = nullptr;          // coarse context (owned)
  int64_t n1 = 0, n1c = 0, nuc = 0;
  int64_t *d_Prow = nullptr, *d_Trow = nullptr;
  int32_t *d_Pcol = nullptr, *d_Tcol = nullptr;
  double *d_Pval = nullptr, *d_Tval = nullptr;
  float *d_Pval32 = nullptr, *d_Tval32 = nullptr;
  int64_t *d_Rrow = nullptr;       // state interpolation R (nuc x fine nfree[u], CSR):
  int32_t *d_Rcol = nullptr;       // the fine velocity at every coarse u node
  double *d_Rval = nullptr;
  int64_t nnzP = 0;                // nonzeros of P (and of P^T)
  int coarse_iters = 10;           // Schwarz-GMRES steps on the coarsest level; 0: dense inverse
  // coarsest level solved exactly: explicit inverse of its [u,j] block (row-major)
  double *d_Ainv = nullptr;        // [n1c][n1c]
  double *d_LU = nullptr;          // [n1c][n1c] LU factors (set-up scratch)
  int32_t *d_piv = nullptr, *d_info = nullptr;
  double *d_gjw = nullptr;         // block Gauss-Jordan workspace (column and row panels, pivot block, check)
  int32_t *d_gjflag = nullptr;     // set when a diagonal block of the block Gauss-Jordan has a zero pivot
  int dense_fallbacks = 0;         // coarsest inversions redone with the pivoted LU
  double omega = 0.175;            // damping of the Schwarz smoother inside the V-cycle
  int step_mode = 1;               // coarse correction step: 0 full, 1 minimal residual
  void *rr = nullptr;              // [3 (n1 + 1)] fine residual, A11 yc, yc (T)
  double *d_ab = nullptr;          // the two dots of the minimal-residual coarse step
  void *rc = nullptr;              // [n1c] coarse rhs (T)
  void *ec = nullptr;              // [n1c] coarse correction (T)
  int tprec = -1;                  // element type of rr/rc/ec: 0 double, 1 float
};
void mhd_coarse_free(mhd_ctx *c);

// Slots of the operation counts (mhd_op_counts)
enum {
  OP_SPMV64 = 0, OP_SPMV32, OP_PATCH64, OP_PATCH32, OP_PATCH16, OP_DENSE, OP_VEC,
  OP_SETUP_CELLS, OP_SETUP_PATCH3, OP_SETUP_DENSE3, OP_OUTER
};
void mhd_solver_free(mhd_ctx *c);

// Launchers (mhd_assembly.hip)
int mhd_launch_residual(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s);
int mhd_launch_jacobian(mhd_ctx *c, const double *d_x, double *d_val, hipStream_t s);
int mhd_build_rank_map(mhd_ctx *c);
int mhd_build_tperm(mhd_ctx *c);
int mhd_launch_spmv(mhd_ctx *c, const double *d_val, const double *d_x, double *d_y, hipStream_t s);
int mhd_launch_cell_integrals(mhd_ctx *c, const double *d_x, double *d_out, hipStream_t s);
int mhd_launch_residual_mode(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s, int mode);
int mhd_launch_transpose_values(mhd_ctx *c, const double *d_val, double *d_val_t, hipStream_t s);
int mhd_noalloc_init(mhd_ctx *c);
int mhd_launch_residual_noalloc(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s);
int mhd_build_skeleton(mhd_ctx *c, const int32_t *d_fcell, const int8_t *d_flf, const int8_t *d_fperm,
                       const uint8_t *d_mslot, const double *d_mu);
int mhd_launch_skeleton_residual(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s);
int mhd_launch_skeleton_jacobian(mhd_ctx *c, double *d_val, hipStream_t s);

#define MHD_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      c->err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
      return -2;                                                                      \
    }                                                                                 \
  } while (0)
