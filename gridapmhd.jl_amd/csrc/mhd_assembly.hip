// Cell-wise assembly of the H1-HDiv residual and Newton Jacobian on CDNA4.
//
// Weak form (src/weakforms.jl:255-312, conv at :670):
//   R(x)(y) = int  beta grad u : grad v + alpha v.((grad u)^T u) - p div v - gamma (j x B).v
//                 - div u q + j.s - sigma phi div s - sigma (u x B).s - div j phi_t
//                 - f.v - g.s  [+ zeta_u Pi_p(div u) div v + zeta_j div j div s]
//   J(x)(dx,y) = its derivative (Picard keeps only (grad du)^T u).
// Thin-wall faces (src/weakforms.jl:721-733, residual evaluated at x):
//   R += int_G tau (s.n)(j.n - jw + cw n.(grad j).n),  J += int_G tau (s.n)(dj.n + cw n.(grad dj).n)
//
// One workgroup per cell: quadrature geometry, pushed-forward bases and the
// state of x at the 27 Gauss points are staged in LDS; the 14913 touched
// entries of the element matrix are spread over the workgroup in the
// canonical row-major order of the cell's rank map, so the rank-map reads and
// the row-local scatter into the CSR values are contiguous.
#include <mutex>

#include "mhd_internal.h"

using namespace mhd;

namespace {

constexpr int WG = 256;

struct CellLDS {
  double W[NQ];              // w |det J|
  union {
    struct {
      double Ji[NQ][9];      // (J^-1)[l][i]
      double Jp[NQ][9];      // J[k][l] / det J
    };
    struct {
      double D[4][NU];       // AL-u: int psi_k div v_b
      double MD[4][NU];      // AL-u: M^-1 D
      double Mi[16];         // AL-u: inverse dP1 mass
      double cu[4];          // AL-u: M^-1 int psi div u
      double Mq[16];         // AL-u: dP1 mass
      double dq[4];          // AL-u: int psi div u
    };
  };
  double idet[NQ];           // 1 / det J
  double xq[NQ][3];          // physical point
  double q2[NQ][27];         // reference Q2 values
  double p1[NQ][4];          // reference dP1 values
  double q1[NQ][8];          // reference dQ1 values
  double gphi[NQ][27][3];    // physical Q2 gradients
  double cphi[NQ][27];       // u . grad phi (convective derivative)
  double psi[NQ][36][3];     // Piola-mapped RT basis (with face signs)
  double dpsi[NQ][36];       // its divergence
  double u[NQ][3];           // state of x
  double gu[NQ][9];          // grad u [i][k] = d u_k / d x_i
  double wN[NQ][9];          // W alpha d_d u_c  at [3c + d]  (Newton term weights)
  double p[NQ], jv[NQ][3], divj[NQ], ph[NQ], divu[NQ];
  double xl[NDOF];
  double xyz[24];
  double sg[6];
  double B[3], f0[3], fx[9], g0[3], gx[9];
  int64_t rp[NDOF];
  int32_t g[NDOF];
};

__device__ inline void inv4(const double *m, double *inv) {
  // adjugate / determinant of a 4x4 row-major matrix
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  const double id = 1.0 / det;
  for (int i = 0; i < 16; ++i) inv[i] *= id;
}

// Load one cell and stage everything the integrands need.  `want_state`
// evaluates x at the quadrature points (always needed except for a linear
// Jacobian), `want_al` the zeta_u local projection.
__device__ void stage_cell(CellLDS &S, int c, const double *__restrict__ xyz, const int32_t *__restrict__ dofs,
                           const int8_t *__restrict__ sign, const double *__restrict__ dir,
                           const double *__restrict__ x, const RefTables *__restrict__ T,
                           const int64_t *__restrict__ rowptr, const MhdPhys &ph, bool want_al) {
  const int tid = threadIdx.x;
  if (tid < 24) S.xyz[tid] = xyz[(int64_t)c * 24 + tid];
  if (tid < 6) S.sg[tid] = sign ? (double)sign[(int64_t)c * 6 + tid] : 1.0;
  if (tid == 0) {
    for (int i = 0; i < 3; ++i) {
      S.B[i] = ph.B[i];
      S.f0[i] = ph.f0[i];
      S.g0[i] = ph.g0[i];
    }
    for (int i = 0; i < 9; ++i) {
      S.fx[i] = ph.fx[i];
      S.gx[i] = ph.gx[i];
    }
  }
  if (tid < NDOF) {
    const int32_t g = dofs[(int64_t)c * NDOF + tid];
    S.g[tid] = g;
    S.xl[tid] = g >= 0 ? x[g] : dir[-g - 1];
    S.rp[tid] = (g >= 0 && rowptr) ? rowptr[g] : -1;
  }
  __syncthreads();
  // geometry: one thread per quadrature point
  if (tid < NQ) {
    const int q = tid;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, X[3] = {0, 0, 0};
#pragma unroll
    for (int v = 0; v < 8; ++v)
      for (int i = 0; i < 3; ++i) {
        const double xv = S.xyz[3 * v + i];
        X[i] += T->q1[q][v] * xv;
        for (int l = 0; l < 3; ++l) J[i][l] += xv * T->dq1[q][v][l];
      }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    double *Ji = S.Ji[q];
    Ji[0 * 3 + 0] = c00 * id;
    Ji[1 * 3 + 0] = c01 * id;
    Ji[2 * 3 + 0] = c02 * id;
    Ji[0 * 3 + 1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1 * 3 + 1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2 * 3 + 1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0 * 3 + 2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1 * 3 + 2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2 * 3 + 2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) S.Jp[q][3 * k + l] = J[k][l] * id;
    S.W[q] = T->w[q] * fabs(det);
    S.idet[q] = id;
    for (int i = 0; i < 3; ++i) S.xq[q][i] = X[i];
  }
  for (int idx = tid; idx < NQ * 27; idx += WG) S.q2[idx / 27][idx % 27] = T->q2[idx / 27][idx % 27];
  for (int idx = tid; idx < NQ * 4; idx += WG) S.p1[idx / 4][idx % 4] = T->p1[idx / 4][idx % 4];
  for (int idx = tid; idx < NQ * 8; idx += WG) S.q1[idx / 8][idx % 8] = T->q1[idx / 8][idx % 8];
  __syncthreads();
  // push-forward of the Q2 gradients and the RT basis (fixed trip counts, so
  // the reference-table loads of all rounds are issued together)
#pragma unroll
  for (int k = 0; k < (NQ * 27 + WG - 1) / WG; ++k) {
    const int idx = tid + k * WG;
    if (idx >= NQ * 27) break;
    const int q = idx / 27, a = idx % 27;
    const double *Ji = S.Ji[q];
    const double d0 = T->dq2[q][a][0], d1 = T->dq2[q][a][1], d2 = T->dq2[q][a][2];
    for (int i = 0; i < 3; ++i) S.gphi[q][a][i] = Ji[0 * 3 + i] * d0 + Ji[1 * 3 + i] * d1 + Ji[2 * 3 + i] * d2;
  }
#pragma unroll
  for (int k = 0; k < (NQ * 36 + WG - 1) / WG; ++k) {
    const int idx = tid + k * WG;
    if (idx >= NQ * 36) break;
    const int q = idx / 36, r = idx % 36;
    const double sgn = r < 24 ? S.sg[r >> 2] : 1.0;
    const double *Jp = S.Jp[q];
    const double r0 = T->rt[q][r][0], r1 = T->rt[q][r][1], r2 = T->rt[q][r][2];
    for (int k = 0; k < 3; ++k) S.psi[q][r][k] = sgn * (Jp[3 * k + 0] * r0 + Jp[3 * k + 1] * r1 + Jp[3 * k + 2] * r2);
    S.dpsi[q][r] = sgn * T->rtdiv[q][r] * S.idet[q];
  }
  __syncthreads();
  // state of x at the quadrature points: 27 x 18 scalar reductions of at most
  // 36 terms (div u is the trace of grad u, taken after the barrier)
  for (int idx = tid; idx < NQ * 18; idx += WG) {
    const int k = idx / NQ, q = idx % NQ;  // quantity-major: a wave's lanes share few branches
    double acc = 0.0;
    if (k < 3) {  // u_k
#pragma unroll
      for (int a = 0; a < 27; ++a) acc += S.xl[k * 27 + a] * S.q2[q][a];
      S.u[q][k] = acc;
    } else if (k < 12) {  // grad u [i][kk]
      const int i = (k - 3) / 3, kk = (k - 3) % 3;
#pragma unroll
      for (int a = 0; a < 27; ++a) acc += S.xl[kk * 27 + a] * S.gphi[q][a][i];
      S.gu[q][(k - 3)] = acc;
    } else if (k == 12) {
#pragma unroll
      for (int m = 0; m < 4; ++m) acc += S.xl[OFF_P + m] * S.p1[q][m];
      S.p[q] = acc;
    } else if (k < 16) {
      const int kk = k - 13;
#pragma unroll
      for (int r = 0; r < 36; ++r) acc += S.xl[OFF_J + r] * S.psi[q][r][kk];
      S.jv[q][kk] = acc;
    } else if (k == 16) {
#pragma unroll
      for (int r = 0; r < 36; ++r) acc += S.xl[OFF_J + r] * S.dpsi[q][r];
      S.divj[q] = acc;
    } else {
#pragma unroll
      for (int v = 0; v < 8; ++v) acc += S.xl[OFF_PHI + v] * S.q1[q][v];
      S.ph[q] = acc;
    }
  }
  __syncthreads();
  if (tid < NQ) S.divu[tid] = S.gu[tid][0] + S.gu[tid][4] + S.gu[tid][8];
  if (want_al) {
    __syncthreads();
    for (int idx = tid; idx < 4 * NU; idx += WG) {
      const int k = idx / NU, b = idx % NU, cb = b / 27, nb = b % 27;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc += S.W[q] * S.p1[q][k] * S.gphi[q][nb][cb];
      S.D[k][b] = acc;
    }
    if (tid < 20) {
      double acc = 0.0;
      if (tid < 16) {
        const int k = tid >> 2, l = tid & 3;
        for (int q = 0; q < NQ; ++q) acc += S.W[q] * S.p1[q][k] * S.p1[q][l];
        S.Mq[tid] = acc;
      } else {
        const int k = tid - 16;
        for (int q = 0; q < NQ; ++q) acc += S.W[q] * S.p1[q][k] * S.divu[q];
        S.dq[k] = acc;
      }
    }
    __syncthreads();
    if (tid == 0) {
      inv4(S.Mq, S.Mi);
      for (int k = 0; k < 4; ++k)
        S.cu[k] = S.Mi[4 * k + 0] * S.dq[0] + S.Mi[4 * k + 1] * S.dq[1] + S.Mi[4 * k + 2] * S.dq[2] + S.Mi[4 * k + 3] * S.dq[3];
    }
    __syncthreads();
    for (int idx = tid; idx < 4 * NU; idx += WG) {
      const int k = idx / NU, b = idx % NU;
      double acc = 0.0;
      for (int l = 0; l < 4; ++l) acc += S.Mi[4 * k + l] * S.D[l][b];
      S.MD[k][b] = acc;
    }
  }
  __syncthreads();
  // convective derivative and Newton weights for the Jacobian tiles
  for (int idx = tid; idx < NQ * 27; idx += WG) {
    const int q = idx / 27, a = idx % 27;
    S.cphi[q][a] = S.u[q][0] * S.gphi[q][a][0] + S.u[q][1] * S.gphi[q][a][1] + S.u[q][2] * S.gphi[q][a][2];
  }
  for (int idx = tid; idx < NQ * 9; idx += WG) {
    const int q = idx / 9, cd = idx % 9, cc = cd / 3, dd = cd % 3;
    S.wN[q][cd] = S.W[q] * ph.alpha * S.gu[q][3 * dd + cc];
  }
  __syncthreads();
}

__device__ inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Decode canonical entry e -> (local row a, local column b).
__device__ inline void decode_entry(int e, int &a, int &b) {
  if (e < E_P) {
    a = e / ROWS_U_LEN;
    b = e % ROWS_U_LEN;
  } else if (e < E_J) {
    const int t = e - E_P;
    a = OFF_P + t / ROWS_P_LEN;
    b = t % ROWS_P_LEN;
  } else if (e < E_PHI) {
    const int t = e - E_J;
    a = OFF_J + t / ROWS_J_LEN;
    const int k = t % ROWS_J_LEN;
    b = k < NU ? k : (k < NU + NJ ? OFF_J + (k - NU) : OFF_PHI + (k - NU - NJ));
  } else {
    const int t = e - E_PHI;
    a = OFF_PHI + t / ROWS_PHI_LEN;
    b = OFF_J + t % ROWS_PHI_LEN;
  }
}

// Jacobian entry (test a, trial b) of the fluid terms.
__device__ double jac_entry(const CellLDS &S, const RefTables *__restrict__ T, const MhdPhys &ph, int a, int b) {
  double acc = 0.0;
  if (a < OFF_P) {  // velocity test v = phi_na e_ca
    const int ca = a / 27, na = a % 27;
    if (b < OFF_P) {
      const int cb = b / 27, nb = b % 27;
      const bool diag = ca == cb;
      for (int q = 0; q < NQ; ++q) {
        double t = 0.0;
        if (diag) {
          const double *ga = S.gphi[q][na], *gb = S.gphi[q][nb];
          t = ph.beta * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
          if (ph.convection)
            t += ph.alpha * S.q2[q][na] * (S.u[q][0] * gb[0] + S.u[q][1] * gb[1] + S.u[q][2] * gb[2]);
        }
        if (ph.convection == 2) t += ph.alpha * S.q2[q][na] * S.q2[q][nb] * S.gu[q][3 * cb + ca];
        acc += S.W[q] * t;
      }
      if (ph.zeta_u != 0.0) {
        double al = 0.0;
        for (int k = 0; k < 4; ++k) {
          double m = 0.0;
          for (int l = 0; l < 4; ++l) m += S.Mi[4 * k + l] * S.D[l][b];
          al += S.D[k][a] * m;
        }
        acc += ph.zeta_u * al;
      }
    } else if (b < OFF_J) {  // - dp div v
      const int k = b - OFF_P;
      for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.p1[q][k] * S.gphi[q][na][ca];
    } else {  // - gamma (dj x B).v
      const int r = b - OFF_J;
      const int i1 = ca == 2 ? 0 : ca + 1, i2 = ca == 0 ? 2 : ca - 1;
      const double b1 = S.B[i1], b2 = S.B[i2];
      for (int q = 0; q < NQ; ++q) {
        const double cr = S.psi[q][r][i1] * b2 - S.psi[q][r][i2] * b1;  // (psi x B)_ca
        acc -= ph.gamma * S.W[q] * S.q2[q][na] * cr;
      }
    }
  } else if (a < OFF_J) {  // pressure test: - div du q
    const int k = a - OFF_P;
    const int cb = b / 27, nb = b % 27;
    for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.p1[q][k] * S.gphi[q][nb][cb];
  } else if (a < OFF_PHI) {  // current test s = psi_r
    const int r = a - OFF_J;
    if (b < OFF_P) {  // - sigma (du x B).s
      const int cb = b / 27, nb = b % 27;
      const int i1 = cb == 2 ? 0 : cb + 1, i2 = cb == 0 ? 2 : cb - 1;
      const double b1 = S.B[i1], b2 = S.B[i2];
      for (int q = 0; q < NQ; ++q) {
        // (e_cb x B).s = (B x s)_cb
        const double t = b1 * S.psi[q][r][i2] - b2 * S.psi[q][r][i1];
        acc -= ph.sigma * S.W[q] * S.q2[q][nb] * t;
      }
    } else if (b < OFF_PHI) {  // dj.s + zeta_j div dj div s
      const int rb = b - OFF_J;
      for (int q = 0; q < NQ; ++q) {
        const double *s = S.psi[q][r], *d = S.psi[q][rb];
        double t = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
        if (ph.zeta_j != 0.0) t += ph.zeta_j * S.dpsi[q][r] * S.dpsi[q][rb];
        acc += S.W[q] * t;
      }
    } else {  // - sigma dphi div s
      const int v = b - OFF_PHI;
      for (int q = 0; q < NQ; ++q) acc -= ph.sigma * S.W[q] * S.q1[q][v] * S.dpsi[q][r];
    }
  } else {  // potential test: - div dj phi_t
    const int v = a - OFF_PHI, rb = b - OFF_J;
    for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.q1[q][v] * S.dpsi[q][rb];
  }
  return acc;
}

// Residual entry of test function a.
// mode 0: residual; mode 1: dR/dgamma = -int (j x B).v (u rows only)
__device__ double res_entry(const CellLDS &S, const RefTables *__restrict__ T, const MhdPhys &ph, int a, int mode) {
  double acc = 0.0;
  if (mode == 1) {
    if (a >= OFF_P) return 0.0;
    const int ca = a / 27, na = a % 27;
    const int i1 = ca == 2 ? 0 : ca + 1, i2 = ca == 0 ? 2 : ca - 1;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q)
      acc -= S.W[q] * (S.jv[q][i1] * S.B[i2] - S.jv[q][i2] * S.B[i1]) * S.q2[q][na];
    return acc;
  }
  if (a < OFF_P) {
    const int ca = a / 27, na = a % 27;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const double *gv = S.gphi[q][na];
      const double phv = S.q2[q][na];
      // beta grad u : grad v with v = phi e_ca -> sum_i gu[i][ca] dphi_i
      double t = ph.beta * (S.gu[q][0 * 3 + ca] * gv[0] + S.gu[q][1 * 3 + ca] * gv[1] + S.gu[q][2 * 3 + ca] * gv[2]);
      if (ph.convection) {
        // ((grad u)^T u)_ca = sum_i gu[i][ca] u_i
        const double cv = S.gu[q][0 * 3 + ca] * S.u[q][0] + S.gu[q][1 * 3 + ca] * S.u[q][1] + S.gu[q][2 * 3 + ca] * S.u[q][2];
        t += ph.alpha * phv * cv;
      }
      t -= S.p[q] * gv[ca];
      const int i1 = ca == 2 ? 0 : ca + 1, i2 = ca == 0 ? 2 : ca - 1;
      const double jxb = S.jv[q][i1] * S.B[i2] - S.jv[q][i2] * S.B[i1];  // (j x B)_ca
      t -= ph.gamma * jxb * phv;
      double f = S.f0[ca];
      for (int k = 0; k < 3; ++k) f += S.fx[3 * ca + k] * S.xq[q][k];
      t -= f * phv;
      acc += S.W[q] * t;
    }
    if (ph.zeta_u != 0.0) {
      double al = 0.0;
      for (int k = 0; k < 4; ++k) al += S.D[k][a] * S.cu[k];
      acc += ph.zeta_u * al;
    }
  } else if (a < OFF_J) {
    const int k = a - OFF_P;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.divu[q] * S.p1[q][k];
  } else if (a < OFF_PHI) {
    const int r = a - OFF_J;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const double *s = S.psi[q][r];
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int i1 = i == 2 ? 0 : i + 1, i2 = i == 0 ? 2 : i - 1;
        const double uxb = S.u[q][i1] * S.B[i2] - S.u[q][i2] * S.B[i1];
        double g = S.g0[i];
        for (int k = 0; k < 3; ++k) g += S.gx[3 * i + k] * S.xq[q][k];
        t += (S.jv[q][i] - ph.sigma * uxb - g) * s[i];
      }
      t -= ph.sigma * S.ph[q] * S.dpsi[q][r];
      if (ph.zeta_j != 0.0) t += ph.zeta_j * S.divj[q] * S.dpsi[q][r];
      acc += S.W[q] * t;
    }
  } else {
    const int v = a - OFF_PHI;
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.divj[q] * S.q1[q][v];
  }
  return acc;
}

typedef double d4_t __attribute__((ext_vector_type(4)));

// How a cell's Jacobian entries leave the kernel:
//   SCATTER_ATOMIC  FP64 atomic add into the CSR values;
//   SCATTER_COLOUR  plain read-modify-write (the launch covers one colour class:
//                   no two of its cells share a dof);
//   SCATTER_EBUF    plain store of every canonical entry into the cell's slot of
//                   the element buffer (summed into CSR rows by k_row_assemble).
enum { SCATTER_ATOMIC = 0, SCATTER_COLOUR = 1, SCATTER_EBUF = 2 };

// Entry e (canonical index) of local row lr / local column lc with value v.
template <int MODE>
__device__ inline void emit(const CellLDS &S, const uint16_t *__restrict__ rk, double *__restrict__ out, int e, int lr,
                            int lc, double v) {
  if (MODE == SCATTER_EBUF) {
    out[e] = v;
    return;
  }
  if (S.g[lr] < 0 || S.g[lc] < 0) return;
  double *p = &out[S.rp[lr] + rk[e]];
  if (MODE == SCATTER_ATOMIC) atomicAdd(p, v);
  else *p += v;
}

__device__ inline d4_t mfma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Scatter-add one 16x16 MFMA result tile.  Lane l holds rows l/16 + 4 r
// (r = 0..3), column l%16 of the tile (v_mfma_f64_16x16x4f64 output layout); (row0, col0) are the local dof offsets of the tile,
// `ent(i, j)` maps tile-local (i, j) to the canonical entry of the rank map.
template <int MODE, class EntFn>
__device__ inline void scatter_tile(const CellLDS &S, const d4_t &acc, double scale, int row0, int nrow, int col0,
                                    int ncol, const uint16_t *__restrict__ rk, double *__restrict__ out, EntFn ent) {
  const int lane = threadIdx.x & 63;
  const int j = lane & 15;
  if (j >= ncol) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = (lane >> 4) + 4 * r;
    if (i >= nrow) continue;
    emit<MODE>(S, rk, out, ent(i, j), row0 + i, col0 + j, scale * acc[r]);
  }
}

// Jacobian of one cell as FP64 MFMA tiles (16x16x4):
//   uu(c,d) = delta_cd S + N^{cd} + zeta_u D_c^T M^-1 D_d with
//     S[a][b]    = sum_q W (beta grad phi_a . grad phi_b + alpha phi_a u.grad phi_b)   (K = 27 q x 4)
//     N^{cd}[a][b] = sum_q W alpha d_d u_c phi_a phi_b                                (K = 27 q)
//   X_c[a][r]  = sum_q W phi_a (psi_r x B)_c  ->  uj = -gamma X,  ju = +sigma X^T         (K = 27 q)
//     (-sigma (e_c phi_a x B).psi_r = +sigma phi_a (psi_r x B)_c, so ju = -(sigma/gamma) uj^T)
//   jj[r][s]   = sum_q W (psi_r . psi_s + zeta_j div psi_r div psi_s)                    (K = 27 q x 4)
// The small blocks up / pu (81 x 4) and jphi / phij (36 x 8) run on the vector ALUs.
// `out` is the CSR value array (atomic / colour modes) or the cell's element
// buffer slot (NENT doubles, element-buffer mode).
template <int MODE>
__device__ void cell_jacobian_tiles(const CellLDS &S, const MhdPhys &ph, const uint16_t *__restrict__ rk,
                                    double *__restrict__ val) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const d4_t zero = {0.0, 0.0, 0.0, 0.0};
  // job list: 4 uu super-tiles, 9 jj tiles, 18 cross tiles; wave w takes jobs w, w+4, ...
  for (int job = wave; job < 4 + 9 + 18; job += WG / 64) {
    if (job < 4) {
      const int ti = job >> 1, tj = job & 1;
      const int a = 16 * ti + li, b = 16 * tj + li;
      const bool av = a < 27, bv = b < 27;
      d4_t Sacc = zero;
#pragma unroll 9
      for (int q = 0; q < NQ; ++q) {
        double A = 0.0, Bv = 0.0;
        if (av) A = lk < 3 ? S.gphi[q][a][lk] : S.q2[q][a];
        if (bv) Bv = S.W[q] * (lk < 3 ? ph.beta * S.gphi[q][b][lk] : (ph.convection ? ph.alpha * S.cphi[q][b] : 0.0));
        Sacc = mfma16(A, Bv, Sacc);
      }
      // Q2 values of this lane's rows / columns at its 7 quadrature points,
      // shared by the nine (c, d) blocks
      double qa[7], qb[7];
#pragma unroll
      for (int st = 0; st < 7; ++st) {
        const int q = 4 * st + lk;
        qa[st] = (av && q < NQ) ? S.q2[q][a] : 0.0;
        qb[st] = (bv && q < NQ) ? S.q2[q][b] : 0.0;
      }
#pragma unroll 1
      for (int cd = 0; cd < 9; ++cd) {
        const int c = cd / 3, d = cd % 3;
        d4_t acc = c == d ? Sacc : zero;
        if (ph.convection == 2) {
          double wq[7];
#pragma unroll
          for (int st = 0; st < 7; ++st) wq[st] = 4 * st + lk < NQ ? S.wN[4 * st + lk][cd] : 0.0;
#pragma unroll
          for (int st = 0; st < 7; ++st) acc = mfma16(qa[st], wq[st] * qb[st], acc);
        }
        if (ph.zeta_u != 0.0) {
          const double A = av ? S.D[lk][c * 27 + a] : 0.0;
          const double Bv = bv ? ph.zeta_u * S.MD[lk][d * 27 + b] : 0.0;
          acc = mfma16(A, Bv, acc);
        }
        const int r0 = c * 27 + 16 * ti, c0 = d * 27 + 16 * tj;
        scatter_tile<MODE>(S, acc, 1.0, r0, min(16, 27 - 16 * ti), c0, min(16, 27 - 16 * tj), rk, val,
                     [&](int i, int j) { return (r0 + i) * ROWS_U_LEN + (c0 + j); });
      }
    } else if (job < 13) {
      const int t = job - 4, ti = t / 3, tj = t % 3;
      const int r = 16 * ti + li, sidx = 16 * tj + li;
      const bool rv = r < 36, sv = sidx < 36;
      d4_t acc = zero;
#pragma unroll 9
      for (int q = 0; q < NQ; ++q) {
        double A = 0.0, Bv = 0.0;
        if (rv) A = lk < 3 ? S.psi[q][r][lk] : S.dpsi[q][r];
        if (sv) Bv = S.W[q] * (lk < 3 ? S.psi[q][sidx][lk] : ph.zeta_j * S.dpsi[q][sidx]);
        acc = mfma16(A, Bv, acc);
      }
      const int r0 = 16 * ti, s0 = 16 * tj;
      scatter_tile<MODE>(S, acc, 1.0, OFF_J + r0, min(16, 36 - r0), OFF_J + s0, min(16, 36 - s0), rk, val,
                   [&](int i, int j) { return E_J + (r0 + i) * ROWS_J_LEN + NU + (s0 + j); });
    } else {
      const int t = job - 13, c = t / 6, tt = t % 6, ti = tt / 3, tj = tt % 3;
      const int c1 = c == 2 ? 0 : c + 1, c2 = c == 0 ? 2 : c - 1;
      const double b1 = S.B[c1], b2 = S.B[c2];
      const int a = 16 * ti + li, r = 16 * tj + li;
      const bool av = a < 27, rv = r < 36;
      d4_t acc = zero, accT = zero;
#pragma unroll
      for (int st = 0; st < 7; ++st) {
        const int q = 4 * st + lk;
        const double A = (av && q < NQ) ? S.q2[q][a] : 0.0;
        const double Bv = (rv && q < NQ) ? S.W[q] * (S.psi[q][r][c1] * b2 - S.psi[q][r][c2] * b1) : 0.0;
        acc = mfma16(A, Bv, acc);
        // A and B share the lane -> (index, k) layout, so swapping them gives X^T
        if (MODE == SCATTER_EBUF) accT = mfma16(Bv, A, accT);
      }
      const int a0 = 16 * ti, r0 = 16 * tj;
      const int na = min(16, 27 - a0), nr = min(16, 36 - r0);
      // uj = -gamma X   (u row c*27 + a, j column r)
      scatter_tile<MODE>(S, acc, -ph.gamma, c * 27 + a0, na, OFF_J + r0, nr, rk, val,
                   [&](int i, int j) { return (c * 27 + a0 + i) * ROWS_U_LEN + OFF_J + r0 + j; });
      // ju = +sigma X^T (j row r, u column c*27 + a)
      if (MODE == SCATTER_EBUF) {
        // accT lane holds X^T[r0 + i][a0 + j]: rows of the j block, contiguous u columns
        scatter_tile<MODE>(S, accT, ph.sigma, OFF_J + r0, nr, c * 27 + a0, na, rk, val,
                           [&](int i, int j) { return E_J + (r0 + i) * ROWS_J_LEN + c * 27 + a0 + j; });
      } else {
        // lane holds X[a0+i][r0+j] -> entry (j row r0+j, u col a0+i)
        const int j = lane & 15;
        if (j < nr && S.g[OFF_J + r0 + j] >= 0) {
          const int lr = OFF_J + r0 + j;
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int i = (lane >> 4) + 4 * rr;
            if (i >= na) continue;
            const int lc = c * 27 + a0 + i;
            if (S.g[lc] < 0) continue;
            emit<MODE>(S, rk, val, E_J + (r0 + j) * ROWS_J_LEN + lc, lr, lc, ph.sigma * acc[rr]);
          }
        }
      }
    }
  }
  // up / pu (81 x 4) and jphi / phij (36 x 8) on the vector ALUs
  for (int t = threadIdx.x; t < NU * 4 + NJ * 8; t += WG) {
    if (t < NU * 4) {
      const int row = t >> 2, k = t & 3, ca = row / 27, na = row % 27;
      if (MODE != SCATTER_EBUF && (S.g[row] < 0 || S.g[OFF_P + k] < 0)) continue;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.p1[q][k] * S.gphi[q][na][ca];
      emit<MODE>(S, rk, val, row * ROWS_U_LEN + OFF_P + k, row, OFF_P + k, acc);     // - dp div v
      emit<MODE>(S, rk, val, E_P + k * ROWS_P_LEN + row, OFF_P + k, row, acc);       // - div du q
    } else {
      const int tt = t - NU * 4, r = tt >> 3, v = tt & 7;
      if (MODE != SCATTER_EBUF && (S.g[OFF_J + r] < 0 || S.g[OFF_PHI + v] < 0)) continue;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc -= S.W[q] * S.q1[q][v] * S.dpsi[q][r];
      emit<MODE>(S, rk, val, E_J + r * ROWS_J_LEN + NU + NJ + v, OFF_J + r, OFF_PHI + v, ph.sigma * acc);  // - sigma dphi div s
      emit<MODE>(S, rk, val, E_PHI + v * ROWS_PHI_LEN + r, OFF_PHI + v, OFF_J + r, acc);                   // - div dj phi_t
    }
  }
}

// cells[i] (or i when cells == nullptr) for i < ncells; SCATTER_COLOUR needs
// a colour class: no two listed cells share a dof.  SCATTER_EBUF writes cell c's
// entries to val + c * NENT (val is then the element buffer).
template <int MODE>
__global__ __launch_bounds__(WG, 2) void k_cell_jacobian(int ncells, const int32_t *__restrict__ cells,
                                                         const double *__restrict__ xyz,
                                                         const int32_t *__restrict__ dofs, const int8_t *__restrict__ sign,
                                                         const double *__restrict__ dir, const double *__restrict__ x,
                                                         const RefTables *__restrict__ T,
                                                         const int64_t *__restrict__ rowptr,
                                                         const uint16_t *__restrict__ rank, double *__restrict__ val,
                                                         MhdPhys ph) {
  __shared__ CellLDS S;
  for (int i = blockIdx.x; i < ncells; i += gridDim.x) {
    const int c = cells ? cells[i] : i;
    stage_cell(S, c, xyz, dofs, sign, dir, x, T, rowptr, ph, ph.zeta_u != 0.0);
    if (threadIdx.x < 3) S.B[threadIdx.x] = ph.B[threadIdx.x];
    __syncthreads();
    cell_jacobian_tiles<MODE>(S, ph, rank + (int64_t)c * NENT, MODE == SCATTER_EBUF ? val + (int64_t)c * NENT : val);
    __syncthreads();
  }
}

// Element residuals into rbuf[cell][NDOF] (free rows only); k_res_gather
// then sums them per dof, so the residual does not depend on the order in
// which the cells finish.
__global__ __launch_bounds__(WG, 2) void k_cell_residual(int ncells, const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ dofs, const int8_t *__restrict__ sign,
                                                      const double *__restrict__ dir, const double *__restrict__ x,
                                                      const RefTables *__restrict__ T, double *__restrict__ rbuf,
                                                      MhdPhys ph, int mode) {
  __shared__ CellLDS S;
  for (int c = blockIdx.x; c < ncells; c += gridDim.x) {
    stage_cell(S, c, xyz, dofs, sign, dir, x, T, nullptr, ph, ph.zeta_u != 0.0);
    const int a = threadIdx.x;
    if (a < NDOF && S.g[a] >= 0) rbuf[(int64_t)c * NDOF + a] = res_entry(S, T, ph, a, mode);
    __syncthreads();
  }
}

// r[i] = sum of the element residual rows holding dof i, in increasing cell
// order (the holder lists of the row assembly)
__global__ void k_res_gather(int64_t n, const int32_t *__restrict__ crow, const int32_t *__restrict__ centry,
                             const double *__restrict__ rbuf, double *__restrict__ r) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int32_t h = crow[i]; h < crow[i + 1]; ++h) s += rbuf[centry[h]];
  r[i] = s;
}

// Per-cell integrals of the state (post-processing, src/Applications/hunt.jl:247-260):
// out[c] = (int u.u, int grad u : grad u, int j.j, int (div j)^2, int p, |K|).  The 3x3x3
// Gauss rule integrates these polynomial quantities exactly on affine cells.
__global__ __launch_bounds__(WG, 2) void k_cell_integrals(int ncells, const double *__restrict__ xyz,
                                                         const int32_t *__restrict__ dofs,
                                                         const int8_t *__restrict__ sign,
                                                         const double *__restrict__ dir,
                                                         const double *__restrict__ x,
                                                         const RefTables *__restrict__ T, double *__restrict__ out,
                                                         MhdPhys ph) {
  __shared__ CellLDS S;
  for (int c = blockIdx.x; c < ncells; c += gridDim.x) {
    stage_cell(S, c, xyz, dofs, sign, dir, x, T, nullptr, ph, false);
    const int k = threadIdx.x;
    if (k < 6) {
      double acc = 0.0;
      for (int q = 0; q < NQ; ++q) {
        double v;
        if (k == 0) v = S.u[q][0] * S.u[q][0] + S.u[q][1] * S.u[q][1] + S.u[q][2] * S.u[q][2];
        else if (k == 1) {
          v = 0.0;
          for (int i = 0; i < 9; ++i) v += S.gu[q][i] * S.gu[q][i];
        } else if (k == 2) v = S.jv[q][0] * S.jv[q][0] + S.jv[q][1] * S.jv[q][1] + S.jv[q][2] * S.jv[q][2];
        else if (k == 3) v = S.divj[q] * S.divj[q];
        else if (k == 4) v = S.p[q];
        else v = 1.0;
        acc += S.W[q] * v;
      }
      out[(int64_t)c * 6 + k] = acc;
    }
    __syncthreads();
  }
}

// Thin-wall face terms: one 64-wide workgroup per boundary face.
//
// Only the face's own four RT dofs (local 4 lf .. 4 lf + 3) have a nonzero
// normal flux on it: every other reference field is either tangential to the
// face or carries a 1-D Lagrange factor that vanishes there.  The normal flux
// s.n is taken from the reference field, s.n = sgn rt_d (+-1) / (det J |J^-T n^|)
// (the contravariant Piola identity), which is exactly zero for the other 32
// dofs, so a face adds to the rows of its own dofs only.  A boundary face dof
// belongs to one face, so every residual and Jacobian entry receives at most
// one face contribution: the sums do not depend on the launch order, on
// non-affine cells too.
struct FaceLDS {
  double W[NQF];
  double ndn[NQF];           // +-1 / (det J |J^-T n^|): s.n = sgn * rt_d * ndn
  double n[NQF][3];
  double psi[NQF][36][3];
  double gpsi[NQF][36][9];   // [i][k] = d psi_k / d x_i
  double jn[NQF], jnn[NQF];
  double xl[NJ];
  double xyz[24];
  int32_t g[NJ];
  int64_t rp[NJ];
};

__device__ void stage_face(FaceLDS &S, int f, const double *__restrict__ xyz, const int32_t *__restrict__ dofs,
                           const int8_t *__restrict__ sign, const double *__restrict__ dir,
                           const double *__restrict__ x, const FaceTables *__restrict__ F,
                           const int32_t *__restrict__ fcell, const int8_t *__restrict__ flf,
                           const int64_t *__restrict__ rowptr, int &cell, int &lf) {
  const int tid = threadIdx.x, nt = blockDim.x;
  cell = fcell[f];
  lf = flf[f];
  if (tid < 24) S.xyz[tid] = xyz[(int64_t)cell * 24 + tid];
  if (tid < NJ) {
    const int32_t g = dofs[(int64_t)cell * NDOF + OFF_J + tid];
    S.g[tid] = g;
    S.xl[tid] = g >= 0 ? x[g] : dir[-g - 1];
    S.rp[tid] = (g >= 0 && rowptr) ? rowptr[g] : -1;
  }
  __syncthreads();
  __shared__ double Jq[NQF][9], Jiq[NQF][9], detq[NQF];
  if (tid < NQF) {
    const int q = tid;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int v = 0; v < 8; ++v)
      for (int i = 0; i < 3; ++i)
        for (int l = 0; l < 3; ++l) J[i][l] += S.xyz[3 * v + i] * F->dq1[lf][q][v][l];
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double id = 1.0 / det;
    double Ji[3][3];
    Ji[0][0] = c00 * id;
    Ji[1][0] = c01 * id;
    Ji[2][0] = c02 * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    const int d = 2 - lf / 2;
    const double nh = (lf & 1) ? 1.0 : -1.0;
    double n[3], nn = 0.0;
    for (int i = 0; i < 3; ++i) {
      n[i] = Ji[d][i] * nh;
      nn += n[i] * n[i];
    }
    nn = sqrt(nn);
    for (int i = 0; i < 3; ++i) S.n[q][i] = n[i] / nn;
    S.W[q] = F->w[q] * fabs(det) * nn;
    S.ndn[q] = nh / (det * nn);
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 3; ++l) {
        Jq[q][3 * i + l] = J[i][l];
        Jiq[q][3 * i + l] = Ji[i][l];
      }
    detq[q] = det;
  }
  __syncthreads();
  for (int idx = tid; idx < NQF * 36; idx += nt) {
    const int q = idx / 36, r = idx % 36;
    const double sgn = r < 24 ? (sign ? (double)sign[(int64_t)cell * 6 + (r >> 2)] : 1.0) : 1.0;
    const double id = sgn / detq[q];
    const double *J = Jq[q], *Ji = Jiq[q];
    const double *rt = F->rt[lf][q][r];
    for (int k = 0; k < 3; ++k) S.psi[q][r][k] = id * (J[3 * k + 0] * rt[0] + J[3 * k + 1] * rt[1] + J[3 * k + 2] * rt[2]);
    // grad psi [i][k] = id * sum_l J[k][l] sum_m drt[l][m] Ji[m][i]  (affine push-forward)
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
        for (int l = 0; l < 3; ++l) {
          double dl = 0.0;
          for (int m = 0; m < 3; ++m) dl += F->drt[lf][q][r][l][m] * Ji[3 * m + i];
          acc += J[3 * k + l] * dl;
        }
        S.gpsi[q][r][3 * i + k] = id * acc;
      }
  }
  __syncthreads();
  if (tid < NQF) {
    const int q = tid;
    const double *n = S.n[q];
    const int d = 2 - lf / 2;
    double jn = 0.0, jnn = 0.0;
    for (int r = 4 * lf; r < 4 * lf + 4; ++r) {
      const double sgn = sign ? (double)sign[(int64_t)cell * 6 + lf] : 1.0;
      jn += S.xl[r] * sgn * F->rt[lf][q][r][d] * S.ndn[q];
    }
    for (int r = 0; r < 36; ++r) {
      const double *gs = S.gpsi[q][r];
      double t = 0.0;
      for (int i = 0; i < 3; ++i) t += n[i] * (gs[3 * i + 0] * n[0] + gs[3 * i + 1] * n[1] + gs[3 * i + 2] * n[2]);
      jnn += S.xl[r] * t;
    }
    S.jn[q] = jn;
    S.jnn[q] = jnn;
  }
  __syncthreads();
}

__global__ __launch_bounds__(64) void k_face_residual(int nfaces, const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ dofs, const int8_t *__restrict__ sign,
                                                      const double *__restrict__ dir, const double *__restrict__ x,
                                                      const FaceTables *__restrict__ F, const int32_t *__restrict__ fcell,
                                                      const int8_t *__restrict__ flf, const double *__restrict__ ftau,
                                                      const double *__restrict__ fcw, const double *__restrict__ fjw,
                                                      double *__restrict__ r, int mode) {
  __shared__ FaceLDS S;
  for (int f = blockIdx.x; f < nfaces; f += gridDim.x) {
    int cell, lf;
    stage_face(S, f, xyz, dofs, sign, dir, x, F, fcell, flf, nullptr, cell, lf);
    const int a = 4 * lf + threadIdx.x;
    if (threadIdx.x < 4 && S.g[a] >= 0) {
      const double tau = ftau[f], cw = fcw[f], jw = fjw[f];
      const int d = 2 - lf / 2;
      const double sgn = sign ? (double)sign[(int64_t)cell * 6 + lf] : 1.0;
      double acc = 0.0;
      for (int q = 0; q < NQF; ++q) {
        const double sn = sgn * F->rt[lf][q][a][d] * S.ndn[q];
        acc += mode == 2 ? S.W[q] * tau * sn * S.jnn[q] : S.W[q] * tau * sn * (S.jn[q] - jw + cw * S.jnn[q]);
      }
      atomicAdd(&r[S.g[a]], acc);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_face_jacobian(int nfaces, const double *__restrict__ xyz,
                                                      const int32_t *__restrict__ dofs, const int8_t *__restrict__ sign,
                                                      const double *__restrict__ dir, const double *__restrict__ x,
                                                      const FaceTables *__restrict__ F, const int32_t *__restrict__ fcell,
                                                      const int8_t *__restrict__ flf, const double *__restrict__ ftau,
                                                      const double *__restrict__ fcw, const int64_t *__restrict__ rowptr,
                                                      const uint16_t *__restrict__ rank, double *__restrict__ val) {
  __shared__ FaceLDS S;
  for (int f = blockIdx.x; f < nfaces; f += gridDim.x) {
    int cell, lf;
    stage_face(S, f, xyz, dofs, sign, dir, x, F, fcell, flf, rowptr, cell, lf);
    const double tau = ftau[f], cw = fcw[f];
    const uint16_t *rk = rank + (int64_t)cell * NENT;
    const int dax = 2 - lf / 2;
    const double sgn = sign ? (double)sign[(int64_t)cell * 6 + lf] : 1.0;
    for (int idx = threadIdx.x; idx < 4 * NJ; idx += blockDim.x) {
      const int a = 4 * lf + idx / NJ, b = idx % NJ;
      if (S.g[a] < 0 || S.g[b] < 0) continue;
      const bool bface = (b >> 2) == lf;
      double acc = 0.0;
      for (int q = 0; q < NQF; ++q) {
        const double *n = S.n[q], *gd = S.gpsi[q][b];
        const double sn = sgn * F->rt[lf][q][a][dax] * S.ndn[q];
        const double dn = bface ? sgn * F->rt[lf][q][b][dax] * S.ndn[q] : 0.0;
        double dnn = 0.0;
        for (int i = 0; i < 3; ++i) dnn += n[i] * (gd[3 * i + 0] * n[0] + gd[3 * i + 1] * n[1] + gd[3 * i + 2] * n[2]);
        acc += S.W[q] * tau * sn * (dn + cw * dnn);
      }
      // jj block entry of j row a, j column b inside the cell's rank map
      const int e = E_J + a * ROWS_J_LEN + NU + b;
      atomicAdd(&val[S.rp[a] + rk[e]], acc);
    }
    __syncthreads();
  }
}

// Canonical entry range [start, start + len) of local row a of a cell.
__device__ inline void row_segment(int a, int &start, int &len) {
  if (a < OFF_P) {
    start = a * ROWS_U_LEN;
    len = ROWS_U_LEN;
  } else if (a < OFF_J) {
    start = E_P + (a - OFF_P) * ROWS_P_LEN;
    len = ROWS_P_LEN;
  } else if (a < OFF_PHI) {
    start = E_J + (a - OFF_J) * ROWS_J_LEN;
    len = ROWS_J_LEN;
  } else {
    start = E_PHI + (a - OFF_PHI) * ROWS_PHI_LEN;
    len = ROWS_PHI_LEN;
  }
}

// Row assembly from the element buffer: one 64-lane wave per CSR row.  The row
// is accumulated in LDS -- every (cell, local row) holding the row's dof adds
// its contiguous element-buffer segment at the entries' ranks, in increasing
// cell order -- and then written once, coalesced.  Every entry of the row is
// written, so val needs no clearing.  A cell holds a dof at most once, so the
// ranks of one segment are distinct and its lanes never collide in LDS.
// All holders of a dof hold it as the same field, so the segments of a row
// share one length (<= 125 = 2 x 64 lanes); up to RA_BATCH segments are loaded
// into registers together before they are added in order.
constexpr int RA_WAVES = 4;
constexpr int RA_BATCH = 8;
__global__ __launch_bounds__(64 * RA_WAVES) void k_row_assemble(int64_t n, const int64_t *__restrict__ rowptr,
                                                                const int32_t *__restrict__ crow,
                                                                const int32_t *__restrict__ centry,
                                                                const double *__restrict__ ebuf,
                                                                const uint16_t *__restrict__ rank,
                                                                double *__restrict__ val, int maxrow) {
  extern __shared__ double ra_lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double *buf = ra_lds + (size_t)wave * maxrow;
  for (int64_t row = (int64_t)blockIdx.x * RA_WAVES + wave; row < n; row += (int64_t)gridDim.x * RA_WAVES) {
    const int64_t b = rowptr[row];
    const int len = (int)(rowptr[row + 1] - b);
    const int p0 = crow[row], p1 = crow[row + 1];
    for (int t = lane; t < len; t += 64) buf[t] = 0.0;
    __builtin_amdgcn_wave_barrier();
    for (int pb = p0; pb < p1; pb += RA_BATCH) {
      double v0[RA_BATCH], v1[RA_BATCH];
      uint16_t k0[RA_BATCH], k1[RA_BATCH];
#pragma unroll
      for (int u = 0; u < RA_BATCH; ++u) {
        k0[u] = k1[u] = RANK_NONE;
        v0[u] = v1[u] = 0.0;
        if (pb + u < p1) {
          const int32_t ce = centry[pb + u];
          const int cell = ce / NDOF, a = ce - cell * NDOF;
          int start, slen;
          row_segment(a, start, slen);
          const int64_t base = (int64_t)cell * NENT + start;
          if (lane < slen) {
            k0[u] = rank[base + lane];
            v0[u] = ebuf[base + lane];
          }
          if (lane + 64 < slen) {
            k1[u] = rank[base + lane + 64];
            v1[u] = ebuf[base + lane + 64];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < RA_BATCH; ++u) {
        if (k0[u] != RANK_NONE) buf[k0[u]] += v0[u];
        if (k1[u] != RANK_NONE) buf[k1[u]] += v1[u];
        __builtin_amdgcn_wave_barrier();
      }
    }
    for (int t = lane; t < len; t += 64) val[b + t] = buf[t];
    __builtin_amdgcn_wave_barrier();
  }
}

// Rank of every touched cell entry inside its CSR row (binary search).
__global__ void k_rank_map(int ncells, const int32_t *__restrict__ dofs, const int64_t *__restrict__ rowptr,
                           const int32_t *__restrict__ colind, uint16_t *__restrict__ rank, int *__restrict__ bad) {
  const int64_t total = (int64_t)ncells * NENT;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t / NENT), e = (int)(t % NENT);
    int a, b;
    decode_entry(e, a, b);
    const int32_t ga = dofs[(int64_t)c * NDOF + a], gb = dofs[(int64_t)c * NDOF + b];
    uint16_t out = RANK_NONE;
    if (ga >= 0 && gb >= 0) {
      int64_t lo = rowptr[ga], hi = rowptr[ga + 1];
      const int64_t base = lo;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (colind[mid] < gb) lo = mid + 1;
        else hi = mid;
      }
      if (lo < rowptr[ga + 1] && colind[lo] == gb && lo - base < RANK_NONE) out = (uint16_t)(lo - base);
      else atomicAdd(bad, 1);
    }
    rank[t] = out;
  }
}

// tperm[p] = CSR position of the transposed entry of p.
__global__ void k_tperm(int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colind,
                        int64_t *__restrict__ tperm, int *__restrict__ bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int32_t j = colind[p];
      int64_t lo = rowptr[j], hi = rowptr[j + 1];
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (colind[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      if (lo < rowptr[j + 1] && colind[lo] == i) tperm[p] = lo;
      else {
        tperm[p] = -1;
        atomicAdd(bad, 1);
      }
    }
  }
}

// CSR SpMV, one 64-lane wavefront per group of rows, `LPR` lanes per row.
template <int LPR>
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int64_t *__restrict__ rowptr,
                                              const int32_t *__restrict__ colind, const double *__restrict__ val,
                                              const double *__restrict__ x, double *__restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t row = gid / LPR;
  const int lane = threadIdx.x % LPR;
  double acc = 0.0;
  if (row < n) {
    // four independent accumulators: four loads in flight per lane
    const int64_t e = rowptr[row + 1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int64_t p = rowptr[row] + lane;
    for (; p + 3 * LPR < e; p += 4 * LPR) {
      a0 += val[p] * x[colind[p]];
      a1 += val[p + LPR] * x[colind[p + LPR]];
      a2 += val[p + 2 * LPR] * x[colind[p + 2 * LPR]];
      a3 += val[p + 3 * LPR] * x[colind[p + 3 * LPR]];
    }
    for (; p < e; p += LPR) a0 += val[p] * x[colind[p]];
    acc = (a0 + a1) + (a2 + a3);
  }
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LPR);
  if (row < n && lane == 0) y[row] = acc;
}

// Skeleton stabilisation (src/weakforms.jl:685-688):
//   a_Lambda(u, v) = int_Lambda (1/2) mu h^2 [grad u] : [grad v],  h^2 = |face|.
// [grad u] : [grad v] = sum_c sum_i [d_i u_c][d_i v_c]: the three velocity
// components share one scalar face matrix over the face's 54 node slots,
//   K_f[a][b] = (mu/2) |f| sum_q W_q G_q[a] . G_q[b],
// G_q[a] the jump of the physical gradient of node a's Q2 function at face
// point q (first cell's minus the second cell's, at the coinciding points),
// W_q = w_q |det J| |J^-T n^| on the first cell.  a_Lambda is linear and
// depends on the geometry only, so K_f is computed once per context.
__global__ __launch_bounds__(64) void k_skel_matrix(int nf, const double *__restrict__ xyz,
                                                    const int32_t *__restrict__ fcell,
                                                    const int8_t *__restrict__ flf, const int8_t *__restrict__ fperm,
                                                    const uint8_t *__restrict__ mslot, const double *__restrict__ mu,
                                                    const FaceTables *__restrict__ F, double *__restrict__ K) {
  __shared__ double X[2][24];
  __shared__ double Jis[2][NQF][9];      // (J^-1)[l][i] of each side at the first side's point q
  __shared__ double W[NQF];
  __shared__ double G[NQF][SKEL_SLOTS][3];
  const int tid = threadIdx.x;
  for (int f = blockIdx.x; f < nf; f += gridDim.x) {
    const int c0 = fcell[2 * f], c1 = fcell[2 * f + 1];
    const int lf0 = flf[2 * f], lf1 = flf[2 * f + 1];
    if (tid < 48) X[tid / 24][tid % 24] = xyz[(int64_t)(tid < 24 ? c0 : c1) * 24 + tid % 24];
    for (int e = tid; e < NQF * SKEL_SLOTS * 3; e += 64) (&G[0][0][0])[e] = 0.0;
    __syncthreads();
    if (tid < 2 * NQF) {
      const int s = tid / NQF, q = tid % NQF;
      const int lf = s ? lf1 : lf0, qs = s ? fperm[(int64_t)NQF * f + q] : q;
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int v = 0; v < 8; ++v)
        for (int i = 0; i < 3; ++i)
          for (int l = 0; l < 3; ++l) J[i][l] += X[s][3 * v + i] * F->dq1[lf][qs][v][l];
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      double *Ji = Jis[s][q];
      Ji[0] = c00 * id;
      Ji[3] = c01 * id;
      Ji[6] = c02 * id;
      Ji[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      Ji[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      Ji[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      Ji[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      if (s == 0) {
        const int d = 2 - lf / 2;
        const double nn = sqrt(Ji[3 * d] * Ji[3 * d] + Ji[3 * d + 1] * Ji[3 * d + 1] + Ji[3 * d + 2] * Ji[3 * d + 2]);
        W[q] = F->w[q] * fabs(det) * nn;
      }
    }
    __syncthreads();
    // first cell: slots 0..26 (distinct); then the second cell subtracts its
    // gradients at the slots mslot (distinct among themselves)
    for (int s = 0; s < 2; ++s) {
      for (int e = tid; e < NQF * 27; e += 64) {
        const int q = e / 27, a = e % 27;
        const int lf = s ? lf1 : lf0, qs = s ? fperm[(int64_t)NQF * f + q] : q;
        const int slot = s ? mslot[(int64_t)27 * f + a] : a;
        const double *g = F->dq2[lf][qs][a];
        const double *Ji = Jis[s][q];
        const double sg = s ? -1.0 : 1.0;
        for (int i = 0; i < 3; ++i) G[q][slot][i] += sg * (Ji[i] * g[0] + Ji[3 + i] * g[1] + Ji[6 + i] * g[2]);
      }
      __syncthreads();
    }
    double area = 0.0;
    for (int q = 0; q < NQF; ++q) area += W[q];
    const double coef = 0.5 * mu[f] * area;
    double *Kf = K + (int64_t)f * SKEL_SLOTS * SKEL_SLOTS;
    for (int e = tid; e < SKEL_SLOTS * SKEL_SLOTS; e += 64) {
      const int a = e / SKEL_SLOTS, b = e % SKEL_SLOTS;
      double acc = 0.0;
      for (int q = 0; q < NQF; ++q)
        acc += W[q] * (G[q][a][0] * G[q][b][0] + G[q][a][1] * G[q][b][1] + G[q][a][2] * G[q][b][2]);
      Kf[e] = coef * acc;
    }
    __syncthreads();
  }
}

// Skeleton residual, in two steps.  k_skel_face_product: one wave per face,
// lane a < 54 forms (K_f x_c)[a] for the three components c from the slot
// values x_c (Dirichlet values for non-free slots, 0 for unused ones), reading
// K_f by rows b (K_f is symmetric bit for bit: its entries are sums of the
// same products in the same order), so the loads are coalesced.
// k_skel_res_gather: one thread per row adds the face products of the row's
// (face, component, slot) entries, whose index is the product's own index, in
// increasing face order: the sums do not depend on the launch.
__global__ __launch_bounds__(64) void k_skel_face_product(int nf, const int32_t *__restrict__ sdofs,
                                                          const double *__restrict__ K, const double *__restrict__ x,
                                                          const double *__restrict__ dir, double *__restrict__ fres) {
  __shared__ double xs[SKEL_DOFS];
  const int lane = threadIdx.x;
  for (int f = blockIdx.x; f < nf; f += gridDim.x) {
    for (int e = lane; e < SKEL_DOFS; e += 64) {
      const int32_t g = sdofs[(int64_t)f * SKEL_DOFS + e];
      xs[e] = g == SKEL_NONE ? 0.0 : (g >= 0 ? x[g] : dir[-(int64_t)g - 1]);
    }
    __syncthreads();
    if (lane < SKEL_SLOTS) {
      const double *Kf = K + (int64_t)f * SKEL_SLOTS * SKEL_SLOTS;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0;
      for (int b = 0; b < SKEL_SLOTS; ++b) {
        const double k = Kf[b * SKEL_SLOTS + lane];
        a0 += k * xs[b];
        a1 += k * xs[SKEL_SLOTS + b];
        a2 += k * xs[2 * SKEL_SLOTS + b];
      }
      double *o = fres + (int64_t)f * SKEL_DOFS;
      o[lane] = a0;
      o[SKEL_SLOTS + lane] = a1;
      o[2 * SKEL_SLOTS + lane] = a2;
    }
    __syncthreads();
  }
}

__global__ void k_skel_res_gather(int64_t n, const int32_t *__restrict__ srow, const int32_t *__restrict__ sentry,
                                  const double *__restrict__ fres, double *__restrict__ r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int p0 = srow[i], p1 = srow[i + 1];
    if (p0 == p1) continue;
    double acc = 0.0;
    for (int p = p0; p < p1; ++p) acc += fres[sentry[p]];
    r[i] += acc;
  }
}

// Rank of every slot dof of an entry's face (same component) in the entry's
// CSR row, found once at creation by binary search (RANK_NONE for non-free
// slots).  One wave per row.
__global__ __launch_bounds__(64) void k_skel_rank(int64_t n, const int64_t *__restrict__ rowptr,
                                                  const int32_t *__restrict__ colind, const int32_t *__restrict__ srow,
                                                  const int32_t *__restrict__ sentry,
                                                  const int32_t *__restrict__ sdofs, uint16_t *__restrict__ rank) {
  const int lane = threadIdx.x;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int p0 = srow[i], p1 = srow[i + 1];
    if (p0 == p1) continue;
    const int64_t rb = rowptr[i];
    const int len = (int)(rowptr[i + 1] - rb);
    for (int p = p0; p < p1; ++p) {
      if (lane >= SKEL_SLOTS) continue;
      const int32_t e = sentry[p];
      const int f = e / SKEL_DOFS, comp = (e % SKEL_DOFS) / SKEL_SLOTS;
      const int32_t g = sdofs[(int64_t)f * SKEL_DOFS + comp * SKEL_SLOTS + lane];
      uint16_t rk = RANK_NONE;
      if (g >= 0) {
        int lo = 0, hi = len;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (colind[rb + mid] < g) lo = mid + 1;
          else hi = mid;
        }
        if (lo < len && colind[rb + lo] == g) rk = (uint16_t)lo;
      }
      rank[(int64_t)p * SKEL_SLOTS + lane] = rk;
    }
  }
}

// Skeleton rows of the Jacobian, added to the assembled CSR values: one
// 64-lane wave per row; the row is accumulated in LDS (lane b adds K_f[slot][b]
// at the rank of slot b's dof, from the rank table -- the slot dofs of one
// entry are distinct, so lanes never collide), entries in increasing face
// order, then added to the row once.
__global__ __launch_bounds__(64) void k_skel_jacobian(int64_t n, const int64_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ srow,
                                                      const int32_t *__restrict__ sentry,
                                                      const uint16_t *__restrict__ rank,
                                                      const double *__restrict__ K, double *__restrict__ val) {
  extern __shared__ double sk_lds[];
  const int lane = threadIdx.x;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int p0 = srow[i], p1 = srow[i + 1];
    if (p0 == p1) continue;
    const int64_t rb = rowptr[i];
    const int len = (int)(rowptr[i + 1] - rb);
    for (int t = lane; t < len; t += 64) sk_lds[t] = 0.0;
    __builtin_amdgcn_wave_barrier();
    for (int p = p0; p < p1; ++p) {
      if (lane < SKEL_SLOTS) {
        const uint16_t rk = rank[(int64_t)p * SKEL_SLOTS + lane];
        if (rk != RANK_NONE) {
          const int32_t e = sentry[p];
          const int f = e / SKEL_DOFS, a = (e % SKEL_DOFS) % SKEL_SLOTS;
          sk_lds[rk] += K[((int64_t)f * SKEL_SLOTS + a) * SKEL_SLOTS + lane];
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    for (int t = lane; t < len; t += 64) val[rb + t] += sk_lds[t];
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace

int mhd_launch_jacobian(mhd_ctx *c, const double *d_x, double *d_val, hipStream_t s) {
  // a cell holding one dof twice (a periodic direction one cell thick) needs
  // the atomic scatter: its own entries may coincide
  int mode = c->dup_dofs ? 0 : c->assembly_mode;
  if (mode == 2 && sizeof(double) * (size_t)RA_WAVES * c->max_row > 65536) mode = 1;  // row too long for LDS
  if (mode == 2 && c->ncells > 0) {
    if (!c->d_ebuf) MHD_HIP(hipMalloc(&c->d_ebuf, sizeof(double) * (size_t)c->ncells * NENT));
    hipLaunchKernelGGL((k_cell_jacobian<SCATTER_EBUF>), dim3(c->ncells), dim3(WG), 0, s, c->ncells,
                       (const int32_t *)nullptr, c->d_xyz, c->d_dofs, c->d_sign, c->d_dir, d_x, c->d_ref, c->d_rowptr,
                       c->d_rank, c->d_ebuf, c->phys);
    MHD_HIP(hipGetLastError());
    if (c->n > 0) {
      const int64_t blocks = std::min<int64_t>((c->n + RA_WAVES - 1) / RA_WAVES, 1 << 20);
      const size_t lds = sizeof(double) * (size_t)RA_WAVES * c->max_row;
      hipLaunchKernelGGL(k_row_assemble, dim3((unsigned)blocks), dim3(64 * RA_WAVES), lds, s, c->n, c->d_rowptr,
                         c->d_crow, c->d_centry, c->d_ebuf, c->d_rank, d_val, c->max_row);
      MHD_HIP(hipGetLastError());
    }
  } else {
    MHD_HIP(hipMemsetAsync(d_val, 0, sizeof(double) * c->nnz, s));
  }
  if (mode == 2 || c->ncells == 0) {
  } else if (mode == 0 || c->color_ptr.size() < 2) {
    hipLaunchKernelGGL((k_cell_jacobian<SCATTER_ATOMIC>), dim3(c->ncells), dim3(WG), 0, s, c->ncells,
                       (const int32_t *)nullptr, c->d_xyz, c->d_dofs, c->d_sign, c->d_dir, d_x, c->d_ref, c->d_rowptr,
                       c->d_rank, d_val, c->phys);
    MHD_HIP(hipGetLastError());
  } else {
    for (size_t k = 0; k + 1 < c->color_ptr.size(); ++k) {
      const int n = c->color_ptr[k + 1] - c->color_ptr[k];
      if (n == 0) continue;
      hipLaunchKernelGGL((k_cell_jacobian<SCATTER_COLOUR>), dim3(n), dim3(WG), 0, s, n,
                         c->d_color_cells + c->color_ptr[k], c->d_xyz, c->d_dofs, c->d_sign, c->d_dir, d_x, c->d_ref,
                         c->d_rowptr, c->d_rank, d_val, c->phys);
      MHD_HIP(hipGetLastError());
    }
  }
  if (c->nfaces > 0) {
    hipLaunchKernelGGL(k_face_jacobian, dim3(c->nfaces), dim3(64), 0, s, c->nfaces, c->d_xyz, c->d_dofs, c->d_sign,
                       c->d_dir, d_x, c->d_fref, c->d_fcell, c->d_flf, c->d_ftau, c->d_fcw, c->d_rowptr, c->d_rank,
                       d_val);
    MHD_HIP(hipGetLastError());
  }
  return mhd_launch_skeleton_jacobian(c, d_val, s);
}

int mhd_build_skeleton(mhd_ctx *c, const int32_t *d_fcell, const int8_t *d_flf, const int8_t *d_fperm,
                       const uint8_t *d_mslot, const double *d_mu) {
  if (c->nskel <= 0) return 0;
  MHD_HIP(hipMalloc(&c->d_skel_K, sizeof(double) * (size_t)c->nskel * SKEL_SLOTS * SKEL_SLOTS));
  MHD_HIP(hipMalloc(&c->d_skel_res, sizeof(double) * (size_t)c->nskel * SKEL_DOFS));
  MHD_HIP(hipMalloc(&c->d_skel_rank, sizeof(uint16_t) * (size_t)std::max<int64_t>(c->nsentry, 1) * SKEL_SLOTS));
  const int blocks = std::min(c->nskel, 1 << 20);
  hipLaunchKernelGGL(k_skel_matrix, dim3(blocks), dim3(64), 0, c->stream, c->nskel, c->d_xyz, d_fcell, d_flf,
                     d_fperm, d_mslot, d_mu, c->d_fref, c->d_skel_K);
  MHD_HIP(hipGetLastError());
  if (c->n > 0) {
    hipLaunchKernelGGL(k_skel_rank, dim3((unsigned)std::min<int64_t>(c->n, 1 << 20)), dim3(64), 0, c->stream, c->n,
                       c->d_rowptr, c->d_colind, c->d_srow, c->d_sentry, c->d_skel_dofs, c->d_skel_rank);
    MHD_HIP(hipGetLastError());
  }
  return 0;
}

int mhd_launch_skeleton_residual(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s) {
  if (c->nskel <= 0 || c->n <= 0) return 0;
  hipLaunchKernelGGL(k_skel_face_product, dim3((unsigned)std::min(c->nskel, 1 << 20)), dim3(64), 0, s, c->nskel,
                     c->d_skel_dofs, c->d_skel_K, d_x, c->d_dir, c->d_skel_res);
  const int64_t blocks = std::min<int64_t>((c->n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_skel_res_gather, dim3((unsigned)blocks), dim3(256), 0, s, c->n, c->d_srow, c->d_sentry,
                     c->d_skel_res, d_r);
  MHD_HIP(hipGetLastError());
  return 0;
}

int mhd_launch_skeleton_jacobian(mhd_ctx *c, double *d_val, hipStream_t s) {
  if (c->nskel <= 0 || c->n <= 0) return 0;
  const size_t lds = sizeof(double) * (size_t)c->max_row;
  if (lds > 65536) {
    c->err = "skeleton Jacobian: a CSR row longer than 8192 entries";
    return -1;
  }
  const int64_t blocks = std::min<int64_t>(c->n, 1 << 20);
  hipLaunchKernelGGL(k_skel_jacobian, dim3((unsigned)blocks), dim3(64), lds, s, c->n, c->d_rowptr, c->d_srow,
                     c->d_sentry, c->d_skel_rank, c->d_skel_K, d_val);
  MHD_HIP(hipGetLastError());
  return 0;
}

int mhd_launch_residual_mode(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s, int mode) {
  if (c->ncells > 0 && mode != 2) {
    if (!c->d_rbuf) MHD_HIP(hipMalloc(&c->d_rbuf, sizeof(double) * (size_t)c->ncells * NDOF));
    hipLaunchKernelGGL(k_cell_residual, dim3(c->ncells), dim3(WG), 0, s, c->ncells, c->d_xyz, c->d_dofs, c->d_sign,
                       c->d_dir, d_x, c->d_ref, c->d_rbuf, c->phys, mode);
    hipLaunchKernelGGL(k_res_gather, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, s, c->n, c->d_crow,
                       c->d_centry, c->d_rbuf, d_r);
    MHD_HIP(hipGetLastError());
  } else {
    MHD_HIP(hipMemsetAsync(d_r, 0, sizeof(double) * c->n, s));
  }
  if (c->nfaces > 0 && mode != 1) {
    hipLaunchKernelGGL(k_face_residual, dim3(c->nfaces), dim3(64), 0, s, c->nfaces, c->d_xyz, c->d_dofs, c->d_sign,
                       c->d_dir, d_x, c->d_fref, c->d_fcell, c->d_flf, c->d_ftau, c->d_fcw, c->d_fjw, d_r, mode);
    MHD_HIP(hipGetLastError());
  }
  // the skeleton term depends on neither gamma nor cw (modes 1, 2: dR/dtheta)
  return mode == 0 ? mhd_launch_skeleton_residual(c, d_x, d_r, s) : 0;
}

int mhd_launch_residual(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s) {
  return mhd_launch_residual_mode(c, d_x, d_r, s, 0);
}

// ---------------------------------------------------------------------------
// Allocation-free residual.  Its element residuals live in `scratch`, device
// memory of the code object: every GPU the library runs on holds its own
// copy (NOALLOC_DOUBLES doubles = 128 MiB) until the process ends.  Host code
// never passes `scratch` itself to a kernel (on the host it only names the
// symbol): the device address of each GPU's copy is looked up once per device
// with hipGetSymbolAddress while that device is current, and that pointer is
// what the launches take.  Every call on one GPU shares that GPU's copy, so
// calls that overlap on one GPU race.  k_cell_residual writes every element
// residual entry that k_res_gather reads, so nothing is cleared between calls.

namespace {
constexpr int64_t NOALLOC_DOUBLES = int64_t(1) << 24;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

__global__ void k_fill(int64_t n, double v, double *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}
}  // namespace

int mhd_noalloc_init(mhd_ctx *c) {
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES) {
    c->err = "device index beyond the allocation-free workspace table";
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void *p = nullptr;
    MHD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[c->device] = static_cast<double *>(p);
  }
  c->d_scratch = g_scratch_ptr[c->device];
  return 0;
}

int mhd_launch_residual_noalloc(mhd_ctx *c, const double *d_x, double *d_r, hipStream_t s) {
  if ((int64_t)c->ncells * NDOF > NOALLOC_DOUBLES) {
    c->err = "mhd_residual_d_noalloc: " + std::to_string(c->ncells) + " cells exceed the allocation-free workspace (" +
             std::to_string(NOALLOC_DOUBLES / NDOF) + " cells at most)";
    return -1;
  }
  if (!c->d_scratch) {
    c->err = "mhd_residual_d_noalloc: workspace address not set up for this device";
    return -2;
  }
  if (c->ncells > 0) {
    hipLaunchKernelGGL(k_cell_residual, dim3(c->ncells), dim3(WG), 0, s, c->ncells, c->d_xyz, c->d_dofs, c->d_sign,
                       c->d_dir, d_x, c->d_ref, c->d_scratch, c->phys, 0);
    hipLaunchKernelGGL(k_res_gather, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, s, c->n, c->d_crow,
                       c->d_centry, c->d_scratch, d_r);
  } else if (c->n > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)std::min<int64_t>((c->n + 255) / 256, 65536)), dim3(256), 0, s, c->n,
                       0.0, d_r);
  }
  MHD_HIP(hipGetLastError());
  if (c->nfaces > 0) {
    hipLaunchKernelGGL(k_face_residual, dim3(c->nfaces), dim3(64), 0, s, c->nfaces, c->d_xyz, c->d_dofs, c->d_sign,
                       c->d_dir, d_x, c->d_fref, c->d_fcell, c->d_flf, c->d_ftau, c->d_fcw, c->d_fjw, d_r, 0);
    MHD_HIP(hipGetLastError());
  }
  // the skeleton face matrices and face products are context memory made at creation
  return mhd_launch_skeleton_residual(c, d_x, d_r, s);
}

// val_t[p] = val[tperm[p]]: the CSR values of A^T on the (symmetric) pattern
__global__ void k_gather_t(int64_t nnz, const int64_t *__restrict__ tperm, const double *__restrict__ val,
                           double *__restrict__ val_t) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x)
    val_t[p] = val[tperm[p]];
}

int mhd_launch_transpose_values(mhd_ctx *c, const double *d_val, double *d_val_t, hipStream_t s) {
  int rc;
  if (!c->d_tperm) {
    MHD_HIP(hipMalloc(&c->d_tperm, sizeof(int64_t) * std::max<int64_t>(c->nnz, 1)));
    if ((rc = mhd_build_tperm(c))) return rc;
  }
  const int grid = (int)std::min<int64_t>((c->nnz + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(k_gather_t, dim3(grid), dim3(256), 0, s, c->nnz, c->d_tperm, d_val, d_val_t);
  MHD_HIP(hipGetLastError());
  return 0;
}

int mhd_launch_cell_integrals(mhd_ctx *c, const double *d_x, double *d_out, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_integrals, dim3(c->ncells), dim3(WG), 0, s, c->ncells, c->d_xyz, c->d_dofs, c->d_sign,
                     c->d_dir, d_x, c->d_ref, d_out, c->phys);
  MHD_HIP(hipGetLastError());
  return 0;
}

int mhd_build_rank_map(mhd_ctx *c) {
  int *d_bad = nullptr;
  MHD_HIP(hipMalloc(&d_bad, sizeof(int)));
  MHD_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), c->stream));
  const int64_t total = (int64_t)c->ncells * NENT;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_rank_map, dim3(grid), dim3(256), 0, c->stream, c->ncells, c->d_dofs, c->d_rowptr, c->d_colind,
                     c->d_rank, d_bad);
  MHD_HIP(hipGetLastError());
  int bad = 0;
  MHD_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  MHD_HIP(hipStreamSynchronize(c->stream));
  MHD_HIP(hipFree(d_bad));
  if (bad) {
    c->err = "rank map: " + std::to_string(bad) + " entries missing from the pattern";
    return -3;
  }
  return 0;
}

int mhd_build_tperm(mhd_ctx *c) {
  int *d_bad = nullptr;
  MHD_HIP(hipMalloc(&d_bad, sizeof(int)));
  MHD_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), c->stream));
  const int grid = (int)std::min<int64_t>((c->n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_tperm, dim3(grid), dim3(256), 0, c->stream, c->n, c->d_rowptr, c->d_colind, c->d_tperm, d_bad);
  MHD_HIP(hipGetLastError());
  int bad = 0;
  MHD_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  MHD_HIP(hipStreamSynchronize(c->stream));
  MHD_HIP(hipFree(d_bad));
  if (bad) {
    c->err = "pattern is not structurally symmetric (" + std::to_string(bad) + " entries)";
    return -3;
  }
  return 0;
}

int mhd_launch_spmv(mhd_ctx *c, const double *d_val, const double *d_x, double *d_y, hipStream_t s) {
  constexpr int LPR = 16;
  const int64_t threads = c->n * LPR;
  const int grid = (int)((threads + 255) / 256);
  hipLaunchKernelGGL(k_spmv<LPR>, dim3(grid), dim3(256), 0, s, c->n, c->d_rowptr, c->d_colind, d_val, d_x, d_y);
  MHD_HIP(hipGetLastError());
  return 0;
}
