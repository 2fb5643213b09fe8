/*
 * mhd_oracle.c -- CPU FP64 restatement of GridapMHD's H1-HDiv residual and
 * Newton Jacobian, cell by cell.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product library links or calls this file; only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load it, as the checker.
 *
 * What it restates (reference files, read as text):
 *   src/weakforms.jl:39-43    setup_variable (u, p, j, phi, grad u, div u, grad j, div j)
 *   src/weakforms.jl:255-281  res_fluid_h1_hdiv
 *   src/weakforms.jl:283-312  jac_fluid_h1_hdiv (Picard / Newton convection)
 *   src/weakforms.jl:670      conv(u, grad u) = (grad u)^T . u
 *   src/weakforms.jl:672-681  local projection Pi_p used by the zeta_u AL term
 *   src/weakforms.jl:721-733  res/jac_thin_wall (the residual reads x, not the
 *                             undefined dx of weakforms.jl:722-724; the Jacobian
 *                             at :728-733 is its exact derivative)
 *   src/parameters.jl:380-389 quadrature degree q = 5 -> 3x3x3 Gauss cells, 3x3 faces
 *
 * Discrete spaces (src/parameters.jl:436-441, 517-526): Q2 vector Lagrangian u,
 * discontinuous P1 pressure, Raviart-Thomas order 1 current (36 dofs / hex),
 * discontinuous Q1 potential.  Local dof order of one cell (129 dofs):
 *   u   [0,81)    l = c*27 + a, a = i + 3j + 9k the Q2 node at (i/2, j/2, k/2)
 *   p   [81,85)   1 - x - y - z, x, y, z   (reference coordinates)
 *   j   [85,121)  24 face dofs (face f = 0..5 in order z-, z+, y-, y+, x-, x+,
 *                 4 each), then 12 interior dofs (component d = 0..2, 4 each)
 *   phi [121,129) trilinear at vertex v = i + 2j + 4k
 * The RT basis is nodal: dof = normal component (+axis direction) of the
 * reference field at the 2x2 Gauss points of its face (interior: at the
 * mid-plane); component d of a dof of axis d is L_m(x_d) G_g1(x_t1) G_g2(x_t2)
 * with L_m the 1-D Q2 Lagrange polynomial on {0, 1/2, 1} and G the 1-D
 * Lagrange polynomial on the 2 Gauss points.  Fields are pushed forward with
 * the contravariant Piola map j = J jhat / det J; grad j uses the affine form
 * (derivatives of J are dropped, exact for affine cells).
 *
 * Parity status: the element integrands are pinned by closed forms and
 * invariants checked in tests/test_oracle.py (1-D tensor-product stiffness and
 * mass, block transpose identities, finite-difference Jacobian, manufactured
 * solution with zero residual, Hunt solution norms against the reference's own
 * results).  Gridap's own basis/dof ordering is not reproduced: parity
 * unpinned for bit-exact agreement with Gridap's matrices.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NU 81
#define NP 4
#define NJ 36
#define NPHI 8
#define ND 129
#define OFF_P 81
#define OFF_J 85
#define OFF_PHI 121

enum { T_U = 0, T_P = 1, T_J = 2, T_PHI = 3 };

typedef struct {
  int32_t ncells;
  int32_t nfree;
  int32_t ndir;
  int32_t convection; /* 0 none, 1 picard, 2 newton */
  const double *cell_xyz;  /* [ncells][8][3] */
  const int32_t *cell_dofs; /* [ncells][129]; >=0 free id, <0 -> dirichlet -(k+1) */
  const int8_t *rt_sign;   /* [ncells][6] or NULL (all +1) */
  const double *dir_vals;  /* [ndir] */
  double alpha, beta, gamma, sigma;
  double B[3];
  double f0[3], fx[9];     /* f(x) = f0 + fx . x  (fx row-major, f_i = f0_i + sum_k fx[i*3+k] x_k) */
  double g0[3], gx[9];
  double zeta_u, zeta_j;
  int32_t nfaces;          /* thin-wall faces */
  int32_t pad_;
  const int32_t *face_cell;
  const int8_t *face_lface; /* 0..5 */
  const double *face_tau, *face_cw, *face_jw;
} orc_problem;

/* ---------------- 1-D bases and quadrature ---------------- */

static void lag2(double t, double v[3], double d[3]) {
  v[0] = (2.0 * t - 1.0) * (t - 1.0);
  v[1] = 4.0 * t * (1.0 - t);
  v[2] = t * (2.0 * t - 1.0);
  d[0] = 4.0 * t - 3.0;
  d[1] = 4.0 - 8.0 * t;
  d[2] = 4.0 * t - 1.0;
}

static void lag1(double t, double v[2], double d[2]) {
  v[0] = 1.0 - t;
  v[1] = t;
  d[0] = -1.0;
  d[1] = 1.0;
}

static void lagg(double t, double v[2], double d[2]) {
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  const double h = g1 - g0;
  v[0] = (g1 - t) / h;
  v[1] = (t - g0) / h;
  d[0] = -1.0 / h;
  d[1] = 1.0 / h;
}

/* n-point Gauss-Legendre on [0,1], n = 3 or 4 */
static void gauss01(int n, double *pt, double *w) {
  if (n == 3) {
    double a = sqrt(0.15);
    pt[0] = 0.5 - a; pt[1] = 0.5; pt[2] = 0.5 + a;
    w[0] = 5.0 / 18.0; w[1] = 8.0 / 18.0; w[2] = 5.0 / 18.0;
  } else {
    double a = sqrt(3.0 / 7.0 - 2.0 / 7.0 * sqrt(6.0 / 5.0));
    double b = sqrt(3.0 / 7.0 + 2.0 / 7.0 * sqrt(6.0 / 5.0));
    double wa = (18.0 + sqrt(30.0)) / 36.0, wb = (18.0 - sqrt(30.0)) / 36.0;
    pt[0] = 0.5 - 0.5 * b; pt[1] = 0.5 - 0.5 * a; pt[2] = 0.5 + 0.5 * a; pt[3] = 0.5 + 0.5 * b;
    w[0] = 0.5 * wb; w[1] = 0.5 * wa; w[2] = 0.5 * wa; w[3] = 0.5 * wb;
  }
}

/* RT dof r -> (axis d, normal-position m in {0,1,2}, tangential gauss g1, g2, face or -1) */
static void rt_dof_info(int r, int *d, int *m, int *g1, int *g2, int *face) {
  if (r < 24) {
    int f = r / 4, k = r % 4;
    *d = 2 - f / 2;
    *m = (f % 2) ? 2 : 0;
    *g1 = k % 2;
    *g2 = k / 2;
    *face = f;
  } else {
    int rr = r - 24;
    *d = rr / 4;
    *m = 1;
    *g1 = (rr % 4) % 2;
    *g2 = (rr % 4) / 2;
    *face = -1;
  }
}

static void tang_axes(int d, int *t1, int *t2) {
  if (d == 0) { *t1 = 1; *t2 = 2; }
  else if (d == 1) { *t1 = 0; *t2 = 2; }
  else { *t1 = 0; *t2 = 1; }
}

/* ---------------- reference tabulation at one point ---------------- */

typedef struct {
  double q2[27], dq2[27][3];
  double p1[4];
  double q1[8];
  double rt[36][3];      /* reference field */
  double drt[36][3][3];  /* drt[r][l][m] = d rt_l / d xi_m */
  double nq1[8], dnq1[8][3]; /* geometry shape functions (trilinear) */
} reftab;

static void tabulate(const double xi[3], reftab *T) {
  double L[3][3], dL[3][3], l1[3][2], dl1[3][2], G[3][2], dG[3][2];
  for (int k = 0; k < 3; ++k) {
    lag2(xi[k], L[k], dL[k]);
    lag1(xi[k], l1[k], dl1[k]);
    lagg(xi[k], G[k], dG[k]);
  }
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        int a = i + 3 * j + 9 * k;
        T->q2[a] = L[0][i] * L[1][j] * L[2][k];
        T->dq2[a][0] = dL[0][i] * L[1][j] * L[2][k];
        T->dq2[a][1] = L[0][i] * dL[1][j] * L[2][k];
        T->dq2[a][2] = L[0][i] * L[1][j] * dL[2][k];
      }
  T->p1[0] = 1.0 - xi[0] - xi[1] - xi[2];
  T->p1[1] = xi[0];
  T->p1[2] = xi[1];
  T->p1[3] = xi[2];
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        int v = i + 2 * j + 4 * k;
        T->q1[v] = l1[0][i] * l1[1][j] * l1[2][k];
        T->nq1[v] = T->q1[v];
        T->dnq1[v][0] = dl1[0][i] * l1[1][j] * l1[2][k];
        T->dnq1[v][1] = l1[0][i] * dl1[1][j] * l1[2][k];
        T->dnq1[v][2] = l1[0][i] * l1[1][j] * dl1[2][k];
      }
  for (int r = 0; r < 36; ++r) {
    int d, m, g1, g2, face, t1, t2;
    rt_dof_info(r, &d, &m, &g1, &g2, &face);
    tang_axes(d, &t1, &t2);
    memset(T->rt[r], 0, sizeof(T->rt[r]));
    memset(T->drt[r], 0, sizeof(T->drt[r]));
    double vn = L[d][m], vt1 = G[t1][g1], vt2 = G[t2][g2];
    T->rt[r][d] = vn * vt1 * vt2;
    T->drt[r][d][d] = dL[d][m] * vt1 * vt2;
    T->drt[r][d][t1] = vn * dG[t1][g1] * vt2;
    T->drt[r][d][t2] = vn * vt1 * dG[t2][g2];
  }
}

/* ---------------- geometry ---------------- */

typedef struct {
  double x[3];
  double J[3][3];    /* J[i][l] = d x_i / d xi_l */
  double Ji[3][3];   /* inverse */
  double det;
} geo;

static void geometry(const double *X /*8x3*/, const reftab *T, geo *g) {
  memset(g, 0, sizeof(*g));
  for (int v = 0; v < 8; ++v)
    for (int i = 0; i < 3; ++i) {
      g->x[i] += T->nq1[v] * X[3 * v + i];
      for (int l = 0; l < 3; ++l) g->J[i][l] += X[3 * v + i] * T->dnq1[v][l];
    }
  double (*J)[3] = g->J;
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  g->det = det;
  double id = 1.0 / det;
  g->Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  g->Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  g->Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  g->Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  g->Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  g->Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  g->Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  g->Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  g->Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

/* One basis function pushed to the physical cell, as a full multifield tuple
 * (v, grad v, div v, q, s, grad s, div s, phi): exactly one field is nonzero. */
typedef struct {
  int type;
  double v[3], gv[3][3], divv; /* gv[i][k] = d v_k / d x_i (Gridap's grad convention) */
  double q;
  double s[3], gs[3][3], divs;
  double ph;
} bfun;

static void push_basis(const reftab *T, const geo *g, const int8_t *sign6, bfun *bf) {
  memset(bf, 0, sizeof(bfun) * ND);
  for (int c = 0; c < 3; ++c)
    for (int a = 0; a < 27; ++a) {
      bfun *b = &bf[c * 27 + a];
      b->type = T_U;
      b->v[c] = T->q2[a];
      double gr[3];
      for (int i = 0; i < 3; ++i) {
        gr[i] = 0.0;
        for (int l = 0; l < 3; ++l) gr[i] += g->Ji[l][i] * T->dq2[a][l];
      }
      for (int i = 0; i < 3; ++i) b->gv[i][c] = gr[i];
      b->divv = gr[c];
    }
  for (int k = 0; k < 4; ++k) {
    bf[OFF_P + k].type = T_P;
    bf[OFF_P + k].q = T->p1[k];
  }
  for (int r = 0; r < 36; ++r) {
    bfun *b = &bf[OFF_J + r];
    b->type = T_J;
    double sg = 1.0;
    if (r < 24 && sign6) sg = (double)sign6[r / 4];
    double id = sg / g->det;
    double dhat = 0.0;
    for (int l = 0; l < 3; ++l) dhat += T->drt[r][l][l];
    b->divs = dhat * id;
    for (int k = 0; k < 3; ++k) {
      double acc = 0.0;
      for (int l = 0; l < 3; ++l) acc += g->J[k][l] * T->rt[r][l];
      b->s[k] = acc * id;
    }
    /* grad s [i][k] = d s_k / d x_i = (1/det) sum_l J[k][l] sum_m d rt_l/d xi_m Ji[m][i] */
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
        for (int l = 0; l < 3; ++l) {
          double dl = 0.0;
          for (int m = 0; m < 3; ++m) dl += T->drt[r][l][m] * g->Ji[m][i];
          acc += g->J[k][l] * dl;
        }
        b->gs[i][k] = acc * id;
      }
  }
  for (int v = 0; v < 8; ++v) {
    bf[OFF_PHI + v].type = T_PHI;
    bf[OFF_PHI + v].ph = T->q1[v];
  }
}

static inline void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Touched field blocks (row = test field, col = trial field) of the H1-HDiv
 * Jacobian: uu, up, uj, pu, jj, jphi, ju, phij (src/weakforms.jl:311). */
static const int TOUCHED[4][4] = {
    {1, 1, 1, 0},
    {1, 0, 0, 0},
    {1, 0, 1, 1},
    {0, 0, 1, 0},
};

int orc_touched(int row_type, int col_type) { return TOUCHED[row_type][col_type]; }

static int dof_type(int l) {
  if (l < OFF_P) return T_U;
  if (l < OFF_J) return T_P;
  if (l < OFF_PHI) return T_J;
  return T_PHI;
}

/* Local values of x on one cell (free values + Dirichlet lift). */
static void gather_local(const orc_problem *P, const double *xfree, int c, double *xl) {
  const int32_t *cd = P->cell_dofs + (size_t)c * ND;
  for (int l = 0; l < ND; ++l) {
    int32_t g = cd[l];
    xl[l] = (g >= 0) ? xfree[g] : P->dir_vals[-g - 1];
  }
}

/* Multifield state of x at a quadrature point (setup_variable). */
typedef struct {
  double u[3], gu[3][3], divu, p, j[3], gj[3][3], divj, ph;
} state;

static void eval_state(const bfun *bf, const double *xl, state *s) {
  memset(s, 0, sizeof(*s));
  for (int l = 0; l < ND; ++l) {
    double a = xl[l];
    if (a == 0.0) continue;
    const bfun *b = &bf[l];
    for (int i = 0; i < 3; ++i) {
      s->u[i] += a * b->v[i];
      s->j[i] += a * b->s[i];
      for (int k = 0; k < 3; ++k) {
        s->gu[i][k] += a * b->gv[i][k];
        s->gj[i][k] += a * b->gs[i][k];
      }
    }
    s->divu += a * b->divv;
    s->divj += a * b->divs;
    s->p += a * b->q;
    s->ph += a * b->ph;
  }
}

/* (grad a)^T . b : ((grad a)^T b)_k = sum_i grad a[i][k] b_i */
static inline void gradT_dot(const double ga[3][3], const double b[3], double out[3]) {
  for (int k = 0; k < 3; ++k) out[k] = ga[0][k] * b[0] + ga[1][k] * b[1] + ga[2][k] * b[2];
}

/* Residual integrand of res_fluid_h1_hdiv for test function y at state x (no AL-u term). */
static double res_integrand(const orc_problem *P, const state *x, const bfun *y, const double xp[3]) {
  double val = 0.0;
  /* u_block = beta grad u : grad v (+ alpha v.conv(u, grad u)) */
  double gg = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) gg += x->gu[i][k] * y->gv[i][k];
  val += P->beta * gg;
  if (P->convection != 0) {
    double cv[3];
    gradT_dot(x->gu, x->u, cv);
    val += P->alpha * dot3(y->v, cv);
  }
  val -= x->p * y->divv;
  double jxb[3], uxb[3];
  cross(x->j, P->B, jxb);
  cross(x->u, P->B, uxb);
  val -= P->gamma * dot3(jxb, y->v);
  val -= x->divu * y->q;
  val += dot3(x->j, y->s);
  if (P->zeta_j != 0.0) val += P->zeta_j * x->divj * y->divs;
  val -= P->sigma * x->ph * y->divs;
  val -= P->sigma * dot3(uxb, y->s);
  val -= x->divj * y->ph;
  double f[3], g[3];
  for (int i = 0; i < 3; ++i) {
    f[i] = P->f0[i];
    g[i] = P->g0[i];
    for (int k = 0; k < 3; ++k) {
      f[i] += P->fx[i * 3 + k] * xp[k];
      g[i] += P->gx[i * 3 + k] * xp[k];
    }
  }
  val -= dot3(f, y->v);
  val -= dot3(g, y->s);
  return val;
}

/* Jacobian integrand of jac_fluid_h1_hdiv for trial dx, test y, at state x (no AL-u term). */
static double jac_integrand(const orc_problem *P, const state *x, const bfun *dx, const bfun *y) {
  double val = 0.0;
  double gg = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) gg += dx->gv[i][k] * y->gv[i][k];
  val += P->beta * gg;
  if (P->convection == 1 || P->convection == 2) {
    double c1[3];
    gradT_dot(dx->gv, x->u, c1); /* conv(u, grad du) */
    if (P->convection == 2) {
      double c2[3];
      gradT_dot(x->gu, dx->v, c2); /* conv(du, grad u) */
      for (int k = 0; k < 3; ++k) c1[k] += c2[k];
    }
    val += P->alpha * dot3(y->v, c1);
  }
  val -= dx->q * y->divv;
  double djxb[3], duxb[3];
  cross(dx->s, P->B, djxb);
  cross(dx->v, P->B, duxb);
  val -= P->gamma * dot3(djxb, y->v);
  val -= dx->divv * y->q;
  val += dot3(dx->s, y->s);
  if (P->zeta_j != 0.0) val += P->zeta_j * dx->divs * y->divs;
  val -= P->sigma * dx->ph * y->divs;
  val -= P->sigma * dot3(duxb, y->s);
  val -= dx->divs * y->ph;
  return val;
}

static void inv4(const double A[16], double Ai[16]) {
  /* Gauss-Jordan with partial pivoting on a 4x4 SPD-ish matrix */
  double M[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 8; ++j) M[i][j] = (j < 4) ? A[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
    if (p != c)
      for (int j = 0; j < 8; ++j) { double t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
    double d = 1.0 / M[c][c];
    for (int j = 0; j < 8; ++j) M[c][j] *= d;
    for (int r = 0; r < 4; ++r)
      if (r != c) {
        double f = M[r][c];
        for (int j = 0; j < 8; ++j) M[r][j] -= f * M[c][j];
      }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) Ai[i * 4 + j] = M[i][j + 4];
}

/* Element residual (129) and, if Ke != NULL, element Jacobian (129x129, row =
 * test dof, col = trial dof) of cell c, fluid terms only. */
static void cell_kernel(const orc_problem *P, const double *xfree, int c, double *re, double *Ke) {
  double gp[3], gw[3];
  gauss01(3, gp, gw);
  const double *X = P->cell_xyz + (size_t)c * 24;
  const int8_t *sg = P->rt_sign ? P->rt_sign + (size_t)c * 6 : NULL;
  double xl[ND];
  gather_local(P, xfree, c, xl);
  if (re) memset(re, 0, sizeof(double) * ND);
  if (Ke) memset(Ke, 0, sizeof(double) * ND * ND);
  /* AL-u projection data: D[k][b] = int psi_k div(v_b), M[k][l] = int psi_k psi_l, du_k = int psi_k div u */
  double D[4][NU], M[16], dproj[4];
  memset(D, 0, sizeof(D));
  memset(M, 0, sizeof(M));
  memset(dproj, 0, sizeof(dproj));
  static __thread bfun bf[ND];
  for (int qz = 0; qz < 3; ++qz)
    for (int qy = 0; qy < 3; ++qy)
      for (int qx = 0; qx < 3; ++qx) {
        double xi[3] = {gp[qx], gp[qy], gp[qz]};
        reftab T;
        tabulate(xi, &T);
        geo g;
        geometry(X, &T, &g);
        double W = gw[qx] * gw[qy] * gw[qz] * fabs(g.det);
        push_basis(&T, &g, sg, bf);
        state s;
        eval_state(bf, xl, &s);
        if (re)
          for (int a = 0; a < ND; ++a) re[a] += W * res_integrand(P, &s, &bf[a], g.x);
        if (Ke)
          for (int a = 0; a < ND; ++a) {
            int ta = bf[a].type;
            for (int b = 0; b < ND; ++b) {
              if (!TOUCHED[ta][bf[b].type]) continue;
              Ke[a * ND + b] += W * jac_integrand(P, &s, &bf[b], &bf[a]);
            }
          }
        if (P->zeta_u != 0.0) {
          for (int k = 0; k < 4; ++k) {
            for (int l = 0; l < 4; ++l) M[k * 4 + l] += W * bf[OFF_P + k].q * bf[OFF_P + l].q;
            for (int b = 0; b < NU; ++b) D[k][b] += W * bf[OFF_P + k].q * bf[b].divv;
            dproj[k] += W * bf[OFF_P + k].q * s.divu;
          }
        }
      }
  if (P->zeta_u != 0.0) {
    /* zeta_u * int Pi_p(div u) div v, Pi_p = L2 projection onto dP1 on the cell */
    double Mi[16], cu[4];
    inv4(M, Mi);
    for (int k = 0; k < 4; ++k) {
      cu[k] = 0.0;
      for (int l = 0; l < 4; ++l) cu[k] += Mi[k * 4 + l] * dproj[l];
    }
    if (re)
      for (int a = 0; a < NU; ++a) {
        double acc = 0.0;
        for (int k = 0; k < 4; ++k) acc += D[k][a] * cu[k];
        re[a] += P->zeta_u * acc;
      }
    if (Ke) {
      double MD[4][NU];
      for (int k = 0; k < 4; ++k)
        for (int b = 0; b < NU; ++b) {
          double acc = 0.0;
          for (int l = 0; l < 4; ++l) acc += Mi[k * 4 + l] * D[l][b];
          MD[k][b] = acc;
        }
      for (int a = 0; a < NU; ++a)
        for (int b = 0; b < NU; ++b) {
          double acc = 0.0;
          for (int k = 0; k < 4; ++k) acc += D[k][a] * MD[k][b];
          Ke[a * ND + b] += P->zeta_u * acc;
        }
    }
  }
}

/* Thin-wall face terms (res_thin_wall / jac_thin_wall) for face index fi.
 * Output restricted to the 36 j dofs of the owning cell: rf[36], Kf[36x36]. */
static void face_kernel(const orc_problem *P, const double *xfree, int fi, double *rf, double *Kf) {
  int c = P->face_cell[fi];
  int lf = P->face_lface[fi];
  int d = 2 - lf / 2, side = lf % 2, t1, t2;
  tang_axes(d, &t1, &t2);
  double tau = P->face_tau[fi], cw = P->face_cw[fi], jw = P->face_jw[fi];
  double gp[3], gw[3];
  gauss01(3, gp, gw);
  const double *X = P->cell_xyz + (size_t)c * 24;
  const int8_t *sg = P->rt_sign ? P->rt_sign + (size_t)c * 6 : NULL;
  double xl[ND];
  gather_local(P, xfree, c, xl);
  if (rf) memset(rf, 0, sizeof(double) * NJ);
  if (Kf) memset(Kf, 0, sizeof(double) * NJ * NJ);
  static __thread bfun bf[ND];
  for (int b2 = 0; b2 < 3; ++b2)
    for (int b1 = 0; b1 < 3; ++b1) {
      double xi[3];
      xi[d] = side ? 1.0 : 0.0;
      xi[t1] = gp[b1];
      xi[t2] = gp[b2];
      reftab T;
      tabulate(xi, &T);
      geo g;
      geometry(X, &T, &g);
      double nh[3] = {0, 0, 0};
      nh[d] = side ? 1.0 : -1.0;
      double n[3];
      for (int i = 0; i < 3; ++i) {
        n[i] = 0.0;
        for (int l = 0; l < 3; ++l) n[i] += g.Ji[l][i] * nh[l];
      }
      double nn = sqrt(dot3(n, n));
      for (int i = 0; i < 3; ++i) n[i] /= nn;
      double W = gw[b1] * gw[b2] * fabs(g.det) * nn;
      push_basis(&T, &g, sg, bf);
      state s;
      eval_state(bf, xl, &s);
      double jn = dot3(s.j, n);
      double gjn[3];
      for (int i = 0; i < 3; ++i) gjn[i] = s.gj[i][0] * n[0] + s.gj[i][1] * n[1] + s.gj[i][2] * n[2];
      double jnn = dot3(n, gjn);
      for (int a = 0; a < NJ; ++a) {
        const bfun *y = &bf[OFF_J + a];
        double sn = dot3(y->s, n);
        if (rf) rf[a] += W * tau * sn * (jn - jw + cw * jnn);
        if (Kf)
          for (int b = 0; b < NJ; ++b) {
            const bfun *dx = &bf[OFF_J + b];
            double djn = dot3(dx->s, n);
            double t[3];
            for (int i = 0; i < 3; ++i) t[i] = dx->gs[i][0] * n[0] + dx->gs[i][1] * n[1] + dx->gs[i][2] * n[2];
            double djnn = dot3(n, t);
            Kf[a * NJ + b] += W * tau * sn * (djn + cw * djnn);
          }
      }
    }
}

/* ---------------- exported entry points ---------------- */

int orc_cell_residuals(const orc_problem *P, const double *xfree, int c0, int c1, double *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int c = c0; c < c1; ++c) cell_kernel(P, xfree, c, out + (size_t)(c - c0) * ND, NULL);
  return 0;
}

int orc_cell_jacobians(const orc_problem *P, const double *xfree, int c0, int c1, double *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = c0; c < c1; ++c) cell_kernel(P, xfree, c, NULL, out + (size_t)(c - c0) * ND * ND);
  return 0;
}

/* Residual assembled over cells [c0, c1): r[gi] += re[a] for the free dofs of
 * each cell (the vector assembly of src/main.jl:218-233, summed with atomics
 * across threads). */
int orc_assemble_residual(const orc_problem *P, const double *xfree, int c0, int c1, double *r) {
#pragma omp parallel
  {
    double re[ND];
#pragma omp for schedule(dynamic, 4)
    for (int c = c0; c < c1; ++c) {
      cell_kernel(P, xfree, c, re, NULL);
      const int32_t *cd = P->cell_dofs + (size_t)c * ND;
      for (int a = 0; a < ND; ++a)
        if (cd[a] >= 0) {
#pragma omp atomic
          r[cd[a]] += re[a];
        }
    }
  }
  return 0;
}

/* Jacobian assembled over cells [c0, c1) into the values of an existing CSC
 * pattern (0-based colptr / rowval, rows sorted inside each column): each
 * touched (test a, trial b) entry is located by binary search in column
 * cell_dofs[b], the way a sparse assembler adds a cell matrix into a fixed
 * pattern (src/main.jl:218-233).  Returns the number of entries not found. */
int64_t orc_assemble_csc(const orc_problem *P, const double *xfree, int c0, int c1, const int64_t *colptr,
                         const int64_t *rowval, double *nzval) {
  int64_t missing = 0;
#pragma omp parallel reduction(+ : missing)
  {
    double *Ke = (double *)malloc(sizeof(double) * ND * ND);
#pragma omp for schedule(dynamic, 1)
    for (int c = c0; c < c1; ++c) {
      cell_kernel(P, xfree, c, NULL, Ke);
      const int32_t *cd = P->cell_dofs + (size_t)c * ND;
      for (int b = 0; b < ND; ++b) {
        const int32_t gj = cd[b];
        if (gj < 0) continue;
        const int tb = dof_type(b);
        for (int a = 0; a < ND; ++a) {
          const int32_t gi = cd[a];
          if (gi < 0 || !TOUCHED[dof_type(a)][tb]) continue;
          int64_t lo = colptr[gj], hi = colptr[gj + 1];
          while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (rowval[mid] < gi) lo = mid + 1; else hi = mid;
          }
          if (lo < colptr[gj + 1] && rowval[lo] == gi) {
#pragma omp atomic
            nzval[lo] += Ke[a * ND + b];
          } else {
            ++missing;
          }
        }
      }
    }
    free(Ke);
  }
  return missing;
}

int orc_face_residuals(const orc_problem *P, const double *xfree, double *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int f = 0; f < P->nfaces; ++f) face_kernel(P, xfree, f, out + (size_t)f * NJ, NULL);
  return 0;
}

int orc_face_jacobians(const orc_problem *P, const double *xfree, double *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int f = 0; f < P->nfaces; ++f) face_kernel(P, xfree, f, NULL, out + (size_t)f * NJ * NJ);
  return 0;
}

/* Per-cell integrals with an n-point Gauss rule (n = 3 or 4; 4 = Gridap's
 * degree-6 measure used for the Hunt post-processing, src/Applications/hunt.jl:247-260):
 * out[c] = (int u.u, int grad u : grad u, int j.j, int div j^2, int p, int |Omega_c|). */
int orc_cell_norms(const orc_problem *P, const double *xfree, int nq, double *out) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int c = 0; c < P->ncells; ++c) {
    double gp[4], gw[4];
    gauss01(nq, gp, gw);
    const double *X = P->cell_xyz + (size_t)c * 24;
    const int8_t *sg = P->rt_sign ? P->rt_sign + (size_t)c * 6 : NULL;
    double xl[ND];
    gather_local(P, xfree, c, xl);
    bfun *bf = (bfun *)malloc(sizeof(bfun) * ND);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int qz = 0; qz < nq; ++qz)
      for (int qy = 0; qy < nq; ++qy)
        for (int qx = 0; qx < nq; ++qx) {
          double xi[3] = {gp[qx], gp[qy], gp[qz]};
          reftab T;
          tabulate(xi, &T);
          geo g;
          geometry(X, &T, &g);
          double W = gw[qx] * gw[qy] * gw[qz] * fabs(g.det);
          push_basis(&T, &g, sg, bf);
          state s;
          eval_state(bf, xl, &s);
          double gg = 0.0;
          for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) gg += s.gu[i][k] * s.gu[i][k];
          acc[0] += W * dot3(s.u, s.u);
          acc[1] += W * gg;
          acc[2] += W * dot3(s.j, s.j);
          acc[3] += W * s.divj * s.divj;
          acc[4] += W * s.p;
          acc[5] += W;
        }
    free(bf);
    for (int k = 0; k < 6; ++k) out[(size_t)c * 6 + k] = acc[k];
  }
  return 0;
}

/* Field values at reference points of given cells (post-processing / tests):
 * out[i] = (u[3], j[3], p, phi, grad u[9]) at cell cells[i], reference point xi[i]. */
int orc_point_values(const orc_problem *P, const double *xfree, int n, const int32_t *cells,
                     const double *xi, double *out) {
  bfun *bf = (bfun *)malloc(sizeof(bfun) * ND);
  for (int i = 0; i < n; ++i) {
    int c = cells[i];
    const double *X = P->cell_xyz + (size_t)c * 24;
    const int8_t *sg = P->rt_sign ? P->rt_sign + (size_t)c * 6 : NULL;
    double xl[ND];
    gather_local(P, xfree, c, xl);
    reftab T;
    tabulate(xi + 3 * i, &T);
    geo g;
    geometry(X, &T, &g);
    push_basis(&T, &g, sg, bf);
    state s;
    eval_state(bf, xl, &s);
    double *o = out + (size_t)i * 17;
    for (int k = 0; k < 3; ++k) { o[k] = s.u[k]; o[3 + k] = s.j[k]; }
    o[6] = s.p;
    o[7] = s.ph;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) o[8 + 3 * a + b] = s.gu[a][b];
  }
  free(bf);
  return 0;
}

/* Reference tabulation export, for closed-form checks of the reference elements. */
int orc_tabulate(const double *xi, double *q2, double *dq2, double *p1, double *q1, double *rt, double *drt) {
  reftab T;
  tabulate(xi, &T);
  memcpy(q2, T.q2, sizeof(T.q2));
  memcpy(dq2, T.dq2, sizeof(T.dq2));
  memcpy(p1, T.p1, sizeof(T.p1));
  memcpy(q1, T.q1, sizeof(T.q1));
  memcpy(rt, T.rt, sizeof(T.rt));
  memcpy(drt, T.drt, sizeof(T.drt));
  return 0;
}

int orc_struct_size(void) { return (int)sizeof(orc_problem); }
