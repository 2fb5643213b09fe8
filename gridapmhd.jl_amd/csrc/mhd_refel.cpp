// Reference-element tables of the four hexahedral spaces at the quadrature
// points: Q2 vector Lagrangian (u), discontinuous P1 (p), Raviart-Thomas
// order 1 (j), discontinuous Q1 (phi), and the trilinear geometry map.
// Quadrature: tensor Gauss-Legendre with 3 points per direction, the rule of
// degree q = 5 chosen by params_fespaces (src/parameters.jl:380-389).
#include <cmath>
#include <cstring>

#include "mhd_internal.h"

namespace mhd {
namespace {

// 1-D Lagrange polynomial through `nodes` (n of them), index k, value and derivative at t.
void lagrange1d(int n, const double *nodes, int k, double t, double *v, double *d) {
  double val = 1.0, der = 0.0;
  for (int m = 0; m < n; ++m) {
    if (m == k) continue;
    double den = nodes[k] - nodes[m];
    // product rule accumulated incrementally
    der = der * (t - nodes[m]) / den + val / den;
    val = val * (t - nodes[m]) / den;
  }
  *v = val;
  *d = der;
}

const double kQ2Nodes[3] = {0.0, 0.5, 1.0};
const double kQ1Nodes[2] = {0.0, 1.0};

void gauss3(double *x, double *w) {
  const double a = std::sqrt(3.0 / 5.0);
  x[0] = 0.5 * (1.0 - a); x[1] = 0.5; x[2] = 0.5 * (1.0 + a);
  w[0] = 5.0 / 18.0; w[1] = 8.0 / 18.0; w[2] = 5.0 / 18.0;
}

// RT order-1 local dof r -> axis d, normal-node index m, tangential gauss indices.
void rt_layout(int r, int &d, int &m, int &ga, int &gb) {
  if (r < 24) {
    const int face = r >> 2, k = r & 3;
    d = 2 - face / 2;           // faces z-, z+, y-, y+, x-, x+
    m = (face & 1) ? 2 : 0;
    ga = k & 1;
    gb = k >> 1;
  } else {
    const int rr = r - 24;
    d = rr >> 2;
    m = 1;
    ga = rr & 1;
    gb = (rr & 3) >> 1;
  }
}

void tangents(int d, int &ta, int &tb) {
  ta = (d == 0) ? 1 : 0;
  tb = (d == 2) ? 1 : 2;
}

struct PointTab {
  double q2[27], dq2[27][3], p1[4], q1[8], dq1[8][3], rt[36][3], drt[36][3][3];
};

void tabulate_point(const double xi[3], PointTab &P) {
  double gn[2];
  {
    const double s = 0.5 / std::sqrt(3.0);
    gn[0] = 0.5 - s;
    gn[1] = 0.5 + s;
  }
  double L2[3][3], dL2[3][3], L1[3][2], dL1[3][2], Lg[3][2], dLg[3][2];
  for (int a = 0; a < 3; ++a) {
    for (int k = 0; k < 3; ++k) lagrange1d(3, kQ2Nodes, k, xi[a], &L2[a][k], &dL2[a][k]);
    for (int k = 0; k < 2; ++k) lagrange1d(2, kQ1Nodes, k, xi[a], &L1[a][k], &dL1[a][k]);
    for (int k = 0; k < 2; ++k) lagrange1d(2, gn, k, xi[a], &Lg[a][k], &dLg[a][k]);
  }
  for (int n = 0; n < 27; ++n) {
    const int i = n % 3, j = (n / 3) % 3, k = n / 9;
    P.q2[n] = L2[0][i] * L2[1][j] * L2[2][k];
    P.dq2[n][0] = dL2[0][i] * L2[1][j] * L2[2][k];
    P.dq2[n][1] = L2[0][i] * dL2[1][j] * L2[2][k];
    P.dq2[n][2] = L2[0][i] * L2[1][j] * dL2[2][k];
  }
  P.p1[0] = 1.0 - xi[0] - xi[1] - xi[2];
  P.p1[1] = xi[0];
  P.p1[2] = xi[1];
  P.p1[3] = xi[2];
  for (int v = 0; v < 8; ++v) {
    const int i = v & 1, j = (v >> 1) & 1, k = v >> 2;
    P.q1[v] = L1[0][i] * L1[1][j] * L1[2][k];
    P.dq1[v][0] = dL1[0][i] * L1[1][j] * L1[2][k];
    P.dq1[v][1] = L1[0][i] * dL1[1][j] * L1[2][k];
    P.dq1[v][2] = L1[0][i] * L1[1][j] * dL1[2][k];
  }
  std::memset(P.rt, 0, sizeof(P.rt));
  std::memset(P.drt, 0, sizeof(P.drt));
  for (int r = 0; r < 36; ++r) {
    int d, m, ga, gb, ta, tb;
    rt_layout(r, d, m, ga, gb);
    tangents(d, ta, tb);
    const double fn = L2[d][m], fa = Lg[ta][ga], fb = Lg[tb][gb];
    P.rt[r][d] = fn * fa * fb;
    P.drt[r][d][d] = dL2[d][m] * fa * fb;
    P.drt[r][d][ta] = fn * dLg[ta][ga] * fb;
    P.drt[r][d][tb] = fn * fa * dLg[tb][gb];
  }
}

}  // namespace

void build_ref_tables(RefTables *T) {
  double x[3], w[3];
  gauss3(x, w);
  for (int q = 0; q < NQ; ++q) {
    const int qx = q % 3, qy = (q / 3) % 3, qz = q / 9;
    const double xi[3] = {x[qx], x[qy], x[qz]};
    T->w[q] = w[qx] * w[qy] * w[qz];
    PointTab P;
    tabulate_point(xi, P);
    std::memcpy(T->q2[q], P.q2, sizeof(P.q2));
    std::memcpy(T->dq2[q], P.dq2, sizeof(P.dq2));
    std::memcpy(T->p1[q], P.p1, sizeof(P.p1));
    std::memcpy(T->q1[q], P.q1, sizeof(P.q1));
    std::memcpy(T->dq1[q], P.dq1, sizeof(P.dq1));
    std::memcpy(T->rt[q], P.rt, sizeof(P.rt));
    for (int r = 0; r < 36; ++r) T->rtdiv[q][r] = P.drt[r][0][0] + P.drt[r][1][1] + P.drt[r][2][2];
  }
}

void build_face_tables(FaceTables *F) {
  double x[3], w[3];
  gauss3(x, w);
  for (int f = 0; f < 6; ++f) {
    const int d = 2 - f / 2;
    int ta, tb;
    tangents(d, ta, tb);
    for (int q = 0; q < NQF; ++q) {
      const int qa = q % 3, qb = q / 3;
      double xi[3];
      xi[d] = (f & 1) ? 1.0 : 0.0;
      xi[ta] = x[qa];
      xi[tb] = x[qb];
      F->w[q] = w[qa] * w[qb];
      PointTab P;
      tabulate_point(xi, P);
      std::memcpy(F->q1[f][q], P.q1, sizeof(P.q1));
      std::memcpy(F->dq1[f][q], P.dq1, sizeof(P.dq1));
      std::memcpy(F->rt[f][q], P.rt, sizeof(P.rt));
      std::memcpy(F->drt[f][q], P.drt, sizeof(P.drt));
      std::memcpy(F->dq2[f][q], P.dq2, sizeof(P.dq2));
    }
  }
}

}  // namespace mhd
