/* CPU kernels of the Newton-step baseline -- TEST INFRASTRUCTURE ONLY.
 *
 * Only the CPU-baseline leg of bench.py loads this (as part of
 * oracle/_build/libcpustep.so, together with the element kernels of
 * mhd_oracle.c).  It prices on the host cores the operations one Newton step
 * of the Badia2024 solve performs (src/Solvers/badia2024.jl:2-48,
 * src/Solvers/gmg.jl:62-81, 323-355): sparse products, patch-inverse
 * products of the vertex-star smoother, dense coarsest products, Krylov
 * vector work, patch and coarsest inversions.  Every routine is an OpenMP
 * loop over independent rows / patches; none is used to check a result.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* y = A x, CSR with 64-bit row pointers and 32-bit column indices. */
int cpu_spmv_f64(int64_t n, const int64_t *rowptr, const int32_t *col, const double *val, const double *x,
                 double *y) {
#pragma omp parallel for schedule(static, 4096)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] = s;
  }
  return 0;
}

int cpu_spmv_f32(int64_t n, const int64_t *rowptr, const int32_t *col, const float *val, const float *x,
                 float *y) {
#pragma omp parallel for schedule(static, 4096)
  for (int64_t i = 0; i < n; ++i) {
    float s = 0.0f;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] = s;
  }
  return 0;
}

/* part[p][i] = sum_j A_p[i][j] r[idx[p][j]] for np dense PS x PS patch
 * inverses (row-major): the product half of one smoother application. */
int cpu_patch_apply_f64(int np, int PS, const int32_t *idx, const double *A, const double *r, double *part) {
#pragma omp parallel
  {
    double *rl = (double *)malloc(sizeof(double) * PS);
#pragma omp for schedule(static)
    for (int p = 0; p < np; ++p) {
      for (int j = 0; j < PS; ++j) rl[j] = r[idx[(int64_t)p * PS + j]];
      const double *Ap = A + (int64_t)p * PS * PS;
      for (int i = 0; i < PS; ++i) {
        double s = 0.0;
        for (int j = 0; j < PS; ++j) s += Ap[(int64_t)i * PS + j] * rl[j];
        part[(int64_t)p * PS + i] = s;
      }
    }
    free(rl);
  }
  return 0;
}

int cpu_patch_apply_f32(int np, int PS, const int32_t *idx, const float *A, const float *r, float *part) {
#pragma omp parallel
  {
    float *rl = (float *)malloc(sizeof(float) * PS);
#pragma omp for schedule(static)
    for (int p = 0; p < np; ++p) {
      for (int j = 0; j < PS; ++j) rl[j] = r[idx[(int64_t)p * PS + j]];
      const float *Ap = A + (int64_t)p * PS * PS;
      for (int i = 0; i < PS; ++i) {
        float s = 0.0f;
        for (int j = 0; j < PS; ++j) s += Ap[(int64_t)i * PS + j] * rl[j];
        part[(int64_t)p * PS + i] = s;
      }
    }
    free(rl);
  }
  return 0;
}

/* In-place Gauss-Jordan inversion of np PS x PS patches with diagonal pivots
 * (the patch matrices of the smoother set-up are diagonally dominant enough
 * for it); returns the number of zero pivots met. */
int cpu_patch_invert_f64(int np, int PS, double *A) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int p = 0; p < np; ++p) {
    double *M = A + (int64_t)p * PS * PS;
    for (int k = 0; k < PS; ++k) {
      double *rk = M + (int64_t)k * PS;
      const double piv = rk[k];
      if (piv == 0.0) {
        ++bad;
        break;
      }
      const double d = 1.0 / piv;
      rk[k] = 1.0;
      for (int j = 0; j < PS; ++j) rk[j] *= d;
      for (int i = 0; i < PS; ++i) {
        if (i == k) continue;
        double *ri = M + (int64_t)i * PS;
        const double f = ri[k];
        ri[k] = 0.0;
        for (int j = 0; j < PS; ++j) ri[j] -= f * rk[j];
      }
    }
  }
  return bad;
}

/* y = A x with a dense row-major m x m A (the coarsest-level inverse). */
int cpu_dense_matvec(int64_t m, const double *A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    double s = 0.0;
    const double *a = A + i * m;
    for (int64_t j = 0; j < m; ++j) s += a[j] * x[j];
    y[i] = s;
  }
  return 0;
}

/* out[k] = V_k . w for K vectors of length n (stride n): the dot products of
 * a classical Gram-Schmidt step. */
int cpu_mdot_f64(int64_t n, int K, const double *V, const double *w, double *out) {
  for (int k = 0; k < K; ++k) out[k] = 0.0;
#pragma omp parallel
  {
    double acc[64];
    for (int k = 0; k < K && k < 64; ++k) acc[k] = 0.0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double wi = w[i];
      for (int k = 0; k < K && k < 64; ++k) acc[k] += V[(int64_t)k * n + i] * wi;
    }
#pragma omp critical
    for (int k = 0; k < K && k < 64; ++k) out[k] += acc[k];
  }
  return 0;
}
