/*
 * ORACLE — test infrastructure only (see oracle/__init__.py).
 *
 * C/OpenMP restatement of the loop of
 * code/internal/fn_explicit_dynamic_solver_v6.m:180-212 for CPU timing and
 * large-size checks. The caller (oracle/march.py) eliminates the diagonal
 * part of X = I + dt/2 M^-1 C into per-row coefficients:
 *   w_i = beta_i u1_i + gamma_i u2_i - alpha_i sum_j K_ij u1_j + alpha_i f_i
 * and, for rows coupled by off-diagonal C (fluid-solid interface),
 *   y_k = w_k + sq sum_j Coff_kj u1_j + sr sum_j Coff_kj u2_j
 *   u_I = Xinv y          (Xinv = identity when xptr == NULL)
 * f holds F(:, ti) at the forcing DOFs (last duplicate wins, v6:182).
 * hist is q x n_t column-major; columns before ti_start stay 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_march(int64_t n, const int64_t *kptr, const int32_t *kcol, const double *kval,
                 const double *alpha, const double *beta, const double *gamma,
                 int32_t nI, const int32_t *iface, const int64_t *cptr, const int32_t *ccol,
                 const double *cval, double sq, double sr, const int64_t *xptr,
                 const int32_t *xcol, const double *xval, int32_t n_t, int32_t t0,
                 int32_t t_end, int32_t p, const int32_t *fidx, const double *F,
                 int32_t F_rows, int32_t q, const int32_t *hidx, double *hist) {
  double *U = (double *)calloc((size_t)n * 3, sizeof(double));
  double *f = (double *)calloc((size_t)n, sizeof(double));
  double *y = (double *)calloc((size_t)(nI > 0 ? nI : 1), sizeof(double));
  if (!U || !f || !y) {
    free(U); free(f); free(y);
    return 1;
  }
  double *u = U, *u1 = U + n, *u2 = U + 2 * n;
  if (t_end > n_t) t_end = n_t;
  for (int32_t t = t0; t < t_end; ++t) {
    for (int32_t k = 0; k < p; ++k) f[fidx[k]] = F[(int64_t)t * F_rows + (F_rows == 1 ? 0 : k)];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t e = kptr[i]; e < kptr[i + 1]; ++e) s += kval[e] * u1[kcol[e]];
      u[i] = beta[i] * u1[i] + gamma[i] * u2[i] + alpha[i] * (f[i] - s);
    }
    if (nI > 0) {
      for (int32_t k = 0; k < nI; ++k) {
        double a1 = 0.0, a2 = 0.0;
        for (int64_t e = cptr[k]; e < cptr[k + 1]; ++e) {
          a1 += cval[e] * u1[ccol[e]];
          a2 += cval[e] * u2[ccol[e]];
        }
        y[k] = u[iface[k]] + sq * a1 + sr * a2;
      }
      for (int32_t k = 0; k < nI; ++k) {
        double acc = 0.0;
        if (xptr) {
          for (int64_t e = xptr[k]; e < xptr[k + 1]; ++e) acc += xval[e] * y[xcol[e]];
        } else {
          acc = y[k];
        }
        u[iface[k]] = acc;
      }
    }
    for (int32_t k = 0; k < q; ++k) hist[(int64_t)t * q + k] = u[hidx[k]];
    double *tmp = u2;
    u2 = u1;
    u1 = u;
    u = tmp;
  }
  free(U);
  free(f);
  free(y);
  return 0;
}
