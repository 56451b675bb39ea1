/* CPU oracle: linearized power-series Newton on real multiple-double series.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (phcpack_amd/) may link or
 * call this code; only tests/ and the benchmark's CPU baseline use it.
 *
 * Restates, for K = 1, 2, 4, 8 limbs with the operation order of the
 * reference's double-double host code (the quad and octo double files repeat
 * the same sequences with their own arithmetic):
 *   Householder vector     CPU_dbl2_factors_house (Matrices/dbl2_factorizations.cpp:374-440)
 *   R and Q updates        CPU_dbl2_factors_leftRupdate (:534-566),
 *                          CPU_dbl2_factors_rightQupdate (:629-660)
 *   QR with explicit Q     CPU_dbl2_factors_houseqr (:723-765)
 *   Q^T b, back subst.     CPU_dbl2_factors_qrbs (:823-840), CPU_dbl2_factors_backward (:121-140)
 *   max-norm of leading    CPU_dbl_onenorm (Newton/dbl_onenorms_host.cpp:9-17)
 *   Toeplitz solve         CPU_dbl2_qrbs_head (Newton/dbl2_bals_host.cpp:15-105),
 *                          CPU_dbl2_qrbs_tail (:228-335), CPU_dbl2_qrbs_solve (:479-506)
 *   residual               CPU_dbl2_linear_residue (:548-600)
 *   Newton step            dbl2_row_newton_qrstep host branch (Newton/dbl2_newton_method.cpp:429-594),
 *                          with exponents > 1 by the common-factor trick of
 *                          CPU_dbl2_evaluate_monomials (Newton/dbl2_systems_host.cpp:12-111)
 *                          applied per monomial (series.c, osr_poly_evaldiff_exp),
 *                          dbl2_update_newton_qrstep host branch (:134-202),
 *                          dbl2_map_evaldiff_output (Newton/dbl2_systems_host.cpp:709-757),
 *                          dbl2_update_series (Newton/dbl2_newton_testers.cpp:116-149)
 *   staggered degrees      test_dbl2_row_newton loop (Newton/dbl2_newton_method.cpp:1447-1494)
 *   column Newton          test_dbl2_column_newton loop (:1036-1287) over
 *                          dbl2_column_newton_qrstep host branch (:259-427):
 *                          CPU_dbl2_evaluate_monomials (Newton/dbl2_systems_host.cpp:12-111)
 *                          with dbl2_linearize_evaldiff_output (:443-523), or
 *                          CPU_dbl2_evaluate_columns (:232-329) with dbl2_define_rhs
 *                          (:626-659); monomials by CPU_dbl2_evaldiff
 *                          (Convolutions/dbl2_monomials_host.cpp:175-283)
 * Skip thresholds per precision: Newton/dbl_bals_host.cpp:199 (1e-15; its
 * head tests nrm + 1 == 1 instead, :44),
 * dbl2_bals_host.cpp:43 (1e-28), dbl4_bals_host.cpp:51 (1e-56),
 * dbl8_bals_host.cpp:70 (1e-120).
 *
 * Complex (K = 2, 4, 8; the operation order of the complex double-double
 * host code):
 *   QR with explicit Q     CPU_cmplx2_factors_houseqr (Matrices/dbl2_factorizations.cpp:766-822)
 *                          with _house, _leftRupdate (baqr.c) and _rightQupdate (:661-722)
 *   Q^H b, back subst.     CPU_cmplx2_factors_qrbs (:842-879), _backward (:142-200)
 *   norm                   CPU_cmplx_onenorm (Newton/dbl_onenorms_host.cpp:19-28)
 *   Toeplitz solve         CPU_cmplx2_qrbs_head, _tail, _solve (Newton/dbl2_bals_host.cpp)
 *   residual               CPU_cmplx2_linear_residue (Newton/dbl2_bals_host.cpp)
 *   Newton step            cmplx2_row_newton_qrstep host branch (Newton/cmplx2_newton_method.cpp:565-790),
 *                          cmplx2_update_newton_qrstep (:163-328),
 *                          cmplx2_update_series (Newton/dbl2_newton_testers.cpp:151-180)
 *   staggered degrees      test_cmplx2_row_newton (Newton/cmplx2_newton_method.cpp:1759)
 *
 * Parity pinned by tests/test_oracle_newton.py and test_oracle_columns.py:
 * bit-exact against the reference functions compiled in oracle/_ref (QR,
 * Toeplitz solve, residual, whole staggered row and column runs at K = 1, 2,
 * 4, 8) and against committed fixtures of the reference's own cyclic row
 * system and its monomial, two-column and cyclic column systems; QR identities Q^T Q = I, QR = A;
 * a planted exponential-series solution recovered to working precision.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "md.h"
#include "series.h"
#include "newton.h"
#include "cxlin.h"

#define MAXK 8
#define E(a, i) ((a) + (size_t)(i) * K)

double onw_threshold(int K) {
  switch (K) {
    case 1: return 1.0e-15;
    case 2: return 1.0e-28;
    case 4: return 1.0e-56;
    default: return 1.0e-120;
  }
}

/* planar <-> interleaved (number j at j*K) */
static void from_planar(int K, size_t n, const double *p, double *a) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < K; q++) a[j * K + q] = p[q * n + j];
}

static void to_planar(int K, size_t n, const double *a, double *p) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < K; q++) p[q * n + j] = a[j * K + q];
}

static int is_zero(int K, const double *a) {
  int q;
  for (q = 0; q < K; q++)
    if (a[q] != 0.0) return 0;
  return 1;
}

static void house(int K, int n, const double *x, double *v, double *beta) {
  double sigma[MAXK], acc[MAXK], mu[MAXK], v0p2[MAXK];
  int i;
  omd_zero(K, sigma);
  omd_from_d(K, 1.0, E(v, 0));
  for (i = 1; i < n; i++) {
    omd_sqr(K, E(x, i), acc);
    omd_add(K, sigma, acc, sigma);
    omd_copy(K, E(x, i), E(v, i));
  }
  if (is_zero(K, sigma)) {
    omd_zero(K, beta);
    return;
  }
  omd_sqr(K, E(x, 0), acc);
  omd_add(K, acc, sigma, acc);
  omd_sqrt(K, acc, mu);
  if (x[0] <= 0.0) {
    omd_sub(K, E(x, 0), mu, E(v, 0));
  } else {
    omd_add(K, E(x, 0), mu, acc);
    omd_div(K, sigma, acc, E(v, 0));
    omd_neg(K, E(v, 0), E(v, 0));
  }
  omd_sqr(K, E(v, 0), v0p2);
  omd_add(K, sigma, v0p2, acc);
  omd_div(K, v0p2, acc, beta);
  omd_mul_d(K, beta, 2.0, beta);
  for (i = 1; i < n; i++) omd_div(K, E(v, i), E(v, 0), E(v, i));
  omd_from_d(K, 1.0, E(v, 0));
}

/* R: nrows x ncols interleaved */
static void left_r_update(int K, int nrows, int ncols, int k, double *R, const double *v,
                          const double *beta, double *w) {
  double acc[MAXK];
  int i, j;
  for (j = k; j < ncols; j++) {
    omd_zero(K, E(w, j - k));
    for (i = k; i < nrows; i++) {
      omd_mul(K, E(R, (size_t)i * ncols + j), E(v, i - k), acc);
      omd_add(K, E(w, j - k), acc, E(w, j - k));
    }
    omd_mul(K, beta, E(w, j - k), E(w, j - k));
  }
  for (i = k; i < nrows; i++)
    for (j = k; j < ncols; j++) {
      omd_mul(K, E(v, i - k), E(w, j - k), acc);
      omd_sub(K, E(R, (size_t)i * ncols + j), acc, E(R, (size_t)i * ncols + j));
    }
}

static void right_q_update(int K, int n, int k, double *Q, const double *v, const double *beta,
                           double *w) {
  double acc[MAXK];
  int i, j;
  for (i = 0; i < n; i++) {
    omd_zero(K, E(w, i));
    for (j = k; j < n; j++) {
      omd_mul(K, E(Q, (size_t)i * n + j), E(v, j - k), acc);
      omd_add(K, E(w, i), acc, E(w, i));
    }
    omd_mul(K, beta, E(w, i), E(w, i));
  }
  for (i = 0; i < n; i++)
    for (j = k; j < n; j++) {
      omd_mul(K, E(w, i), E(v, j - k), acc);
      omd_sub(K, E(Q, (size_t)i * n + j), acc, E(Q, (size_t)i * n + j));
    }
}

static void houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R) {
  double *x = malloc(sizeof(double) * K * nrows);
  double *v = malloc(sizeof(double) * K * nrows);
  double *w = malloc(sizeof(double) * K * (nrows > ncols ? nrows : ncols));
  double beta[MAXK];
  int i, j, k;
  for (i = 0; i < nrows; i++) {
    for (j = 0; j < nrows; j++) omd_zero(K, E(Q, (size_t)i * nrows + j));
    omd_from_d(K, 1.0, E(Q, (size_t)i * nrows + i));
    for (j = 0; j < ncols; j++) omd_copy(K, E(A, (size_t)i * ncols + j), E(R, (size_t)i * ncols + j));
  }
  for (k = 0; k < ncols; k++) {
    if (nrows - k > 0) {
      for (i = k; i < nrows; i++) omd_copy(K, E(R, (size_t)i * ncols + k), E(x, i - k));
      house(K, nrows - k, x, v, beta);
      left_r_update(K, nrows, ncols, k, R, v, beta, w);
      right_q_update(K, nrows, k, Q, v, beta, w);
    }
  }
  free(x);
  free(v);
  free(w);
}

static void backward(int K, int dim, const double *U, const double *b, double *x) {
  double acc[MAXK];
  int i, j;
  for (i = dim - 1; i >= 0; i--) {
    omd_copy(K, E(b, i), E(x, i));
    for (j = dim - 1; j > i; j--) {
      omd_mul(K, E(U, (size_t)i * dim + j), E(x, j), acc);
      omd_sub(K, E(x, i), acc, E(x, i));
    }
    omd_div(K, E(x, i), E(U, (size_t)i * dim + i), E(x, i));
  }
}

static void qrbs(int K, int dim, const double *Q, const double *R, const double *b, double *x,
                 double *wrk) {
  double acc[MAXK];
  int i, j;
  for (i = 0; i < dim; i++) {
    omd_zero(K, E(wrk, i));
    for (j = 0; j < dim; j++) {
      omd_mul(K, E(Q, (size_t)j * dim + i), E(b, j), acc);
      omd_add(K, E(wrk, i), acc, E(wrk, i));
    }
  }
  backward(K, dim, R, wrk, x);
}

/* largest |leading limb| of a vector of dim numbers */
static double maxnorm(int K, int dim, const double *v) {
  double r = fabs(v[0]);
  int i;
  for (i = 1; i < dim; i++)
    if (fabs(v[(size_t)i * K]) > r) r = fabs(v[(size_t)i * K]);
  return r;
}

static void qrbs_head(int K, int dim, const double *mat, const double *rhs, double *sol,
                      double *Q, double *R, double *wrk, int *zeroQ, int *noqr) {
  const double nrm = maxnorm(K, dim, rhs);
  /* the double version tests nrm + 1 == 1 (Newton/dbl_bals_host.cpp:44) */
  const int small = (K == 1) ? (nrm + 1.0 == 1.0) : (nrm < onw_threshold(K));
  if (small) *noqr = *zeroQ ? 0 : 1;
  if (!*noqr) {
    houseqr(K, dim, dim, mat, Q, R);
    *zeroQ = 0;
    qrbs(K, dim, Q, R, rhs, sol, wrk);
  }
}

static void qrbs_tail(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                      double *sol, const double *Q, const double *R, double *wrk, int *upidx,
                      int *bsidx, int *newtail) {
  const double thr = onw_threshold(K);
  const size_t nn = (size_t)dim * dim;
  const double prevnorm = 1.0e+99;
  double acc[MAXK];
  int skipup = 0, skipbs = 0, firstbs = 1, i, j, k, L;
  *newtail = degp1;
  for (i = tailidx; i < degp1; i++) {
    const double *xi = E(sol, (size_t)(i - 1) * dim);
    if (maxnorm(K, dim, xi) < thr) {
      skipup++;
    } else {
      for (j = i; j < degp1; j++) {
        const double *Aj = E(mat, (size_t)(j - i + 1) * nn);
        double *wj = E(rhs, (size_t)j * dim);
        for (k = 0; k < dim; k++)
          for (L = 0; L < dim; L++) {
            omd_mul(K, E(Aj, (size_t)k * dim + L), E(xi, L), acc);
            omd_sub(K, E(wj, k), acc, E(wj, k));
          }
      }
    }
    {
      double *x = E(sol, (size_t)i * dim);
      const double *b = E(rhs, (size_t)i * dim);
      const double nrm = maxnorm(K, dim, b);
      if (nrm < thr || nrm > prevnorm) {
        skipbs++;
        for (j = 0; j < dim; j++) omd_zero(K, E(x, j));
      } else {
        if (firstbs) {
          *newtail = i;
          firstbs = 0;
        }
        qrbs(K, dim, Q, R, b, x, wrk);
      }
    }
  }
  *upidx = skipup;
  *bsidx = skipbs;
}

static void qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                       double *sol, double *Q, double *R, double *wrk, int *zeroQ, int *noqr,
                       int *upidx, int *bsidx, int *newtail) {
  if (!(*noqr && !*zeroQ)) qrbs_head(K, dim, mat, rhs, sol, Q, R, wrk, zeroQ, noqr);
  if (degp1 > 1)
    qrbs_tail(K, dim, degp1, tailidx, mat, rhs, sol, Q, R, wrk, upidx, bsidx, newtail);
}

/* ---------------------------------------------------------- public API -- */

void onw_houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R) {
  double *a = malloc(sizeof(double) * K * nrows * ncols);
  double *q = malloc(sizeof(double) * K * nrows * nrows);
  double *r = malloc(sizeof(double) * K * nrows * ncols);
  from_planar(K, (size_t)nrows * ncols, A, a);
  houseqr(K, nrows, ncols, a, q, r);
  to_planar(K, (size_t)nrows * nrows, q, Q);
  to_planar(K, (size_t)nrows * ncols, r, R);
  free(a);
  free(q);
  free(r);
}

void onw_qrbs(int K, int dim, const double *Q, const double *R, const double *b, double *x) {
  const size_t nn = (size_t)dim * dim;
  double *q = malloc(sizeof(double) * K * nn), *r = malloc(sizeof(double) * K * nn);
  double *bb = malloc(sizeof(double) * K * dim), *xx = malloc(sizeof(double) * K * dim);
  double *w = malloc(sizeof(double) * K * dim);
  from_planar(K, nn, Q, q);
  from_planar(K, nn, R, r);
  from_planar(K, dim, b, bb);
  qrbs(K, dim, q, r, bb, xx, w);
  to_planar(K, dim, xx, x);
  free(q);
  free(r);
  free(bb);
  free(xx);
  free(w);
}

void onw_qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                    double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                    int *bsidx, int *newtail) {
  const size_t nn = (size_t)dim * dim, nv = (size_t)degp1 * dim;
  double *m = malloc(sizeof(double) * K * nn * degp1);
  double *b = malloc(sizeof(double) * K * nv), *x = malloc(sizeof(double) * K * nv);
  double *q = malloc(sizeof(double) * K * nn), *r = malloc(sizeof(double) * K * nn);
  double *w = malloc(sizeof(double) * K * dim);
  from_planar(K, nn * degp1, mat, m);
  from_planar(K, nv, rhs, b);
  from_planar(K, nv, sol, x);
  from_planar(K, nn, Q, q);
  from_planar(K, nn, R, r);
  qrbs_solve(K, dim, degp1, tailidx, m, b, x, q, r, w, zeroQ, noqr, upidx, bsidx, newtail);
  to_planar(K, nv, b, rhs);
  to_planar(K, nv, x, sol);
  to_planar(K, nn, q, Q);
  to_planar(K, nn, r, R);
  free(m);
  free(b);
  free(x);
  free(q);
  free(r);
  free(w);
}

void onw_linear_residue(int K, int dim, int degp1, int tailidx, const double *mat,
                        const double *rhs, const double *sol, double *resvec, double *resmax) {
  const size_t nn = (size_t)dim * dim, nv = (size_t)degp1 * dim;
  double *m = malloc(sizeof(double) * K * nn * degp1);
  double *b = malloc(sizeof(double) * K * nv), *x = malloc(sizeof(double) * K * nv);
  double *r = calloc((size_t)K * nv, sizeof(double));
  double acc[MAXK];
  int i, j, k, L, q;
  from_planar(K, nn * degp1, mat, m);
  from_planar(K, nv, rhs, b);
  from_planar(K, nv, sol, x);
  from_planar(K, nv, resvec, r);
  for (q = 0; q < K; q++) resmax[q] = 0.0;
  for (i = tailidx; i < degp1; i++) {
    double *ri = E(r, (size_t)i * dim);
    for (j = 0; j < dim; j++) omd_copy(K, E(b, (size_t)i * dim + j), E(ri, j));
    for (j = 0; j <= i - tailidx; j++) {
      const double *Aj = E(m, (size_t)j * nn);
      const double *xj = E(x, (size_t)(i - j) * dim);
      for (k = 0; k < dim; k++)
        for (L = 0; L < dim; L++) {
          omd_mul(K, E(Aj, (size_t)L * dim + k), E(xj, k), acc);
          omd_sub(K, E(ri, L), acc, E(ri, L));
        }
    }
    for (j = 0; j < dim; j++)
      if (fabs(E(ri, j)[0]) > resmax[0])
        for (q = 0; q < K; q++) resmax[q] = fabs(E(ri, j)[q]);
  }
  to_planar(K, nv, r, resvec);
  free(m);
  free(b);
  free(x);
  free(r);
}

int onw_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                   const int *idxptr, const int *idx, const int *exps, const double *cst,
                   const double *cff, double *input, int *wdeg_out, int *tail_out, int *up_out,
                   int *bs_out, double *history, double *Q, double *R) {
  const int D = deg + 1;
  const int nmono = monoptr[dim];
  const size_t nn = (size_t)dim * dim;
  /* interleaved working copies */
  double *x = malloc(sizeof(double) * K * dim * D);          /* x[i][k] */
  double *jac = malloc(sizeof(double) * K * nn * D);          /* jac[k][i][j] */
  double *rhs = malloc(sizeof(double) * K * dim * D);        /* rhs[k][i] */
  double *sol = malloc(sizeof(double) * K * dim * D);
  double *q = calloc((size_t)K * nn, sizeof(double)), *r = calloc((size_t)K * nn, sizeof(double));
  double *w = malloc(sizeof(double) * K * dim);
  /* planar buffers for the polynomial evaluations at working degree */
  double *pin = malloc(sizeof(double) * K * dim * D);
  double *pcst = malloc(sizeof(double) * K * D);
  double *pcff = malloc(sizeof(double) * K * (nmono > 0 ? nmono : 1) * D);
  double *pout = malloc(sizeof(double) * K * (dim + 1) * D);
  int *nv = malloc(sizeof(int) * (nmono > 0 ? nmono : 1));
  int *ip = malloc(sizeof(int) * (nmono > 0 ? nmono : 1));
  int tailidx = 1, wrkdeg = 0, zeroQ = 1, noqr = 0, upidx = 0, bsidx = 0, step, stepcnt = 0;
  int i, j, k, p;

  from_planar(K, (size_t)dim * D, input, x);
  for (step = 0; step < nbsteps; step++) {
    const int Dw = wrkdeg + 1;
    /* evaluate and differentiate every polynomial at degree wrkdeg */
    for (p = 0; p < K; p++)
      for (i = 0; i < dim; i++)
        for (k = 0; k < Dw; k++) pin[((size_t)p * dim + i) * Dw + k] = E(x, (size_t)i * D + k)[p];
    for (i = 0; i < dim; i++) {
      const int m0 = monoptr[i], nb = monoptr[i + 1] - monoptr[i];
      for (p = 0; p < K; p++) {
        for (k = 0; k < Dw; k++) pcst[(size_t)p * Dw + k] = cst[((size_t)p * dim + i) * D + k];
        for (j = 0; j < nb; j++)
          for (k = 0; k < Dw; k++)
            pcff[((size_t)p * nb + j) * Dw + k] = cff[((size_t)p * nmono + m0 + j) * D + k];
      }
      for (j = 0; j < nb; j++) {
        nv[j] = nvr[m0 + j];
        ip[j] = idxptr[m0 + j];
      }
      osr_poly_evaldiff_exp(K, 0, dim, nb, wrkdeg, nv, ip, idx, exps, pcst, pcff, pin, pout);
      /* rhs[k][i] = -value_i[k]; jac[k][i][j] = d value_i / d x_j [k] */
      for (k = 0; k < Dw; k++) {
        double *dst = E(rhs, (size_t)k * dim + i);
        for (p = 0; p < K; p++) dst[p] = -pout[((size_t)p * (dim + 1) + dim) * Dw + k];
        for (j = 0; j < dim; j++) {
          double *jd = E(jac, (size_t)k * nn + (size_t)i * dim + j);
          for (p = 0; p < K; p++) jd[p] = pout[((size_t)p * (dim + 1) + j) * Dw + k];
        }
      }
    }
    /* update_newton_qrstep: solve from zero, then update from tailidx-1 */
    memset(sol, 0, sizeof(double) * K * dim * Dw);
    {
      int newtail = tailidx;
      qrbs_solve(K, dim, Dw, tailidx, jac, rhs, sol, q, r, w, &zeroQ, &noqr, &upidx, &bsidx,
                 &newtail);
      tailidx = newtail;
    }
    for (k = tailidx - 1; k < Dw; k++)
      for (i = 0; i < dim; i++)
        omd_add(K, E(x, (size_t)i * D + k), E(sol, (size_t)k * dim + i), E(x, (size_t)i * D + k));
    if (wdeg_out) wdeg_out[step] = wrkdeg;
    if (tail_out) tail_out[step] = tailidx;
    if (up_out) up_out[step] = upidx;
    if (bs_out) bs_out[step] = bsidx;
    if (history) to_planar(K, (size_t)dim * D, x, history + (size_t)step * K * dim * D);
    stepcnt++;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  to_planar(K, (size_t)dim * D, x, input);
  if (Q) to_planar(K, nn, q, Q);
  if (R) to_planar(K, nn, r, R);
  free(x);
  free(jac);
  free(rhs);
  free(sol);
  free(q);
  free(r);
  free(w);
  free(pin);
  free(pcst);
  free(pcff);
  free(pout);
  free(nv);
  free(ip);
  return stepcnt;
}

/* ========================================================= column systems */

/* z_k = a_0 b_k + sum_{i=1..k} a_i b_{k-i}, k = 0..deg, interleaved series
 * (CPU_dbl2_product, Convolutions/dbl2_convolutions_host.cpp:7-26); z must
 * not overlap a or b */
static void iprod(int K, int deg, const double *a, const double *b, double *z) {
  double t[MAXK];
  int k, i;
  for (k = 0; k <= deg; k++) {
    double *acc = E(z, k);
    omd_mul(K, E(a, 0), E(b, k), acc);
    for (i = 1; i <= k; i++) {
      omd_mul(K, E(a, i), E(b, k - i), t);
      omd_add(K, acc, t, acc);
    }
  }
}

/* value and derivatives of the monomial cf * prod_j x_idx[j] at the series x
 * (series i at x + i*D numbers) into out (series r at out + r*D, row dim the
 * value, row idx[j] the derivative to x_idx[j]; other rows untouched), as
 * CPU_dbl2_evaldiff (Convolutions/dbl2_monomials_host.cpp:175-283): one
 * variable: x cf; two: (x_a x_b) cf, cf x_a, cf x_b; more: the forward,
 * backward and cross products of CPU_dbl2_speel (:28-87).  fw, bw, cr hold
 * nvr series of D numbers each. */
static void mono_evaldiff(int K, int dim, int nvr, int deg, int D, const int *ix, const double *cf,
                          const double *x, double *out, double *fw, double *bw, double *cr) {
  const size_t sD = (size_t)D * K;
#define XS(v) (x + (size_t)(v) * sD)
#define OS(r) (out + (size_t)(r) * sD)
#define WS(a, r) ((a) + (size_t)(r) * sD)
  int k, j;
  if (nvr == 1) {
    iprod(K, deg, XS(ix[0]), cf, OS(dim));
    for (k = 0; k <= deg; k++) omd_copy(K, E(cf, k), E(OS(ix[0]), k));
  } else if (nvr == 2) {
    iprod(K, deg, XS(ix[0]), XS(ix[1]), fw);
    iprod(K, deg, fw, cf, OS(dim));
    iprod(K, deg, cf, XS(ix[0]), OS(ix[1]));
    iprod(K, deg, cf, XS(ix[1]), OS(ix[0]));
  } else if (nvr > 2) {
    iprod(K, deg, cf, XS(ix[0]), WS(fw, 0));
    for (j = 1; j < nvr; j++) iprod(K, deg, WS(fw, j - 1), XS(ix[j]), WS(fw, j));
    iprod(K, deg, XS(ix[nvr - 1]), XS(ix[nvr - 2]), WS(bw, 0));
    for (j = 1; j < nvr - 2; j++) iprod(K, deg, WS(bw, j - 1), XS(ix[nvr - 2 - j]), WS(bw, j));
    iprod(K, deg, WS(bw, nvr - 3), cf, WS(cr, 0));
    memcpy(WS(bw, nvr - 3), WS(cr, 0), sizeof(double) * sD);
    if (nvr == 3) {
      iprod(K, deg, WS(fw, 0), XS(ix[2]), WS(cr, 0));
    } else {
      for (j = 0; j < nvr - 3; j++) iprod(K, deg, WS(fw, j), WS(bw, nvr - 4 - j), WS(cr, j));
      iprod(K, deg, WS(fw, nvr - 3), XS(ix[nvr - 1]), WS(cr, nvr - 3));
    }
    for (k = 0; k <= deg; k++) {
      omd_copy(K, E(WS(fw, nvr - 1), k), E(OS(dim), k));
      omd_copy(K, E(WS(fw, nvr - 2), k), E(OS(ix[nvr - 1]), k));
      omd_copy(K, E(WS(bw, nvr - 3), k), E(OS(ix[0]), k));
      for (j = 1; j < nvr - 1; j++) omd_copy(K, E(WS(cr, j - 1), k), E(OS(ix[j]), k));
    }
  }
#undef XS
#undef OS
#undef WS
}

int onw_column_newton(int K, int dim, int nbrcol, int deg, int nbsteps, const int *nvr,
                      const int *idxptr, const int *idx, const int *exps, const double *cff,
                      const double *mb, double *input, int *wdeg_out, int *tail_out,
                      int *up_out, int *bs_out, double *history, double *Q, double *R) {
  const int D = deg + 1;
  const int nmono = nbrcol * dim;
  const size_t nn = (size_t)dim * dim, sD = (size_t)D * K;
  double *x = malloc(sizeof(double) * K * dim * D);
  double *jac = malloc(sizeof(double) * K * nn * D);
  double *rhs = malloc(sizeof(double) * K * dim * D);
  double *sol = malloc(sizeof(double) * K * dim * D);
  double *q = calloc((size_t)K * nn, sizeof(double)), *r = calloc((size_t)K * nn, sizeof(double));
  double *w = malloc(sizeof(double) * K * dim);
  double *fval = malloc(sizeof(double) * K * dim * D);
  double *out = malloc(sizeof(double) * K * (dim + 1) * D);
  double *fw = malloc(sizeof(double) * K * dim * D), *bw = malloc(sizeof(double) * K * dim * D);
  double *cr = malloc(sizeof(double) * K * dim * D);
  double *cf = malloc(sizeof(double) * K * D), *acc = malloc(sizeof(double) * K * D);
  double *ci = malloc(sizeof(double) * K * (size_t)(nmono > 0 ? nmono : 1) * D);
  double *mbi = malloc(sizeof(double) * K * dim * D);
  int tailidx = 1, wrkdeg = 0, zeroQ = 1, noqr = 0, upidx = 0, bsidx = 0, step, stepcnt = 0;
  int i, j, k, c, e;

  from_planar(K, (size_t)dim * D, input, x);
  from_planar(K, (size_t)dim * D, mb, mbi);
  if (nbrcol > 1) from_planar(K, (size_t)nmono * D, cff, ci);
  for (step = 0; step < nbsteps; step++) {
    const int Dw = wrkdeg + 1;
    memset(jac, 0, sizeof(double) * K * nn * Dw);
    memset(fval, 0, sizeof(double) * K * dim * D);
    if (nbrcol == 1) {
      /* CPU_dbl2_evaluate_monomials (Newton/dbl2_systems_host.cpp:12-111)
       * on unit coefficients (dbl2_unit_series_vector each step), then
       * dbl2_linearize_evaldiff_output (:443-523) */
      for (i = 0; i < dim; i++) {
        const int *ix = idx + idxptr[i];
        const int *ex = exps ? exps + idxptr[i] : NULL;
        memset(cf, 0, sizeof(double) * K * Dw);
        cf[0] = 1.0;
        for (j = 0; ex && j < nvr[i]; j++)
          for (e = 1; e < ex[j]; e++) {
            iprod(K, wrkdeg, x + (size_t)ix[j] * sD, cf, acc);
            memcpy(cf, acc, sizeof(double) * K * Dw);
          }
        mono_evaldiff(K, dim, nvr[i], wrkdeg, D, ix, cf, x, out, fw, bw, cr);
        for (j = 0; ex && j < nvr[i]; j++)
          if (ex[j] > 1)
            for (k = 0; k < Dw; k++) {
              double *o = E(out + (size_t)ix[j] * sD, k);
              omd_mul_d(K, o, (double)ex[j], o);
            }
        for (k = 0; k < Dw; k++) {
          omd_copy(K, E(out + (size_t)dim * sD, k), E(fval, (size_t)i * D + k));
          for (j = 0; j < nvr[i]; j++)
            omd_copy(K, E(out + (size_t)ix[j] * sD, k),
                     E(jac, (size_t)k * nn + (size_t)i * dim + ix[j]));
        }
      }
    } else {
      /* CPU_dbl2_evaluate_columns (Newton/dbl2_systems_host.cpp:232-329):
       * -1 is the constant of the last equation */
      fval[(size_t)(dim - 1) * D * K] = -1.0;
      for (c = 0; c < nbrcol; c++)
        for (j = 0; j < dim; j++) {
          const int m = c * dim + j;
          const int *ix = idx + idxptr[m];
          int v;
          if (nvr[m] <= 0) continue;
          mono_evaldiff(K, dim, nvr[m], wrkdeg, D, ix, ci + (size_t)m * sD, x, out, fw, bw, cr);
          for (k = 0; k < Dw; k++) {
            double *fv = E(fval, (size_t)j * D + k);
            omd_add(K, fv, E(out + (size_t)dim * sD, k), fv);
          }
          for (v = 0; v < nvr[m]; v++)
            for (k = 0; k < Dw; k++) {
              double *jd = E(jac, (size_t)k * nn + (size_t)j * dim + ix[v]);
              omd_add(K, jd, E(out + (size_t)ix[v] * sD, k), jd);
            }
        }
    }
    /* rhs[k][i] = -(funval[i][k] - mb[i][k]) (dbl2_define_rhs, :626-659) */
    for (k = 0; k < Dw; k++)
      for (i = 0; i < dim; i++) {
        double *dst = E(rhs, (size_t)k * dim + i);
        omd_sub(K, E(fval, (size_t)i * D + k), E(mbi, (size_t)i * D + k), dst);
        omd_neg(K, dst, dst);
      }
    memset(sol, 0, sizeof(double) * K * dim * Dw);
    {
      int newtail = tailidx;
      qrbs_solve(K, dim, Dw, tailidx, jac, rhs, sol, q, r, w, &zeroQ, &noqr, &upidx, &bsidx,
                 &newtail);
      tailidx = newtail;
    }
    for (k = tailidx - 1; k < Dw; k++)
      for (i = 0; i < dim; i++)
        omd_add(K, E(x, (size_t)i * D + k), E(sol, (size_t)k * dim + i), E(x, (size_t)i * D + k));
    if (wdeg_out) wdeg_out[step] = wrkdeg;
    if (tail_out) tail_out[step] = tailidx;
    if (up_out) up_out[step] = upidx;
    if (bs_out) bs_out[step] = bsidx;
    if (history) to_planar(K, (size_t)dim * D, x, history + (size_t)step * K * dim * D);
    stepcnt++;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  to_planar(K, (size_t)dim * D, x, input);
  if (Q) to_planar(K, nn, q, Q);
  if (R) to_planar(K, nn, r, R);
  free(x); free(jac); free(rhs); free(sol); free(q); free(r); free(w); free(fval); free(out);
  free(fw); free(bw); free(cr); free(cf); free(acc); free(ci); free(mbi);
  return stepcnt;
}

/* ================================================================ complex */

/* complex number i: 2K doubles, K real limbs then K imaginary limbs */
#define ZE(a, i) ((a) + (size_t)(i) * 2 * K)
#define ZR(z) (z)
#define ZI(z) ((z) + K)

/* CPU_cmplx2_factors_rightQupdate: Q -= (beta Q v) v^H on columns k.. */
static void cx_right_q_update(int K, int n, int k, double *Q, const double *v, const double *beta,
                              double *w) {
  double zre[MAXK], zim[MAXK], acc[MAXK];
  int i, j;
  for (i = 0; i < n; i++) {
    double *wi = ZE(w, i);
    omd_zero(K, ZR(wi));
    omd_zero(K, ZI(wi));
    for (j = k; j < n; j++) {
      const double *q = ZE(Q, (size_t)i * n + j), *vj = ZE(v, j - k);
      omd_mul(K, ZR(q), ZR(vj), zre);
      omd_mul(K, ZI(q), ZI(vj), acc);
      omd_sub(K, zre, acc, zre);
      omd_mul(K, ZI(q), ZR(vj), zim);
      omd_mul(K, ZR(q), ZI(vj), acc);
      omd_add(K, zim, acc, zim);
      omd_add(K, ZR(wi), zre, ZR(wi));
      omd_add(K, ZI(wi), zim, ZI(wi));
    }
    omd_mul(K, beta, ZR(wi), ZR(wi));
    omd_mul(K, beta, ZI(wi), ZI(wi));
  }
  for (i = 0; i < n; i++)
    for (j = k; j < n; j++) {
      double *q = ZE(Q, (size_t)i * n + j);
      const double *wi = ZE(w, i), *vj = ZE(v, j - k);
      omd_mul(K, ZR(wi), ZR(vj), zre);
      omd_mul(K, ZI(wi), ZI(vj), acc);
      omd_add(K, zre, acc, zre);
      omd_mul(K, ZI(wi), ZR(vj), zim);
      omd_mul(K, ZR(wi), ZI(vj), acc);
      omd_sub(K, zim, acc, zim);
      omd_sub(K, ZR(q), zre, ZR(q));
      omd_sub(K, ZI(q), zim, ZI(q));
    }
}

static void cx_houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R) {
  const size_t z = sizeof(double) * 2 * K;
  const int big = nrows > ncols ? nrows : ncols;
  double *x = malloc(z * nrows), *v = malloc(z * nrows), *w = malloc(z * big);
  double beta[MAXK];
  int i, j, k;
  for (i = 0; i < nrows; i++) {
    for (j = 0; j < nrows; j++) memset(ZE(Q, (size_t)i * nrows + j), 0, z);
    omd_from_d(K, 1.0, ZR(ZE(Q, (size_t)i * nrows + i)));
    for (j = 0; j < ncols; j++) memcpy(ZE(R, (size_t)i * ncols + j), ZE(A, (size_t)i * ncols + j), z);
  }
  for (k = 0; k < ncols; k++) {
    if (nrows - k > 0) {
      for (i = k; i < nrows; i++) memcpy(ZE(x, i - k), ZE(R, (size_t)i * ncols + k), z);
      obq_cx_house(K, nrows - k, x, v, beta);
      obq_cx_left_r_update(K, nrows, ncols, ncols, k, R, v, beta, w);
      cx_right_q_update(K, nrows, k, Q, v, beta, w);
    }
  }
  free(x);
  free(v);
  free(w);
}

/* x = R^{-1} Q^H b */
static void cx_qrbs(int K, int dim, const double *Q, const double *R, const double *b, double *x,
                    double *wrk) {
  double acc[MAXK];
  int i, j;
  for (i = 0; i < dim; i++) {
    double *wi = ZE(wrk, i);
    omd_zero(K, ZR(wi));
    omd_zero(K, ZI(wi));
    for (j = 0; j < dim; j++) {
      const double *q = ZE(Q, (size_t)j * dim + i), *bj = ZE(b, j);
      omd_mul(K, ZR(q), ZR(bj), acc);
      omd_add(K, ZR(wi), acc, ZR(wi));
      omd_mul(K, ZI(q), ZI(bj), acc);
      omd_add(K, ZR(wi), acc, ZR(wi));
      omd_mul(K, ZR(q), ZI(bj), acc);
      omd_add(K, ZI(wi), acc, ZI(wi));
      omd_mul(K, ZI(q), ZR(bj), acc);
      omd_sub(K, ZI(wi), acc, ZI(wi));
    }
  }
  obq_cx_backward(K, dim, dim, R, wrk, x);
}

/* max of |re| + |im| of the leading limbs (NaN only when v[0]'s is) */
static double cx_maxnorm(int K, int dim, const double *v) {
  double r = fabs(ZR(v)[0]) + fabs(ZI(v)[0]);
  int i;
  for (i = 1; i < dim; i++) {
    const double a = fabs(ZR(ZE(v, i))[0]) + fabs(ZI(ZE(v, i))[0]);
    if (a > r) r = a;
  }
  return r;
}

static void cx_qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                          double *sol, double *Q, double *R, double *wrk, int *zeroQ, int *noqr,
                          int *upidx, int *bsidx, int *newtail) {
  const double thr = onw_threshold(K);
  const size_t nn = (size_t)dim * dim;
  double acc[MAXK];
  int i, j, k, L;
  if (!(*noqr && !*zeroQ)) {
    if (cx_maxnorm(K, dim, rhs) < thr) *noqr = *zeroQ ? 0 : 1;
    if (!*noqr) {
      cx_houseqr(K, dim, dim, mat, Q, R);
      *zeroQ = 0;
      cx_qrbs(K, dim, Q, R, rhs, sol, wrk);
    }
  }
  if (degp1 > 1) {
    int skipup = 0, skipbs = 0, firstbs = 1;
    *newtail = degp1;
    for (i = tailidx; i < degp1; i++) {
      const double *xi = ZE(sol, (size_t)(i - 1) * dim);
      if (cx_maxnorm(K, dim, xi) < thr) {
        skipup++;
      } else {
        for (j = i; j < degp1; j++) {
          const double *Aj = ZE(mat, (size_t)(j - i + 1) * nn);
          double *wj = ZE(rhs, (size_t)j * dim);
          for (k = 0; k < dim; k++)
            for (L = 0; L < dim; L++) {
              const double *a = ZE(Aj, (size_t)k * dim + L), *x = ZE(xi, L);
              double *wk = ZE(wj, k);
              omd_mul(K, ZR(a), ZR(x), acc);
              omd_sub(K, ZR(wk), acc, ZR(wk));
              omd_mul(K, ZI(a), ZI(x), acc);
              omd_add(K, ZR(wk), acc, ZR(wk));
              omd_mul(K, ZR(a), ZI(x), acc);
              omd_sub(K, ZI(wk), acc, ZI(wk));
              omd_mul(K, ZI(a), ZR(x), acc);
              omd_sub(K, ZI(wk), acc, ZI(wk));
            }
        }
      }
      {
        double *x = ZE(sol, (size_t)i * dim);
        const double *b = ZE(rhs, (size_t)i * dim);
        const double nrm = cx_maxnorm(K, dim, b);
        if (nrm < thr || nrm > 1.0e+99) {
          skipbs++;
          memset(x, 0, sizeof(double) * 2 * K * dim);
        } else {
          if (firstbs) {
            *newtail = i;
            firstbs = 0;
          }
          cx_qrbs(K, dim, Q, R, b, x, wrk);
        }
      }
    }
    *upidx = skipup;
    *bsidx = skipbs;
  }
}

void onw_cx_houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R) {
  const int P = 2 * K;
  double *a = malloc(sizeof(double) * P * nrows * ncols);
  double *q = malloc(sizeof(double) * P * nrows * nrows);
  double *r = malloc(sizeof(double) * P * nrows * ncols);
  from_planar(P, (size_t)nrows * ncols, A, a);
  cx_houseqr(K, nrows, ncols, a, q, r);
  to_planar(P, (size_t)nrows * nrows, q, Q);
  to_planar(P, (size_t)nrows * ncols, r, R);
  free(a);
  free(q);
  free(r);
}

void onw_cx_qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                       double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                       int *bsidx, int *newtail) {
  const int P = 2 * K;
  const size_t nn = (size_t)dim * dim, nv = (size_t)degp1 * dim;
  double *m = malloc(sizeof(double) * P * nn * degp1);
  double *b = malloc(sizeof(double) * P * nv), *x = malloc(sizeof(double) * P * nv);
  double *q = malloc(sizeof(double) * P * nn), *r = malloc(sizeof(double) * P * nn);
  double *w = malloc(sizeof(double) * P * dim);
  from_planar(P, nn * degp1, mat, m);
  from_planar(P, nv, rhs, b);
  from_planar(P, nv, sol, x);
  from_planar(P, nn, Q, q);
  from_planar(P, nn, R, r);
  cx_qrbs_solve(K, dim, degp1, tailidx, m, b, x, q, r, w, zeroQ, noqr, upidx, bsidx, newtail);
  to_planar(P, nv, b, rhs);
  to_planar(P, nv, x, sol);
  to_planar(P, nn, q, Q);
  to_planar(P, nn, r, R);
  free(m);
  free(b);
  free(x);
  free(q);
  free(r);
  free(w);
}

void onw_cx_linear_residue(int K, int dim, int degp1, int tailidx, const double *mat,
                           const double *rhs, const double *sol, double *resvec, double *resmax) {
  const int P = 2 * K;
  const size_t nn = (size_t)dim * dim, nv = (size_t)degp1 * dim;
  double *m = malloc(sizeof(double) * P * nn * degp1);
  double *b = malloc(sizeof(double) * P * nv), *x = malloc(sizeof(double) * P * nv);
  double *r = calloc((size_t)P * nv, sizeof(double));
  double acc[MAXK];
  int i, j, k, L, q;
  from_planar(P, nn * degp1, mat, m);
  from_planar(P, nv, rhs, b);
  from_planar(P, nv, sol, x);
  from_planar(P, nv, resvec, r);
  for (q = 0; q < K; q++) resmax[q] = 0.0;
  for (i = tailidx; i < degp1; i++) {
    double *ri = ZE(r, (size_t)i * dim);
    memcpy(ri, ZE(b, (size_t)i * dim), sizeof(double) * P * dim);
    for (j = 0; j <= i - tailidx; j++) {
      const double *Aj = ZE(m, (size_t)j * nn);
      const double *xj = ZE(x, (size_t)(i - j) * dim);
      for (k = 0; k < dim; k++)
        for (L = 0; L < dim; L++) {
          const double *a = ZE(Aj, (size_t)L * dim + k), *xk = ZE(xj, k);
          double *rl = ZE(ri, L);
          omd_mul(K, ZR(a), ZR(xk), acc);
          omd_sub(K, ZR(rl), acc, ZR(rl));
          omd_mul(K, ZI(a), ZI(xk), acc);
          omd_add(K, ZR(rl), acc, ZR(rl));
          omd_mul(K, ZR(a), ZI(xk), acc);
          omd_sub(K, ZI(rl), acc, ZI(rl));
          omd_mul(K, ZI(a), ZR(xk), acc);
          omd_sub(K, ZI(rl), acc, ZI(rl));
        }
    }
    for (j = 0; j < dim; j++) {
      const double *rj = ZE(ri, j);
      if (fabs(ZR(rj)[0]) + fabs(ZI(rj)[0]) > resmax[0])
        for (q = 0; q < K; q++) resmax[q] = fabs(ZR(rj)[q]) + fabs(ZI(rj)[q]);
    }
  }
  to_planar(P, nv, r, resvec);
  free(m);
  free(b);
  free(x);
  free(r);
}

int onw_cx_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                      const int *idxptr, const int *idx, const double *cst, const double *cff,
                      double *input, int *wdeg_out, int *tail_out, int *up_out, int *bs_out,
                      double *history, double *Q, double *R) {
  const int D = deg + 1, P = 2 * K;
  const int nmono = monoptr[dim];
  const size_t nn = (size_t)dim * dim;
  double *x = malloc(sizeof(double) * P * dim * D);
  double *jac = malloc(sizeof(double) * P * nn * D);
  double *rhs = malloc(sizeof(double) * P * dim * D);
  double *sol = malloc(sizeof(double) * P * dim * D);
  double *q = calloc((size_t)P * nn, sizeof(double)), *r = calloc((size_t)P * nn, sizeof(double));
  double *w = malloc(sizeof(double) * P * dim);
  double *pin = malloc(sizeof(double) * P * dim * D);
  double *pcst = malloc(sizeof(double) * P * D);
  double *pcff = malloc(sizeof(double) * P * (nmono > 0 ? nmono : 1) * D);
  double *pout = malloc(sizeof(double) * P * (dim + 1) * D);
  int *nv = malloc(sizeof(int) * (nmono > 0 ? nmono : 1));
  int *ip = malloc(sizeof(int) * (nmono > 0 ? nmono : 1));
  int tailidx = 1, wrkdeg = 0, zeroQ = 1, noqr = 0, upidx = 0, bsidx = 0, step, stepcnt = 0;
  int i, j, k, p;
  from_planar(P, (size_t)dim * D, input, x);
  for (step = 0; step < nbsteps; step++) {
    const int Dw = wrkdeg + 1;
    for (p = 0; p < P; p++)
      for (i = 0; i < dim; i++)
        for (k = 0; k < Dw; k++) pin[((size_t)p * dim + i) * Dw + k] = ZE(x, (size_t)i * D + k)[p];
    for (i = 0; i < dim; i++) {
      const int m0 = monoptr[i], nb = monoptr[i + 1] - monoptr[i];
      for (p = 0; p < P; p++) {
        for (k = 0; k < Dw; k++) pcst[(size_t)p * Dw + k] = cst[((size_t)p * dim + i) * D + k];
        for (j = 0; j < nb; j++)
          for (k = 0; k < Dw; k++)
            pcff[((size_t)p * nb + j) * Dw + k] = cff[((size_t)p * nmono + m0 + j) * D + k];
      }
      for (j = 0; j < nb; j++) {
        nv[j] = nvr[m0 + j];
        ip[j] = idxptr[m0 + j];
      }
      osr_poly_evaldiff(K, 1, dim, nb, wrkdeg, nv, ip, idx, pcst, pcff, pin, pout);
      for (k = 0; k < Dw; k++) {
        double *dst = ZE(rhs, (size_t)k * dim + i);
        for (p = 0; p < P; p++) dst[p] = -pout[((size_t)p * (dim + 1) + dim) * Dw + k];
        for (j = 0; j < dim; j++) {
          double *jd = ZE(jac, (size_t)k * nn + (size_t)i * dim + j);
          for (p = 0; p < P; p++) jd[p] = pout[((size_t)p * (dim + 1) + j) * Dw + k];
        }
      }
    }
    memset(sol, 0, sizeof(double) * P * dim * Dw);
    {
      int newtail = tailidx;
      cx_qrbs_solve(K, dim, Dw, tailidx, jac, rhs, sol, q, r, w, &zeroQ, &noqr, &upidx, &bsidx,
                    &newtail);
      tailidx = newtail;
    }
    for (k = tailidx - 1; k < Dw; k++)
      for (i = 0; i < dim; i++) {
        double *xi = ZE(x, (size_t)i * D + k);
        const double *d = ZE(sol, (size_t)k * dim + i);
        omd_add(K, ZR(xi), ZR(d), ZR(xi));
        omd_add(K, ZI(xi), ZI(d), ZI(xi));
      }
    if (wdeg_out) wdeg_out[step] = wrkdeg;
    if (tail_out) tail_out[step] = tailidx;
    if (up_out) up_out[step] = upidx;
    if (bs_out) bs_out[step] = bsidx;
    if (history) to_planar(P, (size_t)dim * D, x, history + (size_t)step * P * dim * D);
    stepcnt++;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  to_planar(P, (size_t)dim * D, x, input);
  if (Q) to_planar(P, nn, q, Q);
  if (R) to_planar(P, nn, r, R);
  free(x);
  free(jac);
  free(rhs);
  free(sol);
  free(q);
  free(r);
  free(w);
  free(pin);
  free(pcst);
  free(pcff);
  free(pout);
  free(nv);
  free(ip);
  return stepcnt;
}
