/* CPU oracle: blocked Householder QR and tiled back substitution of complex
 * multiple-double matrices (test infrastructure only).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (phcpack_amd/) may link or
 * call this code; only tests/ and the benchmark's checks use it.
 *
 * Restates, K-generic (K = 2, 4, 8 limbs), with the operation order of the
 * reference's host code (quad-double names; the double-double and
 * octo-double files repeat the sequences):
 *   complex Householder vector  CPU_cmplx2_factors_house (Matrices/dbl2_factorizations.cpp:441-533)
 *   panel column update         CPU_cmplx2_factors_leftRupdate (:567-627)
 *   W of the WY representation  CPU_cmplx4_blocked_VB_to_W (Matrices/dbl4_baqr_host.cpp:91-240)
 *   trailing R update           CPU_cmplx4_blocked_leftRupdate (:357-585)
 *   Q update                    CPU_cmplx4_blocked_rightQupdate (:743-1032)
 *   blocked QR                  CPU_cmplx4_blocked_houseqr (:1206-1437)
 *   back substitution           CPU_cmplx2_factors_backward (Matrices/dbl2_factorizations.cpp:142-200)
 *   tile inverse, tiled solver  CPU_cmplx2_upper_inverse, CPU_cmplx2_upper_tiled_solver
 *                               (Matrices/dbl2_tabs_host.cpp:65-106, 289-513)
 *   Q^H b                       CPU_cmplx2_factors_qrbs (Matrices/dbl2_factorizations.cpp:842-879)
 *   real blocked QR             CPU_dbl4_blocked_houseqr (Matrices/dbl4_baqr_host.cpp:1033-1205)
 *     with CPU_dbl4_factors_house, CPU_dbl4_factors_leftRupdate
 *     (Matrices/dbl4_factorizations.cpp:541-626, 788-832), CPU_dbl4_blocked_VB_to_W,
 *     _leftRupdate, _rightQupdate (dbl4_baqr_host.cpp:13-86, 242-356, 586-742)
 *
 * Parity pinned by tests/test_oracle_baqr.py: bit-exact against the
 * reference functions compiled in oracle/_ref (K = 2, 4, 8), and Q^H Q = I,
 * QR = A, U x = b for planted x.
 */
#include <stdlib.h>
#include <string.h>
#include "md.h"
#include "baqr.h"

#define MAXK 8
/* complex number z: K real limbs then K imaginary limbs (2K doubles) */
#define Z(a, i) ((a) + (size_t)(i) * 2 * K)
#define RE(z) (z)
#define IM(z) ((z) + K)

static void cx_from_planar(int K, size_t n, const double *p, double *a) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < 2 * K; q++) a[j * 2 * K + q] = p[q * n + j];
}

static void cx_to_planar(int K, size_t n, const double *a, double *p) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < 2 * K; q++) p[q * n + j] = a[j * 2 * K + q];
}

static int all_zero(int K, const double *a) {
  int q;
  for (q = 0; q < K; q++)
    if (a[q] != 0.0) return 0;
  return 1;
}

/* acc = a*b; c += acc  /  c -= acc (the inc/dec pairs of the reference) */
static void mac(int K, const double *a, const double *b, double *c) {
  double t[MAXK];
  omd_mul(K, a, b, t);
  omd_add(K, c, t, c);
}

static void msb(int K, const double *a, const double *b, double *c) {
  double t[MAXK];
  omd_mul(K, a, b, t);
  omd_sub(K, c, t, c);
}

static void cx_house(int K, int n, const double *x, double *v, double *beta) {
  double sigma[MAXK], acc[MAXK], mu[MAXK], sqrx0[MAXK], x0rad[MAXK], sqrv0[MAXK];
  double inv0re[MAXK], inv0im[MAXK], zre[MAXK], zim[MAXK];
  int i;
  omd_zero(K, sigma);
  omd_from_d(K, 1.0, RE(Z(v, 0)));
  omd_zero(K, IM(Z(v, 0)));
  for (i = 1; i < n; i++) {
    omd_sqr(K, RE(Z(x, i)), acc);
    omd_add(K, sigma, acc, sigma);
    omd_sqr(K, IM(Z(x, i)), acc);
    omd_add(K, sigma, acc, sigma);
    memcpy(Z(v, i), Z(x, i), sizeof(double) * 2 * K);
  }
  if (all_zero(K, sigma)) {
    omd_zero(K, beta);
    return;
  }
  omd_sqr(K, RE(Z(x, 0)), sqrx0);
  omd_sqr(K, IM(Z(x, 0)), acc);
  omd_add(K, sqrx0, acc, sqrx0);
  omd_sqrt(K, sqrx0, x0rad);
  omd_add(K, sqrx0, sigma, sqrx0);
  omd_sqrt(K, sqrx0, mu);
  if (all_zero(K, x0rad)) {
    omd_neg(K, mu, RE(Z(v, 0)));
    omd_zero(K, IM(Z(v, 0)));
  } else {
    omd_div(K, mu, x0rad, mu);
    omd_mul(K, mu, RE(Z(x, 0)), acc);
    omd_sub(K, RE(Z(x, 0)), acc, RE(Z(v, 0)));
    omd_mul(K, mu, IM(Z(x, 0)), acc);
    omd_sub(K, IM(Z(x, 0)), acc, IM(Z(v, 0)));
  }
  omd_sqr(K, RE(Z(v, 0)), sqrv0);
  omd_sqr(K, IM(Z(v, 0)), acc);
  omd_add(K, sqrv0, acc, sqrv0);
  omd_add(K, sigma, sqrv0, sigma);
  omd_div(K, sqrv0, sigma, beta);
  omd_mul_d(K, beta, 2.0, beta);
  omd_div(K, RE(Z(v, 0)), sqrv0, inv0re);
  omd_div(K, IM(Z(v, 0)), sqrv0, inv0im);
  omd_neg(K, inv0im, inv0im);
  for (i = 1; i < n; i++) {
    double *vi = Z(v, i);
    omd_mul(K, RE(vi), inv0re, zre);
    msb(K, IM(vi), inv0im, zre);
    omd_mul(K, IM(vi), inv0re, zim);
    mac(K, RE(vi), inv0im, zim);
    memcpy(RE(vi), zre, sizeof(double) * K);
    memcpy(IM(vi), zim, sizeof(double) * K);
  }
  omd_from_d(K, 1.0, RE(Z(v, 0)));
  omd_zero(K, IM(Z(v, 0)));
}

/* R (nrows x ldr) columns k..ncols-1 from row k: R -= v (beta v^H R) */
static void cx_left_r_update(int K, int nrows, int ncols, int ldr, int k, double *R,
                             const double *v, const double *beta, double *w) {
  double zre[MAXK], zim[MAXK];
  int i, j;
  for (j = k; j < ncols; j++) {
    double *wj = Z(w, j - k);
    omd_zero(K, RE(wj));
    omd_zero(K, IM(wj));
    for (i = k; i < nrows; i++) {
      const double *r = Z(R, (size_t)i * ldr + j), *vi = Z(v, i - k);
      omd_mul(K, RE(r), RE(vi), zre);
      mac(K, IM(r), IM(vi), zre);
      omd_mul(K, RE(r), IM(vi), zim);
      msb(K, IM(r), RE(vi), zim);
      omd_add(K, RE(wj), zre, RE(wj));
      omd_add(K, IM(wj), zim, IM(wj));
    }
    omd_mul(K, beta, RE(wj), RE(wj));
    omd_mul(K, beta, IM(wj), IM(wj));
  }
  for (i = k; i < nrows; i++)
    for (j = k; j < ncols; j++) {
      double *r = Z(R, (size_t)i * ldr + j);
      const double *vi = Z(v, i - k), *wj = Z(w, j - k);
      omd_mul(K, RE(vi), RE(wj), zre);
      mac(K, IM(vi), IM(wj), zre);
      omd_mul(K, IM(vi), RE(wj), zim);
      msb(K, RE(vi), IM(wj), zim);
      omd_sub(K, RE(r), zre, RE(r));
      omd_sub(K, IM(r), zim, IM(r));
    }
}

/* W (szt columns of rowdim) from the betas B and Householder vectors V */
static void vb_to_w(int K, int rowdim, int szt, const double *B, const double *V, double *W,
                    double *p) {
  double zre[MAXK], zim[MAXK];
  int i, j, k;
  for (i = 0; i < rowdim; i++) {
    const double *vi = Z(V, i);
    double *wi = Z(W, i);
    omd_mul(K, B, RE(vi), RE(wi));
    omd_neg(K, RE(wi), RE(wi));
    omd_mul(K, B, IM(vi), IM(wi));
    omd_neg(K, IM(wi), IM(wi));
  }
  for (j = 1; j < szt; j++) {
    const double *vj = Z(V, (size_t)j * rowdim);
    for (k = 0; k < j; k++) {
      double *pk = Z(p, k);
      const double *vk = Z(V, (size_t)k * rowdim);
      omd_zero(K, RE(pk));
      omd_zero(K, IM(pk));
      for (i = 0; i < rowdim; i++) {
        mac(K, RE(Z(vk, i)), RE(Z(vj, i)), RE(pk));
        mac(K, IM(Z(vk, i)), IM(Z(vj, i)), RE(pk));
        msb(K, IM(Z(vk, i)), RE(Z(vj, i)), IM(pk));
        mac(K, RE(Z(vk, i)), IM(Z(vj, i)), IM(pk));
      }
    }
    for (i = 0; i < rowdim; i++) {
      omd_zero(K, zre);
      omd_zero(K, zim);
      for (k = 0; k < j; k++) {
        const double *wk = Z(W, (size_t)k * rowdim + i), *pk = Z(p, k);
        mac(K, RE(wk), RE(pk), zre);
        msb(K, IM(wk), IM(pk), zre);
        mac(K, IM(wk), RE(pk), zim);
        mac(K, RE(wk), IM(pk), zim);
      }
      omd_add(K, zre, RE(Z(vj, i)), zre);
      omd_add(K, zim, IM(Z(vj, i)), zim);
      omd_mul(K, B + (size_t)j * K, zre, RE(Z(W, (size_t)j * rowdim + i)));
      omd_neg(K, RE(Z(W, (size_t)j * rowdim + i)), RE(Z(W, (size_t)j * rowdim + i)));
      omd_mul(K, B + (size_t)j * K, zim, IM(Z(W, (size_t)j * rowdim + i)));
      omd_neg(K, IM(Z(W, (size_t)j * rowdim + i)), IM(Z(W, (size_t)j * rowdim + i)));
    }
  }
}

/* C = C + (Y W^H) C on rows rowoff.., columns coloff.. of C (nrows x ncols) */
static void blocked_left_r(int K, int nrows, int ncols, int szt, int idx, double *C,
                           const double *Y, const double *W) {
  const int rowoff = idx * szt, rowdim = nrows - rowoff, coloff = (idx + 1) * szt;
  const int coldim = ncols - coloff;
  double *ywt = malloc(sizeof(double) * 2 * K * (size_t)rowdim * rowdim);
  double *prd = malloc(sizeof(double) * 2 * K * (size_t)rowdim * (coldim > 0 ? coldim : 1));
  int i, j, k;
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = Z(ywt, (size_t)i * rowdim + j);
      omd_zero(K, RE(e));
      omd_zero(K, IM(e));
      for (k = 0; k < szt; k++) {
        const double *y = Z(Y, (size_t)k * rowdim + i), *w = Z(W, (size_t)k * rowdim + j);
        mac(K, RE(y), RE(w), RE(e));
        mac(K, IM(y), IM(w), RE(e));
        mac(K, IM(y), RE(w), IM(e));
        msb(K, RE(y), IM(w), IM(e));
      }
    }
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < coldim; j++) {
      double *e = Z(prd, (size_t)i * coldim + j);
      omd_zero(K, RE(e));
      omd_zero(K, IM(e));
      for (k = 0; k < rowdim; k++) {
        const double *a = Z(ywt, (size_t)i * rowdim + k);
        const double *c = Z(C, (size_t)(rowoff + k) * ncols + coloff + j);
        mac(K, RE(a), RE(c), RE(e));
        msb(K, IM(a), IM(c), RE(e));
        mac(K, IM(a), RE(c), IM(e));
        mac(K, RE(a), IM(c), IM(e));
      }
    }
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < coldim; j++) {
      double *c = Z(C, (size_t)(rowoff + i) * ncols + coloff + j);
      const double *e = Z(prd, (size_t)i * coldim + j);
      omd_add(K, RE(c), RE(e), RE(c));
      omd_add(K, IM(c), IM(e), IM(c));
    }
  free(ywt);
  free(prd);
}

/* Q = Q + Q (W Y^H) on columns coloff.. of Q (dim x dim) */
static void blocked_right_q(int K, int dim, int szt, int idx, double *Q, const double *Y,
                            const double *W) {
  const int coloff = idx * szt, rowdim = dim - coloff;
  double *wyt = malloc(sizeof(double) * 2 * K * (size_t)rowdim * rowdim);
  double *prd = malloc(sizeof(double) * 2 * K * (size_t)dim * rowdim);
  int i, j, k;
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = Z(wyt, (size_t)i * rowdim + j);
      omd_zero(K, RE(e));
      omd_zero(K, IM(e));
      for (k = 0; k < szt; k++) {
        const double *w = Z(W, (size_t)k * rowdim + i), *y = Z(Y, (size_t)k * rowdim + j);
        mac(K, RE(w), RE(y), RE(e));
        mac(K, IM(w), IM(y), RE(e));
        mac(K, IM(w), RE(y), IM(e));
        msb(K, RE(w), IM(y), IM(e));
      }
    }
  for (i = 0; i < dim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = Z(prd, (size_t)i * rowdim + j);
      omd_zero(K, RE(e));
      omd_zero(K, IM(e));
      for (k = 0; k < rowdim; k++) {
        const double *q = Z(Q, (size_t)i * dim + coloff + k);
        const double *a = Z(wyt, (size_t)k * rowdim + j);
        mac(K, RE(q), RE(a), RE(e));
        msb(K, IM(q), IM(a), RE(e));
        mac(K, IM(q), RE(a), IM(e));
        mac(K, RE(q), IM(a), IM(e));
      }
    }
  for (i = 0; i < dim; i++)
    for (j = 0; j < rowdim; j++) {
      double *q = Z(Q, (size_t)i * dim + coloff + j);
      const double *e = Z(prd, (size_t)i * rowdim + j);
      omd_add(K, RE(q), RE(e), RE(q));
      omd_add(K, IM(q), IM(e), IM(q));
    }
  free(wyt);
  free(prd);
}

static void blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                            double *Q, double *R) {
  const size_t z = sizeof(double) * 2 * K;
  double *x = malloc(z * nrows), *v = malloc(z * nrows), *w = malloc(z * (nrows + ncols));
  double *B = malloc(sizeof(double) * K * szt);
  double *Y = malloc(z * (size_t)szt * nrows), *W = malloc(z * (size_t)szt * nrows);
  double *p = malloc(z * szt);
  double beta[MAXK];
  int i, j, k, L;
  for (i = 0; i < nrows; i++) {
    for (j = 0; j < nrows; j++) memset(Z(Q, (size_t)i * nrows + j), 0, z);
    omd_from_d(K, 1.0, RE(Z(Q, (size_t)i * nrows + i)));
    for (j = 0; j < ncols; j++) memcpy(Z(R, (size_t)i * ncols + j), Z(A, (size_t)i * ncols + j), z);
  }
  for (k = 0; k < nbt; k++) {
    const int rowdim = nrows - k * szt;
    for (L = 0; L < szt; L++) {
      const int colidx = k * szt + L, endcol = (k + 1) * szt, nrowscol = nrows - colidx;
      for (i = colidx; i < nrows; i++) memcpy(Z(x, i - colidx), Z(R, (size_t)i * ncols + colidx), z);
      cx_house(K, nrowscol, x, v, beta);
      cx_left_r_update(K, nrows, endcol, ncols, colidx, R, v, beta, w);
      memcpy(B + (size_t)L * K, beta, sizeof(double) * K);
      for (i = 0; i < L; i++) memset(Z(Y, (size_t)L * rowdim + i), 0, z);
      for (i = 0; i < nrowscol; i++) memcpy(Z(Y, (size_t)L * rowdim + i + L), Z(v, i), z);
    }
    vb_to_w(K, rowdim, szt, B, Y, W, p);
    if (k < nbt - 1) blocked_left_r(K, nrows, ncols, szt, k, R, Y, W);
    blocked_right_q(K, nrows, szt, k, Q, Y, W);
  }
  free(x);
  free(v);
  free(w);
  free(B);
  free(Y);
  free(W);
  free(p);
}

/* x = U^{-1} b by back substitution (U dim x dim with leading dimension ld) */
static void cx_backward(int K, int dim, int ld, const double *U, const double *b, double *x) {
  double a1re[MAXK], a1im[MAXK], a2re[MAXK], a2im[MAXK], a3re[MAXK], a3im[MAXK], den[MAXK];
  int i, j;
  for (i = dim - 1; i >= 0; i--) {
    double *xi = Z(x, i);
    memcpy(xi, Z(b, i), sizeof(double) * 2 * K);
    for (j = dim - 1; j > i; j--) {
      const double *u = Z(U, (size_t)i * ld + j), *xj = Z(x, j);
      omd_mul(K, RE(u), RE(xj), a1re);
      omd_mul(K, IM(u), IM(xj), a1im);
      omd_mul(K, IM(u), RE(xj), a2re);
      omd_mul(K, RE(u), IM(xj), a2im);
      omd_sub(K, RE(xi), a1re, RE(xi));
      omd_add(K, RE(xi), a1im, RE(xi));
      omd_sub(K, IM(xi), a2re, IM(xi));
      omd_sub(K, IM(xi), a2im, IM(xi));
    }
    {
      const double *u = Z(U, (size_t)i * ld + i);
      omd_mul(K, RE(u), RE(u), den);
      omd_mul(K, IM(u), IM(u), a1re);
      omd_add(K, den, a1re, den);
      omd_div(K, RE(u), den, a1re);
      omd_div(K, IM(u), den, a1im);
      omd_neg(K, a1im, a1im);
      omd_mul(K, RE(xi), a1re, a2re);
      omd_mul(K, IM(xi), a1im, a2im);
      omd_mul(K, IM(xi), a1re, a3re);
      omd_mul(K, RE(xi), a1im, a3im);
      memcpy(RE(xi), a2re, sizeof(double) * K);
      omd_sub(K, RE(xi), a2im, RE(xi));
      memcpy(IM(xi), a3re, sizeof(double) * K);
      omd_add(K, IM(xi), a3im, IM(xi));
    }
  }
}

static void upper_inverse(int K, int dim, int ld, const double *U, double *inv, int ldi) {
  const size_t z = sizeof(double) * 2 * K;
  double *rhs = calloc((size_t)dim, z), *col = malloc(z * dim);
  int i, j;
  for (j = 0; j < dim; j++) {
    omd_from_d(K, 1.0, RE(Z(rhs, j)));
    cx_backward(K, dim, ld, U, rhs, col);
    for (i = 0; i < dim; i++) memcpy(Z(inv, (size_t)i * ldi + j), Z(col, i), z);
    omd_zero(K, RE(Z(rhs, j)));
  }
  free(rhs);
  free(col);
}

/* sum_j T[i][j] * x[j] in the order of the tiled solver's products */
static void tile_matvec(int K, int szt, const double *T, int ld, const double *x, double *y) {
  double a1[MAXK], a2[MAXK], a3[MAXK], a4[MAXK];
  int i, j;
  for (i = 0; i < szt; i++) {
    double *yi = Z(y, i);
    omd_zero(K, RE(yi));
    omd_zero(K, IM(yi));
    for (j = 0; j < szt; j++) {
      const double *t = Z(T, (size_t)i * ld + j), *xj = Z(x, j);
      omd_mul(K, RE(t), RE(xj), a1);
      omd_mul(K, IM(t), IM(xj), a2);
      omd_mul(K, IM(t), RE(xj), a3);
      omd_mul(K, RE(t), IM(xj), a4);
      omd_sub(K, a1, a2, a1);
      omd_add(K, RE(yi), a1, RE(yi));
      omd_add(K, a3, a4, a3);
      omd_add(K, IM(yi), a3, IM(yi));
    }
  }
}

static void upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b,
                               double *x) {
  const size_t z = sizeof(double) * 2 * K;
  double *inv = malloc(z * (size_t)szt * szt), *wb = malloc(z * szt), *prod = malloc(z * szt);
  int idx = (nbt - 1) * szt, i, j, k, L;
  upper_inverse(K, szt, dim, Z(U, (size_t)idx * dim + idx), inv, szt);
  for (i = 0; i < szt; i++)
    for (j = 0; j < szt; j++) memcpy(Z(U, (size_t)(idx + i) * dim + idx + j), Z(inv, (size_t)i * szt + j), z);
  tile_matvec(K, szt, inv, szt, Z(b, idx), Z(x, idx));
  for (k = nbt - 1; k > 0; k--) {
    idx -= szt;
    upper_inverse(K, szt, dim, Z(U, (size_t)idx * dim + idx), inv, szt);
    for (i = 0; i < szt; i++)
      for (j = 0; j < szt; j++)
        memcpy(Z(U, (size_t)(idx + i) * dim + idx + j), Z(inv, (size_t)i * szt + j), z);
    for (L = 0; L < k; L++) {
      const int rowidx = L * szt;
      tile_matvec(K, szt, Z(U, (size_t)rowidx * dim + idx + szt), dim, Z(x, idx + szt), prod);
      for (i = 0; i < szt; i++) {
        double *bi = Z(b, rowidx + i);
        omd_sub(K, RE(bi), RE(Z(prod, i)), RE(bi));
        omd_sub(K, IM(bi), IM(Z(prod, i)), IM(bi));
      }
    }
    tile_matvec(K, szt, inv, szt, Z(b, idx), wb);
    memcpy(Z(x, idx), wb, z * szt);
  }
  free(inv);
  free(wb);
  free(prod);
}

/* ---------------------------------------------------------- public API -- */

void obq_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                         double *Q, double *R) {
  const size_t z = sizeof(double) * 2 * K;
  double *a = malloc(z * nrows * ncols), *q = malloc(z * nrows * nrows);
  double *r = malloc(z * nrows * ncols);
  cx_from_planar(K, (size_t)nrows * ncols, A, a);
  blocked_houseqr(K, nrows, ncols, szt, nbt, a, q, r);
  cx_to_planar(K, (size_t)nrows * nrows, q, Q);
  cx_to_planar(K, (size_t)nrows * ncols, r, R);
  free(a);
  free(q);
  free(r);
}

void obq_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b, double *x) {
  const size_t z = sizeof(double) * 2 * K;
  double *u = malloc(z * dim * dim), *bb = malloc(z * dim), *xx = calloc((size_t)dim, z);
  cx_from_planar(K, (size_t)dim * dim, U, u);
  cx_from_planar(K, dim, b, bb);
  upper_tiled_solver(K, dim, szt, nbt, u, bb, xx);
  cx_to_planar(K, (size_t)dim * dim, u, U);
  cx_to_planar(K, dim, bb, b);
  cx_to_planar(K, dim, xx, x);
  free(u);
  free(bb);
  free(xx);
}

void obq_qhb(int K, int n, const double *Q, const double *b, double *y) {
  const size_t z = sizeof(double) * 2 * K;
  double *q = malloc(z * n * n), *bb = malloc(z * n), *yy = malloc(z * n);
  int i, j;
  cx_from_planar(K, (size_t)n * n, Q, q);
  cx_from_planar(K, n, b, bb);
  for (i = 0; i < n; i++) {
    double *yi = Z(yy, i);
    omd_zero(K, RE(yi));
    omd_zero(K, IM(yi));
    for (j = 0; j < n; j++) {
      const double *qq = Z(q, (size_t)j * n + i), *bj = Z(bb, j);
      mac(K, RE(qq), RE(bj), RE(yi));
      mac(K, IM(qq), IM(bj), RE(yi));
      mac(K, RE(qq), IM(bj), IM(yi));
      msb(K, IM(qq), RE(bj), IM(yi));
    }
  }
  cx_to_planar(K, n, yy, y);
  free(q);
  free(bb);
  free(yy);
}

void obq_matvec(int K, int m, int n, const double *A, const double *x, double *y) {
  const size_t z = sizeof(double) * 2 * K;
  double *a = malloc(z * m * n), *xx = malloc(z * n), *yy = malloc(z * m);
  int i, j;
  cx_from_planar(K, (size_t)m * n, A, a);
  cx_from_planar(K, n, x, xx);
  for (i = 0; i < m; i++) {
    double *yi = Z(yy, i);
    omd_zero(K, RE(yi));
    omd_zero(K, IM(yi));
    for (j = 0; j < n; j++) {
      const double *aa = Z(a, (size_t)i * n + j), *xj = Z(xx, j);
      mac(K, RE(aa), RE(xj), RE(yi));
      msb(K, IM(aa), IM(xj), RE(yi));
      mac(K, IM(aa), RE(xj), IM(yi));
      mac(K, RE(aa), IM(xj), IM(yi));
    }
  }
  cx_to_planar(K, m, yy, y);
  free(a);
  free(xx);
  free(yy);
}

/* ---------------------------------------------------------------- real */

/* real number i of an array: K limbs at a + i*K */
#define RL(a, i) ((a) + (size_t)(i) * K)

static void re_house(int K, int n, const double *x, double *v, double *beta) {
  double sigma[MAXK], acc[MAXK], mu[MAXK], v0p2[MAXK];
  int i;
  omd_zero(K, sigma);
  omd_from_d(K, 1.0, RL(v, 0));
  for (i = 1; i < n; i++) {
    omd_sqr(K, RL(x, i), acc);
    omd_add(K, sigma, acc, sigma);
    memcpy(RL(v, i), RL(x, i), sizeof(double) * K);
  }
  if (all_zero(K, sigma)) {
    omd_zero(K, beta);
    return;
  }
  omd_sqr(K, RL(x, 0), acc);
  omd_add(K, acc, sigma, acc);
  omd_sqrt(K, acc, mu);
  if (RL(x, 0)[0] <= 0.0) {
    omd_sub(K, RL(x, 0), mu, RL(v, 0));
  } else {
    omd_add(K, RL(x, 0), mu, acc);
    omd_div(K, sigma, acc, RL(v, 0));
    omd_neg(K, RL(v, 0), RL(v, 0));
  }
  omd_sqr(K, RL(v, 0), v0p2);
  omd_add(K, sigma, v0p2, acc);
  omd_div(K, v0p2, acc, beta);
  omd_mul_d(K, beta, 2.0, beta);
  for (i = 1; i < n; i++) omd_div(K, RL(v, i), RL(v, 0), RL(v, i));
  omd_from_d(K, 1.0, RL(v, 0));
}

/* R (nrows x ldr) columns k..ncols-1 from row k: R -= v (beta v^T R) */
static void re_left_r_update(int K, int nrows, int ncols, int ldr, int k, double *R,
                             const double *v, const double *beta, double *w) {
  double acc[MAXK];
  int i, j;
  for (j = k; j < ncols; j++) {
    double *wj = RL(w, j - k);
    omd_zero(K, wj);
    for (i = k; i < nrows; i++) mac(K, RL(R, (size_t)i * ldr + j), RL(v, i - k), wj);
    omd_mul(K, beta, wj, acc);
    memcpy(wj, acc, sizeof(double) * K);
  }
  for (i = k; i < nrows; i++)
    for (j = k; j < ncols; j++) msb(K, RL(v, i - k), RL(w, j - k), RL(R, (size_t)i * ldr + j));
}

static void re_vb_to_w(int K, int rowdim, int szt, const double *B, const double *V, double *W,
                       double *p) {
  double z[MAXK];
  int i, j, k;
  for (i = 0; i < rowdim; i++) {
    omd_mul(K, B, RL(V, i), RL(W, i));
    omd_neg(K, RL(W, i), RL(W, i));
  }
  for (j = 1; j < szt; j++) {
    const double *vj = RL(V, (size_t)j * rowdim);
    for (k = 0; k < j; k++) {
      omd_zero(K, RL(p, k));
      for (i = 0; i < rowdim; i++) mac(K, RL(V, (size_t)k * rowdim + i), RL(vj, i), RL(p, k));
    }
    for (i = 0; i < rowdim; i++) {
      double *wji = RL(W, (size_t)j * rowdim + i);
      omd_zero(K, z);
      for (k = 0; k < j; k++) mac(K, RL(W, (size_t)k * rowdim + i), RL(p, k), z);
      omd_add(K, z, RL(vj, i), z);
      omd_mul(K, B + (size_t)j * K, z, wji);
      omd_neg(K, wji, wji);
    }
  }
}

/* C = C + (Y W^T) C on rows rowoff.., columns coloff.. */
static void re_blocked_left_r(int K, int nrows, int ncols, int szt, int idx, double *C,
                              const double *Y, const double *W) {
  const int rowoff = idx * szt, rowdim = nrows - rowoff, coloff = (idx + 1) * szt;
  const int coldim = ncols - coloff;
  double *ywt = malloc(sizeof(double) * K * (size_t)rowdim * rowdim);
  double *prd = malloc(sizeof(double) * K * (size_t)rowdim * (coldim > 0 ? coldim : 1));
  int i, j, k;
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = RL(ywt, (size_t)i * rowdim + j);
      omd_zero(K, e);
      for (k = 0; k < szt; k++)
        mac(K, RL(Y, (size_t)k * rowdim + i), RL(W, (size_t)k * rowdim + j), e);
    }
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < coldim; j++) {
      double *e = RL(prd, (size_t)i * coldim + j);
      omd_zero(K, e);
      for (k = 0; k < rowdim; k++)
        mac(K, RL(ywt, (size_t)i * rowdim + k), RL(C, (size_t)(rowoff + k) * ncols + coloff + j), e);
    }
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < coldim; j++) {
      double *c = RL(C, (size_t)(rowoff + i) * ncols + coloff + j);
      omd_add(K, c, RL(prd, (size_t)i * coldim + j), c);
    }
  free(ywt);
  free(prd);
}

/* Q = Q + Q (W Y^T) on columns coloff.. */
static void re_blocked_right_q(int K, int dim, int szt, int idx, double *Q, const double *Y,
                               const double *W) {
  const int coloff = idx * szt, rowdim = dim - coloff;
  double *wyt = malloc(sizeof(double) * K * (size_t)rowdim * rowdim);
  double *prd = malloc(sizeof(double) * K * (size_t)dim * rowdim);
  int i, j, k;
  for (i = 0; i < rowdim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = RL(wyt, (size_t)i * rowdim + j);
      omd_zero(K, e);
      for (k = 0; k < szt; k++)
        mac(K, RL(W, (size_t)k * rowdim + i), RL(Y, (size_t)k * rowdim + j), e);
    }
  for (i = 0; i < dim; i++)
    for (j = 0; j < rowdim; j++) {
      double *e = RL(prd, (size_t)i * rowdim + j);
      omd_zero(K, e);
      for (k = 0; k < rowdim; k++)
        mac(K, RL(Q, (size_t)i * dim + coloff + k), RL(wyt, (size_t)k * rowdim + j), e);
    }
  for (i = 0; i < dim; i++)
    for (j = 0; j < rowdim; j++) {
      double *q = RL(Q, (size_t)i * dim + coloff + j);
      omd_add(K, q, RL(prd, (size_t)i * rowdim + j), q);
    }
  free(wyt);
  free(prd);
}

static void re_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                               double *Q, double *R) {
  const size_t z = sizeof(double) * K;
  double *x = malloc(z * nrows), *v = malloc(z * nrows), *w = malloc(z * (nrows + ncols));
  double *B = malloc(z * szt);
  double *Y = malloc(z * (size_t)szt * nrows), *W = malloc(z * (size_t)szt * nrows);
  double *p = malloc(z * szt);
  double beta[MAXK];
  int i, j, k, L;
  for (i = 0; i < nrows; i++) {
    for (j = 0; j < nrows; j++) memset(RL(Q, (size_t)i * nrows + j), 0, z);
    omd_from_d(K, 1.0, RL(Q, (size_t)i * nrows + i));
    for (j = 0; j < ncols; j++) memcpy(RL(R, (size_t)i * ncols + j), RL(A, (size_t)i * ncols + j), z);
  }
  for (k = 0; k < nbt; k++) {
    const int rowdim = nrows - k * szt;
    for (L = 0; L < szt; L++) {
      const int colidx = k * szt + L, endcol = (k + 1) * szt, nrowscol = nrows - colidx;
      for (i = colidx; i < nrows; i++) memcpy(RL(x, i - colidx), RL(R, (size_t)i * ncols + colidx), z);
      re_house(K, nrowscol, x, v, beta);
      re_left_r_update(K, nrows, endcol, ncols, colidx, R, v, beta, w);
      memcpy(B + (size_t)L * K, beta, z);
      for (i = 0; i < L; i++) memset(RL(Y, (size_t)L * rowdim + i), 0, z);
      for (i = 0; i < nrowscol; i++) memcpy(RL(Y, (size_t)L * rowdim + i + L), RL(v, i), z);
    }
    re_vb_to_w(K, rowdim, szt, B, Y, W, p);
    if (k < nbt - 1) re_blocked_left_r(K, nrows, ncols, szt, k, R, Y, W);
    re_blocked_right_q(K, nrows, szt, k, Q, Y, W);
  }
  free(x);
  free(v);
  free(w);
  free(B);
  free(Y);
  free(W);
  free(p);
}

static void re_from_planar(int K, size_t n, const double *p, double *a) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < K; q++) a[j * K + q] = p[q * n + j];
}

static void re_to_planar(int K, size_t n, const double *a, double *p) {
  size_t j;
  int q;
  for (j = 0; j < n; j++)
    for (q = 0; q < K; q++) p[q * n + j] = a[j * K + q];
}

void obq_real_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                              double *Q, double *R) {
  const size_t na = (size_t)nrows * ncols, nq = (size_t)nrows * nrows;
  double *a = malloc(sizeof(double) * K * na), *q = malloc(sizeof(double) * K * nq);
  double *r = malloc(sizeof(double) * K * na);
  re_from_planar(K, na, A, a);
  re_blocked_houseqr(K, nrows, ncols, szt, nbt, a, q, r);
  re_to_planar(K, nq, q, Q);
  re_to_planar(K, na, r, R);
  free(a);
  free(q);
  free(r);
}

/* ------------------------------------------- shared with oracle/newton.c */

void obq_cx_house(int K, int n, const double *x, double *v, double *beta) {
  cx_house(K, n, x, v, beta);
}

void obq_cx_left_r_update(int K, int nrows, int ncols, int ldr, int k, double *R,
                          const double *v, const double *beta, double *w) {
  cx_left_r_update(K, nrows, ncols, ldr, k, R, v, beta, w);
}

void obq_cx_backward(int K, int dim, int ld, const double *U, const double *b, double *x) {
  cx_backward(K, dim, ld, U, b, x);
}
