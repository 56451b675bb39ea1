/* CPU oracle: truncated power series products and polynomial evaluation and
 * differentiation at power series (test infrastructure only).
 *
 * Restates, for K = 1, 2, 4, 8 limbs, real and complex:
 *   CPU_dbl2_product / CPU_cmplx2_product
 *       Convolutions/dbl2_convolutions_host.cpp:7-88
 *   CPU_dbl2_speel (forward, backward, cross products)
 *       Convolutions/dbl2_monomials_host.cpp:28-87
 *   CPU_dbl2_poly_speel / CPU_dbl2_poly_evaldiff
 *       Polynomials/dbl2_polynomials_host.cpp:13-318
 * in the same order of operations (bit-exact against the compiled reference,
 * tests/test_oracle_series.py).
 *
 * Layout: a series of degree deg has D = deg+1 coefficients; an array of m
 * series with P = K (real) or 2K (complex) limb planes is stored [P][m][D].
 */
#include <stdlib.h>
#include <string.h>
#include "md.h"
#include "series.h"

/* number j of a planar array with P planes of stride s -> packed buffer */
static void ld(const double *a, int P, size_t s, size_t j, double *v) {
  int p;
  for (p = 0; p < P; p++) v[p] = a[p * s + j];
}
static void st(double *a, int P, size_t s, size_t j, const double *v) {
  int p;
  for (p = 0; p < P; p++) a[p * s + j] = v[p];
}

static void num_mul(int K, int cplx, const double *a, const double *b, double *c) {
  if (cplx) omd_cmul(K, a, b, c); else omd_mul(K, a, b, c);
}
static void num_mac(int K, int cplx, const double *a, const double *b, double *c) {
  double t[8];
  if (cplx) omd_cmac(K, a, b, c);
  else { omd_mul(K, a, b, t); omd_add(K, c, t, c); }
}
static void num_inc(int K, int cplx, double *c, const double *a) {
  omd_add(K, c, a, c);
  if (cplx) omd_add(K, c + K, a + K, c + K);
}

/* x, y, z: series with plane stride sx, sy, sz (in doubles) */
static void product_s(int K, int cplx, int deg, const double *x, size_t sx,
                      const double *y, size_t sy, double *z, size_t sz) {
  const int P = cplx ? 2 * K : K;
  double a[16], b[16], acc[16];
  int k, i;
  for (k = 0; k <= deg; k++) {
    ld(x, P, sx, 0, a);
    ld(y, P, sy, k, b);
    num_mul(K, cplx, a, b, acc);
    for (i = 1; i <= k; i++) {
      ld(x, P, sx, i, a);
      ld(y, P, sy, k - i, b);
      num_mac(K, cplx, a, b, acc);
    }
    st(z, P, sz, k, acc);
  }
}

void osr_product(int K, int cplx, int deg, const double *x, const double *y, double *z) {
  const size_t D = deg + 1;
  double *w = (double *)malloc(sizeof(double) * 16 * D);
  product_s(K, cplx, deg, x, D, y, D, w, D);
  memcpy(z, w, sizeof(double) * (cplx ? 2 * K : K) * D);
  free(w);
}

/* series number r of an array of m series: base pointer with plane stride m*D */
#define SER(a, r) ((a) + (size_t)(r) * D)

/* output[dst] += src (coefficientwise) */
static void ser_inc(int K, int cplx, int deg, double *dst, size_t sd, const double *src,
                    size_t ss) {
  const int P = cplx ? 2 * K : K;
  double a[16], b[16];
  int j;
  for (j = 0; j <= deg; j++) {
    ld(dst, P, sd, j, a);
    ld(src, P, ss, j, b);
    num_inc(K, cplx, a, b);
    st(dst, P, sd, j, a);
  }
}

/* output[dst] += e * src (coefficientwise; the multiplication by e as
 * CPU_dbl2_evaluate_monomials scales a derivative, only for e > 1) */
static void ser_inc_e(int K, int cplx, int deg, double *dst, size_t sd, const double *src,
                      size_t ss, int e) {
  const int P = cplx ? 2 * K : K;
  double a[16], b[16];
  int j;
  if (e <= 1) {
    ser_inc(K, cplx, deg, dst, sd, src, ss);
    return;
  }
  for (j = 0; j <= deg; j++) {
    ld(dst, P, sd, j, a);
    ld(src, P, ss, j, b);
    omd_mul_d(K, b, (double)e, b);
    if (cplx) omd_mul_d(K, b + K, (double)e, b + K);
    num_inc(K, cplx, a, b);
    st(dst, P, sd, j, a);
  }
}

void osr_poly_evaldiff(int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                       const int *idxptr, const int *idx, const double *cst,
                       const double *cff, const double *input, double *output) {
  osr_poly_evaldiff_exp(K, cplx, dim, nbr, deg, nvr, idxptr, idx, NULL, cst, cff, input, output);
}

/* With exponents (exps parallel to idx, NULL for all ones): the coefficient
 * of every monomial times its common factor prod_j x_j^(e_j-1), the factors
 * multiplied in one at a time in the order of idx (the first loop of
 * CPU_dbl2_evaluate_monomials, Newton/dbl2_systems_host.cpp:18-45), then the
 * products of CPU_dbl2_poly_speel on its distinct variables, each
 * derivative contribution multiplied by its exponent (the last loop,
 * :103-110) before it is added. */
void osr_poly_evaldiff_exp(int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                           const int *idxptr, const int *idx, const int *exps, const double *cst,
                           const double *cff, const double *input, double *output) {
  const int P = cplx ? 2 * K : K;
  const size_t D = deg + 1;
  const size_t sx = (size_t)dim * D, so = (size_t)(dim + 1) * D;
  size_t sc = (size_t)nbr * D;
  const size_t sw = (size_t)dim * D; /* work arrays of dim series */
  double *fw = (double *)calloc((size_t)P * sw, sizeof(double));
  double *bw = (double *)calloc((size_t)P * sw, sizeof(double));
  double *cr = (double *)calloc((size_t)P * sw, sizeof(double));
  double *cf = (double *)calloc((size_t)P * D, sizeof(double));
  double *acc = (double *)calloc((size_t)P * D, sizeof(double));
  const size_t sc0 = sc;
  int i, j, p, k, e;

  for (p = 0; p < P; p++)
    for (j = 0; j <= deg; j++) {
      output[p * so + (size_t)dim * D + j] = cst[p * D + j];
      for (i = 0; i < dim; i++) output[p * so + (size_t)i * D + j] = 0.0;
    }
  for (i = 0; i < nbr; i++) {
    const int nv = nvr[i];
    const int *ix = idx + idxptr[i];
    const int *ex = exps ? exps + idxptr[i] : NULL;
    const double *c = SER(cff, i);
    int common = 0;
    sc = sc0;
    for (j = 0; ex && j < nv; j++) common = common || ex[j] > 1;
    if (common) {
      for (p = 0; p < P; p++) memcpy(cf + p * D, c + p * sc0, sizeof(double) * D);
      for (j = 0; j < nv; j++)
        for (e = 1; e < ex[j]; e++) {
          product_s(K, cplx, deg, SER(input, ix[j]), sx, cf, D, acc, D);
          memcpy(cf, acc, sizeof(double) * P * D);
        }
      c = cf;
      sc = D;
    }
#define EXP(j) (ex ? ex[j] : 1)
    if (nv == 1) {
      product_s(K, cplx, deg, SER(input, ix[0]), sx, c, sc, fw, sw);
      ser_inc(K, cplx, deg, SER(output, dim), so, fw, sw);
      ser_inc_e(K, cplx, deg, SER(output, ix[0]), so, c, sc, EXP(0));
    } else if (nv == 2) {
      product_s(K, cplx, deg, c, sc, SER(input, ix[0]), sx, fw, sw);
      ser_inc_e(K, cplx, deg, SER(output, ix[1]), so, fw, sw, EXP(1));
      product_s(K, cplx, deg, c, sc, SER(input, ix[1]), sx, bw, sw);
      ser_inc_e(K, cplx, deg, SER(output, ix[0]), so, bw, sw, EXP(0));
      product_s(K, cplx, deg, fw, sw, SER(input, ix[1]), sx, SER(fw, 1), sw);
      ser_inc(K, cplx, deg, SER(output, dim), so, SER(fw, 1), sw);
    } else if (nv > 2) {
      /* CPU_dbl2_speel: forward, backward, cross products */
      product_s(K, cplx, deg, c, sc, SER(input, ix[0]), sx, fw, sw);
      for (k = 1; k < nv; k++)
        product_s(K, cplx, deg, SER(fw, k - 1), sw, SER(input, ix[k]), sx, SER(fw, k), sw);
      product_s(K, cplx, deg, SER(input, ix[nv - 1]), sx, SER(input, ix[nv - 2]), sx, bw, sw);
      for (k = 1; k < nv - 2; k++)
        product_s(K, cplx, deg, SER(bw, k - 1), sw, SER(input, ix[nv - 2 - k]), sx,
                  SER(bw, k), sw);
      product_s(K, cplx, deg, SER(bw, nv - 3), sw, c, sc, cr, sw);
      for (p = 0; p < P; p++)
        memcpy(bw + p * sw + (size_t)(nv - 3) * D, cr + p * sw, sizeof(double) * D);
      if (nv == 3) {
        product_s(K, cplx, deg, fw, sw, SER(input, ix[2]), sx, cr, sw);
      } else {
        for (k = 0; k < nv - 3; k++)
          product_s(K, cplx, deg, SER(fw, k), sw, SER(bw, nv - 4 - k), sw, SER(cr, k), sw);
        product_s(K, cplx, deg, SER(fw, nv - 3), sw, SER(input, ix[nv - 1]), sx,
                  SER(cr, nv - 3), sw);
      }
      ser_inc(K, cplx, deg, SER(output, dim), so, SER(fw, nv - 1), sw);
      ser_inc_e(K, cplx, deg, SER(output, ix[nv - 1]), so, SER(fw, nv - 2), sw, EXP(nv - 1));
      ser_inc_e(K, cplx, deg, SER(output, ix[0]), so, SER(bw, nv - 3), sw, EXP(0));
      for (k = 1; k < nv - 1; k++)
        ser_inc_e(K, cplx, deg, SER(output, ix[k]), so, SER(cr, k - 1), sw, EXP(k));
    }
#undef EXP
  }
  free(fw);
  free(bw);
  free(cr);
  free(cf);
  free(acc);
}
