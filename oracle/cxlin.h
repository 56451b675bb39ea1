/* CPU oracle (test infrastructure only): the complex Householder vector,
 * panel update and back substitution of baqr.c, shared with the complex
 * Newton of newton.c.  Complex numbers interleaved: 2K doubles each, the K
 * real limbs then the K imaginary limbs; matrices row-major. */
#ifndef PHC_ORACLE_CXLIN_H
#define PHC_ORACLE_CXLIN_H

/* CPU_cmplx2_factors_house (Matrices/dbl2_factorizations.cpp:441-533) */
void obq_cx_house(int K, int n, const double *x, double *v, double *beta);
/* CPU_cmplx2_factors_leftRupdate (:567-627): R (nrows x ldr), columns
 * k..ncols-1 from row k; w scratch of ncols - k numbers */
void obq_cx_left_r_update(int K, int nrows, int ncols, int ldr, int k, double *R,
                          const double *v, const double *beta, double *w);
/* CPU_cmplx2_factors_backward (:142-200): x = U^{-1} b, U dim x dim with
 * leading dimension ld */
void obq_cx_backward(int K, int dim, int ld, const double *U, const double *b, double *x);

#endif
