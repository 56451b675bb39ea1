/* CPU oracle: linearized power-series Newton on real multiple-double series
 * (test infrastructure only; see newton.c for the reference files restated).
 *
 * Array conventions (limb-planar, K limbs, plane p at offset p*stride):
 *   matrix   [K][rows][cols]            row-major
 *   mat      [K][degp1][dim][dim]       matrix series, degree-major
 *   rhs/sol  [K][degp1][dim]            vector series, degree-major
 *   input    [K][dim][deg+1]            the series x_i(t)
 *   cst      [K][dim][deg+1]            constant series of each polynomial
 *   cff      [K][nmono][deg+1]          coefficient series of each monomial
 * Monomial m of the system has nvr[m] distinct variables idx[idxptr[m] ..];
 * polynomial i owns monomials monoptr[i] .. monoptr[i+1]-1.
 */
#ifndef PHC_ORACLE_NEWTON_H
#define PHC_ORACLE_NEWTON_H
#ifdef __cplusplus
extern "C" {
#endif

/* skip threshold of the tail solve for K limbs (1e-15, 1e-28, 1e-56, 1e-120) */
double onw_threshold(int K);

/* Householder QR with explicit Q: A (rows x cols) -> Q (rows x rows), R */
void onw_houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R);

/* x = R^{-1} Q^T b for square Q, R of dimension dim */
void onw_qrbs(int K, int dim, const double *Q, const double *R, const double *b, double *x);

/* lower block-triangular Toeplitz solve; flags in/out as the reference keeps
 * them across Newton steps; rhs is overwritten by the tail updates */
void onw_qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                    double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                    int *bsidx, int *newtail);

/* r_i = b_i - sum_{j<=i-tailidx} A_j x_{i-j} for i = tailidx..degp1-1;
 * resmax = the K limbs of the residual entry of largest |leading limb| */
void onw_linear_residue(int K, int dim, int degp1, int tailidx, const double *mat,
                        const double *rhs, const double *sol, double *resvec, double *resmax);

/* staggered row Newton: nbsteps steps at working degrees 0, 1, 2, 4, 7, ...
 * (exps parallel to idx, NULL for all ones: the exponents of the variables)
 * capped at deg, stopping when the tail index reaches deg.  input is
 * updated in place.  Per step s (when the pointers are non-null):
 * wdeg[s], tail[s] (after the step), up[s], bs[s], and history
 * [nbsteps][K][dim][deg+1] = input after step s.  Q and R ([K][dim][dim])
 * receive the last factorization.  Returns the number of steps run. */
int onw_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                   const int *idxptr, const int *idx, const int *exps, const double *cst,
                   const double *cff, double *input, int *wdeg, int *tail, int *up, int *bs,
                   double *history, double *Q, double *R);

/* staggered column Newton (the loop of test_dbl2_column_newton with
 * dbl2_column_newton_qrstep): nbrcol columns of dim monomials, monomial
 * m = c*dim + j (equation j of column c) has nvr[m] variables (0: none)
 * idx[idxptr[m] ..].  nbrcol = 1 is a monomial system x^A_i = mb_i with unit
 * coefficients and exponents exps parallel to idx (NULL: all ones); for
 * nbrcol > 1 the coefficients are cff [K][nbrcol*dim][deg+1], exps are
 * ignored and the last equation has the constant -1.  The right-hand side
 * of the linearized system is -(f - mb), mb [K][dim][deg+1].  Other
 * arguments as onw_row_newton. */
int onw_column_newton(int K, int dim, int nbrcol, int deg, int nbsteps, const int *nvr,
                      const int *idxptr, const int *idx, const int *exps, const double *cff,
                      const double *mb, double *input, int *wdeg, int *tail, int *up, int *bs,
                      double *history, double *Q, double *R);

/* complex (K = 2, 4, 8), arrays complex limb-planar [2K][...] (K real
 * planes, K imaginary planes); resmax [K] of |re| + |im| per limb */
void onw_cx_houseqr(int K, int nrows, int ncols, const double *A, double *Q, double *R);
void onw_cx_qrbs_solve(int K, int dim, int degp1, int tailidx, const double *mat, double *rhs,
                       double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                       int *bsidx, int *newtail);
void onw_cx_linear_residue(int K, int dim, int degp1, int tailidx, const double *mat,
                           const double *rhs, const double *sol, double *resvec, double *resmax);
/* the complex row Newton of distinct variables (no exponents) */
int onw_cx_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                      const int *idxptr, const int *idx, const double *cst, const double *cff,
                      double *input, int *wdeg_out, int *tail_out, int *up_out, int *bs_out,
                      double *history, double *Q, double *R);

#ifdef __cplusplus
}
#endif
#endif
