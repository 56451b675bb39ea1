/* CPU oracle: power series products and polynomial evaluation/differentiation
 * (test infrastructure only; see series.c for the reference files restated). */
#ifndef PHC_ORACLE_SERIES_H
#define PHC_ORACLE_SERIES_H
#ifdef __cplusplus
extern "C" {
#endif

/* z = x*y truncated at degree deg; [P][deg+1] planar, P = K or 2K */
void osr_product(int K, int cplx, int deg, const double *x, const double *y, double *z);

/* value and gradient of cst + sum_i cff_i * prod_{j in idx_i} x_j at the
 * series input; output [P][dim+1][deg+1] with output[dim] the value */
void osr_poly_evaldiff(int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                       const int *idxptr, const int *idx, const double *cst,
                       const double *cff, const double *input, double *output);

/* the same with exponents exps[] parallel to idx[] (NULL: all ones): the
 * common-factor trick of CPU_dbl2_evaluate_monomials per monomial */
void osr_poly_evaldiff_exp(int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                           const int *idxptr, const int *idx, const int *exps, const double *cst,
                           const double *cff, const double *input, double *output);

#ifdef __cplusplus
}
#endif
#endif
