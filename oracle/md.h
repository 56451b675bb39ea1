/* CPU oracle: multiple-double arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (phcpack_amd/) may link or
 * call this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, as a checker.
 *
 * A number with K limbs (K = 1, 2, 4, 8) is an array double[K], limb 0 the
 * most significant.  K = 2 restates the QD library double-double operations
 * of Norms/double_double_functions.cpp:9-277; K = 4 and K = 8 restate the
 * CAMPARY operations of Norms/quad_double_functions.cpp:12-380 and
 * Norms/octo_double_functions.cpp:13-1530.  two_prod uses Dekker's split,
 * as Norms/double_double_functions.cpp:148-182 does.
 *
 * Parity pinned by: tests/test_oracle_md.py (bit-exact against the compiled
 * reference functions in oracle/_ref when present, against committed golden
 * vectors otherwise; sqrt(2)^2 = 2 and EFT identities).
 */
#ifndef PHC_ORACLE_MD_H
#define PHC_ORACLE_MD_H

#ifdef __cplusplus
extern "C" {
#endif

double omd_two_sum(double a, double b, double *err);
double omd_quick_two_sum(double a, double b, double *err);
double omd_two_prod(double a, double b, double *err);

/* c = a op b, all with K limbs; c may alias a or b */
void omd_add(int K, const double *a, const double *b, double *c);
void omd_sub(int K, const double *a, const double *b, double *c);
void omd_mul(int K, const double *a, const double *b, double *c);
void omd_div(int K, const double *a, const double *b, double *c);
void omd_sqrt(int K, const double *a, double *c);
void omd_mul_d(int K, const double *a, double b, double *c);
/* c = a*a as the reference squares (ddf_sqr differs from ddf_mul(a,a)) */
void omd_sqr(int K, const double *a, double *c);
void omd_add_d(int K, const double *a, double b, double *c);
void omd_neg(int K, const double *a, double *c);
void omd_copy(int K, const double *a, double *c);
void omd_zero(int K, double *c);
void omd_from_d(int K, double a, double *c);

/* complex numbers: re limbs then im limbs, 2K doubles */
void omd_cadd(int K, const double *a, const double *b, double *c);
void omd_csub(int K, const double *a, const double *b, double *c);
void omd_cmul(int K, const double *a, const double *b, double *c);
/* z += x*y in the order of CPU_cmplx2_product's inner loop
 * (Convolutions/dbl2_convolutions_host.cpp:75-85) */
void omd_cmac(int K, const double *x, const double *y, double *z);
void omd_cdiv(int K, const double *a, const double *b, double *c);
/* |a|^2 and |a| */
void omd_cnorm2(int K, const double *a, double *c);

/* batched elementwise entry points for tests:
 * op 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(a); arrays are n numbers of K limbs,
 * limb-planar: limb i of number j at [i*n + j] */
void omd_batch(int op, int K, int n, const double *a, const double *b, double *c);

#ifdef __cplusplus
}
#endif
#endif
