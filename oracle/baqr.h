/* CPU oracle: blocked Householder QR and tiled back substitution of complex
 * multiple-double matrices (test infrastructure only; see baqr.c).
 * Arrays are complex limb-planar: [2K][...] (K real planes, K imaginary
 * planes), matrices row-major. */
#ifndef PHC_ORACLE_BAQR_H
#define PHC_ORACLE_BAQR_H
#ifdef __cplusplus
extern "C" {
#endif

/* A (nrows x ncols) = Q R with Q (nrows x nrows) explicit; nbt tiles of szt columns */
void obq_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                         double *Q, double *R);
/* U x = b for upper triangular U (dim = szt*nbt); U's diagonal tiles are
 * replaced by their inverses and b by the updated right-hand side, as the
 * reference does */
void obq_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b, double *x);
/* y = Q^H b */
void obq_qhb(int K, int n, const double *Q, const double *b, double *y);
/* the real blocked QR: A, Q, R real limb-planar [K][rows][cols] */
void obq_real_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, const double *A,
                              double *Q, double *R);
/* y = A x (m x n) */
void obq_matvec(int K, int m, int n, const double *A, const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
