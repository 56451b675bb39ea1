// Flat extern "C" wrappers around the reference's own host code of the
// linearized power-series Newton method (QR, Toeplitz solve, residual,
// staggered row Newton) and of its test-system generators, compiled with the
// reference sources (read in place) into oracle/_ref/libphcref.so.
//
// TEST INFRASTRUCTURE ONLY: pins the oracle restatement (oracle/newton.c)
// and writes golden vectors.  Never linked into the product.
//
// Arrays are limb-planar as in oracle/newton.h.
#include <cstdlib>
#include <cstring>
#include <vector>
#include "double_double_functions.h"
#include "quad_double_functions.h"
#include "octo_double_functions.h"
#include "random_numbers.h"
#include "random2_vectors.h"
#include "dbl_factorizations.h"
#include "dbl2_factorizations.h"
#include "dbl4_factorizations.h"
#include "dbl8_factorizations.h"
#include "dbl_bals_host.h"
#include "dbl2_bals_host.h"
#include "dbl4_bals_host.h"
#include "dbl8_bals_host.h"
#include "dbl_convolutions_host.h"
#include "dbl2_convolutions_host.h"
#include "dbl4_convolutions_host.h"
#include "dbl8_convolutions_host.h"
#include "dbl_monomials_host.h"
#include "dbl2_monomials_host.h"
#include "dbl4_monomials_host.h"
#include "dbl8_monomials_host.h"
#include "dbl_polynomials_host.h"
#include "dbl2_polynomials_host.h"
#include "dbl4_polynomials_host.h"
#include "dbl8_polynomials_host.h"
#include "cyclic_indices.h"
#include "dbl2_indexed_coefficients.h"
#include "dbl2_monomial_systems.h"
#include "dbl_systems_host.h"
#include "dbl2_systems_host.h"
#include "dbl4_systems_host.h"
#include "dbl8_systems_host.h"
#include "unimodular_matrices.h"
#include "cyclic_columns.h"

namespace {

// row pointers of a [P][nmat][rows][cols] array: m[p][k][i] -> row i of matrix k
struct Mats {
  std::vector<std::vector<double *>> rows;
  std::vector<std::vector<double **>> mats;
  Mats(double *base, int P, int nmat, int nr, int nc) : rows(P * nmat), mats(P) {
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < nmat; ++k) {
        auto &r = rows[p * nmat + k];
        for (int i = 0; i < nr; ++i)
          r.push_back(base + (((size_t)p * nmat + k) * nr + i) * nc);
      }
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < nmat; ++k) mats[p].push_back(rows[p * nmat + k].data());
  }
  double ***ser(int p) { return mats[p].data(); }   // series of matrices, plane p
  double **mat(int p) { return mats[p][0]; }        // nmat == 1, plane p
};

// limb expansions of arrays indexed by plane
#define L1(a) a[0]
#define L2(a) a[0], a[1]
#define L4(a) a[0], a[1], a[2], a[3]
#define L8(a) a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]

template <int K> struct Ref;

#define DEFREF(KK, PFX, LX)                                                                   \
  template <> struct Ref<KK> {                                                                \
    static void houseqr(int nr, int nc, double **A[], double **Q[], double **R[]) {           \
      CPU_##PFX##_factors_houseqr(nr, nc, LX(A), LX(Q), LX(R));                               \
    }                                                                                         \
    static void solve(int dim, int degp1, int tail, double ***M[], double **B[],             \
                      double **X[], double **Q[], double **R[], double *W[], bool *zq,        \
                      bool *nq, int *up, int *bs, int *nt) {                                  \
      CPU_##PFX##_qrbs_solve(dim, degp1, tail, LX(M), LX(B), LX(X), LX(Q), LX(R), LX(W), zq,  \
                             nq, up, bs, nt, 0);                                              \
    }                                                                                         \
    static void residue(int dim, int degp1, int tail, double ***M[], double **B[],           \
                        double **X[], double **Rv[], double *rm[]) {                          \
      CPU_##PFX##_linear_residue(dim, degp1, tail, LX(M), LX(B), LX(X), LX(Rv), LX(rm), 0);   \
    }                                                                                         \
    static void evaldiff(int dim, int nbr, int deg, int *nvr, int **idx, double *c[],        \
                         double **f[], double **in[], double **out[]) {                       \
      double el = 0.0;                                                                        \
      CPU_##PFX##_poly_evaldiff(dim, nbr, deg, nvr, idx, LX(c), LX(f), LX(in), LX(out), &el,  \
                                0);                                                           \
    }                                                                                         \
    static void product(int deg, double *x[], double *y[], double *z[]) {                     \
      CPU_##PFX##_product(deg, LX(x), LX(y), LX(z));                                          \
    }                                                                                         \
    static void speel(int nvr, int deg, int *idx, double *c[], double **in[], double **f[],   \
                      double **b[], double **x[]) {                                           \
      CPU_##PFX##_speel(nvr, deg, idx, LX(c), LX(in), LX(f), LX(b), LX(x));                   \
    }                                                                                         \
    static void monomials(int dim, int deg, int *nvr, int **idx, int **exp, int *nbrfac,      \
                          int **expfac, double **c[], double *acc[], double **in[],           \
                          double ***out[]) {                                                  \
      CPU_##PFX##_evaluate_monomials(dim, deg, nvr, idx, exp, nbrfac, expfac, LX(c), LX(acc), \
                                     LX(in), LX(out), 0);                                     \
    }                                                                                         \
    static void linearize(int dim, int degp1, int *nvr, int **idx, double **mb[],            \
                          double ***out[], double **f[], double **rhs[], double ***jac[]) {   \
      PFX##_linearize_evaldiff_output(dim, degp1, nvr, idx, LX(mb), 1.0, LX(out), LX(f),      \
                                      LX(rhs), LX(jac), 0);                                   \
    }                                                                                         \
    static void columns(int dim, int deg, int nbrcol, int **nvr, int ***idx, double ***c[],   \
                        double **acc[], double **in[], double **f[], double ***jac[]) {       \
      CPU_##PFX##_evaluate_columns(dim, deg, nbrcol, nvr, idx, LX(c), LX(acc), LX(in), LX(f), \
                                   LX(jac), 0);                                               \
    }                                                                                         \
    static void define_rhs(int dim, int degp1, double **mb[], double **f[], double **rhs[]) { \
      PFX##_define_rhs(dim, degp1, LX(mb), LX(f), LX(rhs), 0);                                \
    }                                                                                         \
  };
DEFREF(1, dbl, L1)
DEFREF(2, dbl2, L2)
DEFREF(4, dbl4, L4)
DEFREF(8, dbl8, L8)
#undef DEFREF

// the complex functions (K = 2, 4, 8 limbs, 2K planes: real limbs, then
// imaginary limbs; no octo-double complex polynomial evaluation)
#define L16(a) L8(a), a[8], a[9], a[10], a[11], a[12], a[13], a[14], a[15]
template <int K> struct CRef;
#define DEFCREF(KK, PFX, LX, LK)                                                              \
  template <> struct CRef<KK> {                                                               \
    static void houseqr(int nr, int nc, double **A[], double **Q[], double **R[]) {           \
      CPU_##PFX##_factors_houseqr(nr, nc, LX(A), LX(Q), LX(R));                               \
    }                                                                                         \
    static void solve(int dim, int degp1, int tail, double ***M[], double **B[],             \
                      double **X[], double **Q[], double **R[], double *W[], bool *zq,        \
                      bool *nq, int *up, int *bs, int *nt) {                                  \
      CPU_##PFX##_qrbs_solve(dim, degp1, tail, LX(M), LX(B), LX(X), LX(Q), LX(R), LX(W), zq,  \
                             nq, up, bs, nt, 0);                                              \
    }                                                                                         \
    static void residue(int dim, int degp1, int tail, double ***M[], double **B[],           \
                        double **X[], double **Rv[], double *rm[]) {                          \
      CPU_##PFX##_linear_residue(dim, degp1, tail, LX(M), LX(B), LX(X), LX(Rv), LK(rm), 0);   \
    }                                                                                         \
  };
DEFCREF(2, cmplx2, L4, L2)
DEFCREF(4, cmplx4, L8, L4)
DEFCREF(8, cmplx8, L16, L8)
#undef DEFCREF

template <int K> struct CEval;
template <> struct CEval<2> {
  static void evaldiff(int dim, int nbr, int deg, int *nvr, int **idx, double *c[],
                       double **f[], double **in[], double **out[]) {
    double el = 0.0;
    CPU_cmplx2_poly_evaldiff(dim, nbr, deg, nvr, idx, L4(c), L4(f), L4(in), L4(out), &el, 0);
  }
};
template <> struct CEval<4> {
  static void evaldiff(int dim, int nbr, int deg, int *nvr, int **idx, double *c[],
                       double **f[], double **in[], double **out[]) {
    double el = 0.0;
    CPU_cmplx4_poly_evaldiff(dim, nbr, deg, nvr, idx, L8(c), L8(f), L8(in), L8(out), &el, 0);
  }
};

// x += dx with the reference's in-place addition of each precision
// (dbl2_update_series, Newton/dbl2_newton_testers.cpp:116-149, and its twins)
template <int K> void inc(double *x[], const double *dx[]);
template <> void inc<1>(double *x[], const double *dx[]) { *x[0] = *x[0] + *dx[0]; }
template <> void inc<2>(double *x[], const double *dx[]) { ddf_inc(x[0], x[1], *dx[0], *dx[1]); }
template <> void inc<4>(double *x[], const double *dx[]) {
  qdf_inc(x[0], x[1], x[2], x[3], *dx[0], *dx[1], *dx[2], *dx[3]);
}
template <> void inc<8>(double *x[], const double *dx[]) {
  odf_inc(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], *dx[0], *dx[1], *dx[2], *dx[3],
          *dx[4], *dx[5], *dx[6], *dx[7]);
}

// z = factor * z with the multiplication by a double of each precision, as
// CPU_*_evaluate_monomials scales a derivative by its exponent
// (Newton/dbl2_systems_host.cpp:103-110 and its twins)
template <int K> void mul_d(double factor, double *z[]);
template <> void mul_d<1>(double factor, double *z[]) { *z[0] = factor * *z[0]; }
template <> void mul_d<2>(double factor, double *z[]) {
  double hi, lo;
  ddf_mul_d_dd(factor, *z[0], *z[1], &hi, &lo);
  *z[0] = hi;
  *z[1] = lo;
}
template <> void mul_d<4>(double factor, double *z[]) {
  double c[4];
  qdf_mul_qd_d(*z[0], *z[1], *z[2], *z[3], factor, &c[0], &c[1], &c[2], &c[3]);
  for (int p = 0; p < 4; ++p) *z[p] = c[p];
}
template <> void mul_d<8>(double factor, double *z[]) {
  double c[8];
  odf_mul_od_d(*z[0], *z[1], *z[2], *z[3], *z[4], *z[5], *z[6], *z[7], factor, &c[0], &c[1],
               &c[2], &c[3], &c[4], &c[5], &c[6], &c[7]);
  for (int p = 0; p < 8; ++p) *z[p] = c[p];
}

// Value and gradient of polynomial cst + sum_m cff_m x^e_m at the series in,
// composed from the reference's own functions: the coefficient of every
// monomial times its common factor prod_j x_j^(e_j-1) by CPU_*_product (the
// first loop of CPU_dbl2_evaluate_monomials, Newton/dbl2_systems_host.cpp:
// 18-45), then the products of CPU_*_poly_speel for its distinct variables
// (Polynomials/dbl2_polynomials_host.cpp:13-108; CPU_*_speel for three
// or more), each derivative contribution multiplied by its exponent (the
// last loop of CPU_dbl2_evaluate_monomials, :103-110) before it is added.
// out[p][0..dim] series (rows of dim+1), c/cff/in as row pointers per plane.
template <int K>
void poly_evaldiff_exp(int dim, int nbr, int deg, const int *nvr, int **idx, int **exps,
                       double *c[], double **cff[], double **in[], double **out[]) {
  const int D = deg + 1;
  std::vector<double> buf((size_t)K * (3 * dim + 2) * D);
  std::vector<double *> rows((size_t)K * (3 * dim + 2));
  for (size_t r = 0; r < rows.size(); ++r) rows[r] = buf.data() + r * D;
  double **fw[K], **bw[K], **cr[K], *cf[K], *acc[K];
  for (int p = 0; p < K; ++p) {
    double **base = rows.data() + (size_t)p * (3 * dim + 2);
    fw[p] = base; bw[p] = base + dim; cr[p] = base + 2 * dim;
    cf[p] = base[3 * dim]; acc[p] = base[3 * dim + 1];
  }
  for (int p = 0; p < K; ++p)
    for (int j = 0; j < D; ++j) {
      out[p][dim][j] = c[p][j];
      for (int i = 0; i < dim; ++i) out[p][i][j] = 0.0;
    }
  // out[row] += factor * src[srow] (the multiplication only for factors > 1)
  auto add = [&](int row, double **src[], int srow, int e) {
    const double factor = (double)e;
    for (int j = 0; j < D; ++j) {
      double v[K], *pv[K], *po[K];
      for (int p = 0; p < K; ++p) { v[p] = src[p][srow][j]; pv[p] = &v[p]; po[p] = &out[p][row][j]; }
      if (e > 1) mul_d<K>(factor, pv);
      inc<K>(po, (const double **)pv);
    }
  };
  for (int m = 0; m < nbr; ++m) {
    const int nv = nvr[m];
    int *ix = idx[m];
    for (int p = 0; p < K; ++p) std::memcpy(cf[p], cff[p][m], sizeof(double) * D);
    for (int j = 0; j < nv; ++j)
      for (int e = 1; e < exps[m][j]; ++e) {
        double *xin[K];
        for (int p = 0; p < K; ++p) xin[p] = in[p][ix[j]];
        Ref<K>::product(deg, xin, cf, acc);
        for (int p = 0; p < K; ++p) std::memcpy(cf[p], acc[p], sizeof(double) * D);
      }
    double **cfr[K];
    double *cfrow[K][1];
    for (int p = 0; p < K; ++p) { cfrow[p][0] = cf[p]; cfr[p] = cfrow[p]; }
    auto expo = [&](int j) { return exps[m][j]; };
    if (nv == 1) {
      double *xin[K], *f0[K];
      for (int p = 0; p < K; ++p) { xin[p] = in[p][ix[0]]; f0[p] = fw[p][0]; }
      Ref<K>::product(deg, xin, cf, f0);
      add(dim, fw, 0, 1);
      add(ix[0], cfr, 0, expo(0));
    } else if (nv == 2) {
      double *x0[K], *x1[K], *f0[K], *f1[K], *b0[K];
      for (int p = 0; p < K; ++p) {
        x0[p] = in[p][ix[0]]; x1[p] = in[p][ix[1]];
        f0[p] = fw[p][0]; f1[p] = fw[p][1]; b0[p] = bw[p][0];
      }
      Ref<K>::product(deg, cf, x0, f0);
      add(ix[1], fw, 0, expo(1));
      Ref<K>::product(deg, cf, x1, b0);
      add(ix[0], bw, 0, expo(0));
      Ref<K>::product(deg, f0, x1, f1);
      add(dim, fw, 1, 1);
    } else {
      Ref<K>::speel(nv, deg, ix, cf, in, fw, bw, cr);
      add(dim, fw, nv - 1, 1);
      add(ix[nv - 1], fw, nv - 2, expo(nv - 1));
      add(ix[0], bw, nv - 3, expo(0));
      for (int k = 1; k < nv - 1; ++k) add(ix[k], cr, k - 1, expo(k));
    }
  }
}

template <int K> void houseqr_k(int nr, int nc, double *A, double *Q, double *R) {
  Mats a(A, K, 1, nr, nc), q(Q, K, 1, nr, nr), r(R, K, 1, nr, nc);
  double **pa[K], **pq[K], **pr[K];
  for (int p = 0; p < K; ++p) { pa[p] = a.mat(p); pq[p] = q.mat(p); pr[p] = r.mat(p); }
  Ref<K>::houseqr(nr, nc, pa, pq, pr);
}

template <int K>
void solve_k(int dim, int degp1, int tail, double *mat, double *rhs, double *sol, double *Q,
             double *R, int *zeroQ, int *noqr, int *up, int *bs, int *nt) {
  Mats m(mat, K, degp1, dim, dim), b(rhs, K, 1, degp1, dim), x(sol, K, 1, degp1, dim);
  Mats q(Q, K, 1, dim, dim), r(R, K, 1, dim, dim);
  std::vector<double> wrk((size_t)K * dim);
  double ***pm[K], **pb[K], **px[K], **pq[K], **pr[K], *pw[K];
  for (int p = 0; p < K; ++p) {
    pm[p] = m.ser(p); pb[p] = b.mat(p); px[p] = x.mat(p); pq[p] = q.mat(p); pr[p] = r.mat(p);
    pw[p] = wrk.data() + (size_t)p * dim;
  }
  bool zq = *zeroQ != 0, nq = *noqr != 0;
  Ref<K>::solve(dim, degp1, tail, pm, pb, px, pq, pr, pw, &zq, &nq, up, bs, nt);
  *zeroQ = zq;
  *noqr = nq;
}

template <int K>
void residue_k(int dim, int degp1, int tail, double *mat, double *rhs, double *sol,
               double *resvec, double *resmax) {
  Mats m(mat, K, degp1, dim, dim), b(rhs, K, 1, degp1, dim), x(sol, K, 1, degp1, dim);
  Mats rv(resvec, K, 1, degp1, dim);
  double ***pm[K], **pb[K], **px[K], **pr[K], *rm[K];
  for (int p = 0; p < K; ++p) {
    pm[p] = m.ser(p); pb[p] = b.mat(p); px[p] = x.mat(p); pr[p] = rv.mat(p);
    rm[p] = resmax + p;
  }
  Ref<K>::residue(dim, degp1, tail, pm, pb, px, pr, rm);
}

// The host branch of the staggered row Newton test (test_dbl2_row_newton,
// Newton/dbl2_newton_method.cpp:1447-1494, with dbl2_row_newton_qrstep :429-594
// and dbl2_update_newton_qrstep :134-202), calling the reference's
// CPU_*_poly_evaldiff and CPU_*_qrbs_solve.
template <int K>
int row_newton_k(int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                 const int *idxptr, const int *idxflat, const int *expflat, double *cst,
                 double *cff, double *input, int *wdeg, int *tails, int *ups, int *bss,
                 double *history, double *Qo, double *Ro) {
  const int D = deg + 1, nmono = monoptr[dim];
  std::vector<int> nv(nvr, nvr + (nmono > 0 ? nmono : 1));
  std::vector<int> idxc(idxflat, idxflat + idxptr[nmono]);
  std::vector<int> expc(idxptr[nmono] > 0 ? idxptr[nmono] : 1, 1);
  if (expflat) std::copy(expflat, expflat + idxptr[nmono], expc.begin());
  std::vector<int *> idx(nmono > 0 ? nmono : 1), exps(nmono > 0 ? nmono : 1);
  for (int m = 0; m < nmono; ++m) {
    idx[m] = idxc.data() + idxptr[m];
    exps[m] = expc.data() + idxptr[m];
  }
  Mats in(input, K, 1, dim, D);                     // input[p][i] -> series
  Mats cf(cff, K, 1, nmono > 0 ? nmono : 1, D);     // cff[p][m]
  std::vector<double> out((size_t)K * dim * (dim + 1) * D);
  Mats ou(out.data(), K, dim, dim + 1, D);          // out[p][i][j]
  std::vector<double> jac((size_t)K * D * dim * dim), rhs((size_t)K * D * dim),
      urhs((size_t)K * D * dim), sol((size_t)K * D * dim), Q((size_t)K * dim * dim),
      R((size_t)K * dim * dim), wrk((size_t)K * dim);
  Mats jm(jac.data(), K, D, dim, dim), ub(urhs.data(), K, 1, D, dim), sx(sol.data(), K, 1, D, dim);
  Mats qm(Q.data(), K, 1, dim, dim), rm(R.data(), K, 1, dim, dim);
  double ***pm[K], **pb[K], **px[K], **pq[K], **pr[K], *pw[K], **pin[K];
  for (int p = 0; p < K; ++p) {
    pm[p] = jm.ser(p); pb[p] = ub.mat(p); px[p] = sx.mat(p); pq[p] = qm.mat(p);
    pr[p] = rm.mat(p); pw[p] = wrk.data() + (size_t)p * dim; pin[p] = in.mat(p);
  }
  bool zeroQ = true, noqr = false;
  int tailidx = 1, wrkdeg = 0, up = 0, bs = 0, stepcnt = 0;
  for (int step = 0; step < nbsteps; ++step) {
    const int degp1 = wrkdeg + 1;
    for (int i = 0; i < dim; ++i) {
      double *c[K], **f[K], **o[K];
      for (int p = 0; p < K; ++p) {
        c[p] = cst + ((size_t)p * dim + i) * D;
        f[p] = cf.mat(p) + monoptr[i];
        o[p] = ou.ser(p)[i];
      }
      if (expflat)
        poly_evaldiff_exp<K>(dim, monoptr[i + 1] - monoptr[i], wrkdeg, nv.data() + monoptr[i],
                             idx.data() + monoptr[i], exps.data() + monoptr[i], c, f, pin, o);
      else
        Ref<K>::evaldiff(dim, monoptr[i + 1] - monoptr[i], wrkdeg, nv.data() + monoptr[i],
                         idx.data() + monoptr[i], c, f, pin, o);
    }
    // dbl2_map_evaldiff_output and rhs = -funval, then the saved copy urhs
    for (int p = 0; p < K; ++p)
      for (int i = 0; i < dim; ++i)
        for (int k = 0; k < degp1; ++k) {
          rhs[((size_t)p * D + k) * dim + i] = -ou.ser(p)[i][dim][k];
          for (int j = 0; j < dim; ++j)
            jac[(((size_t)p * D + k) * dim + i) * dim + j] = ou.ser(p)[i][j][k];
        }
    for (int p = 0; p < K; ++p)
      for (int k = 0; k < degp1; ++k)
        for (int i = 0; i < dim; ++i) {
          urhs[((size_t)p * D + k) * dim + i] = rhs[((size_t)p * D + k) * dim + i];
          sol[((size_t)p * D + k) * dim + i] = 0.0;
        }
    int newtail = tailidx;
    Ref<K>::solve(dim, degp1, tailidx, pm, pb, px, pq, pr, pw, &zeroQ, &noqr, &up, &bs, &newtail);
    tailidx = newtail;
    for (int j = tailidx - 1; j < degp1; ++j)
      for (int i = 0; i < dim; ++i) {
        double *x[K];
        const double *dx[K];
        for (int p = 0; p < K; ++p) {
          x[p] = &pin[p][i][j];
          dx[p] = &px[p][j][i];
        }
        inc<K>(x, dx);
      }
    if (wdeg) wdeg[step] = wrkdeg;
    if (tails) tails[step] = tailidx;
    if (ups) ups[step] = up;
    if (bss) bss[step] = bs;
    if (history) std::memcpy(history + (size_t)step * K * dim * D, input, sizeof(double) * K * dim * D);
    ++stepcnt;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  if (Qo) std::memcpy(Qo, Q.data(), sizeof(double) * K * dim * dim);
  if (Ro) std::memcpy(Ro, R.data(), sizeof(double) * K * dim * dim);
  return stepcnt;
}

// The host branch of the staggered column Newton test
// (test_dbl2_column_newton, Newton/dbl2_newton_method.cpp:1036-1287, with
// dbl2_column_newton_qrstep :259-427 and dbl2_update_newton_qrstep
// :134-202): a monomial system (nbrcol = 1) by CPU_*_evaluate_monomials on
// unit coefficients and *_linearize_evaldiff_output, else the columns by
// CPU_*_evaluate_columns and *_define_rhs; then CPU_*_qrbs_solve and the
// update of the series.  nvr[c*dim+j], idx from idxptr[c*dim+j], exps
// parallel to idx (nbrcol = 1 only; NULL: all ones), cff [K][nbrcol*dim][D]
// (nbrcol > 1), mb [K][dim][D], input [K][dim][D] in/out.
template <int K>
int column_newton_k(int dim, int nbrcol, int deg, int nbsteps, const int *nvrflat,
                    const int *idxptr, const int *idxflat, const int *expflat, double *cff,
                    double *mbp, double *input, int *wdeg, int *tails, int *ups, int *bss,
                    double *history, double *Qo, double *Ro) {
  const int D = deg + 1, nmono = nbrcol * dim;
  std::vector<int> nvc(nvrflat, nvrflat + nmono), idxc(idxflat, idxflat + idxptr[nmono]);
  std::vector<int> expc(idxptr[nmono] > 0 ? idxptr[nmono] : 1, 1), facc(expc.size(), 0);
  if (expflat) std::copy(expflat, expflat + idxptr[nmono], expc.begin());
  for (size_t j = 0; j < expc.size(); ++j) facc[j] = expc[j] > 1 ? expc[j] - 1 : 0;
  std::vector<int *> nvcol(nbrcol), ixm(nmono), exm(nmono), fcm(nmono);
  std::vector<int **> ixcol(nbrcol);
  std::vector<int> nbrfac(dim, 0);
  for (int m = 0; m < nmono; ++m) {
    ixm[m] = idxc.data() + idxptr[m];
    exm[m] = expc.data() + idxptr[m];
    fcm[m] = facc.data() + idxptr[m];
  }
  for (int c = 0; c < nbrcol; ++c) {
    nvcol[c] = nvc.data() + (size_t)c * dim;
    ixcol[c] = ixm.data() + (size_t)c * dim;
  }
  for (int i = 0; nbrcol == 1 && i < dim; ++i)
    for (int j = 0; j < nvc[i]; ++j) nbrfac[i] += facc[idxptr[i] + j] > 0;
  Mats in(input, K, 1, dim, D), mb(mbp, K, 1, dim, D);
  std::vector<double> cf1((size_t)K * dim * D), accb((size_t)K * (dim + 1) * D);
  Mats unit(cf1.data(), K, 1, dim, D), acc(accb.data(), K, 1, dim + 1, D);
  std::vector<double> zc((size_t)K * (nmono > 0 ? nmono : 1) * D, 0.0);
  Mats cc(cff ? cff : zc.data(), K, nbrcol, dim, D);
  std::vector<double> out((size_t)K * dim * (dim + 1) * D), fv((size_t)K * dim * D);
  Mats ou(out.data(), K, dim, dim + 1, D), fm(fv.data(), K, 1, dim, D);
  std::vector<double> jac((size_t)K * D * dim * dim), rhs((size_t)K * D * dim),
      sol((size_t)K * D * dim), Q((size_t)K * dim * dim), R((size_t)K * dim * dim),
      wrk((size_t)K * dim);
  Mats jm(jac.data(), K, D, dim, dim), ub(rhs.data(), K, 1, D, dim), sx(sol.data(), K, 1, D, dim);
  Mats qm(Q.data(), K, 1, dim, dim), rm(R.data(), K, 1, dim, dim);
  double ***pm[K], **pb[K], **px[K], **pq[K], **pr[K], *pw[K], **pin[K], **pmb[K], **pfv[K];
  double **pun[K], *pacc[K], ***pou[K], **pac[K], ***pcc[K];
  for (int p = 0; p < K; ++p) {
    pm[p] = jm.ser(p); pb[p] = ub.mat(p); px[p] = sx.mat(p); pq[p] = qm.mat(p);
    pr[p] = rm.mat(p); pw[p] = wrk.data() + (size_t)p * dim; pin[p] = in.mat(p);
    pmb[p] = mb.mat(p); pfv[p] = fm.mat(p); pun[p] = unit.mat(p); pacc[p] = acc.mat(p)[0];
    pou[p] = ou.ser(p); pac[p] = acc.mat(p); pcc[p] = cc.ser(p);
  }
  bool zeroQ = true, noqr = false;
  int tailidx = 1, wrkdeg = 0, up = 0, bs = 0, stepcnt = 0;
  for (int step = 0; step < nbsteps; ++step) {
    const int degp1 = wrkdeg + 1;
    if (nbrcol == 1) {
      // dbl2_unit_series_vector (Newton/dbl2_newton_testers.cpp:66-80)
      for (int p = 0; p < K; ++p)
        for (int i = 0; i < dim; ++i)
          for (int k = 0; k <= wrkdeg; ++k) pun[p][i][k] = (p == 0 && k == 0) ? 1.0 : 0.0;
      Ref<K>::monomials(dim, wrkdeg, nvc.data(), ixm.data(), exm.data(), nbrfac.data(),
                        fcm.data(), pun, pacc, pin, pou);
      for (int p = 0; p < K; ++p)
        for (int k = 0; k < degp1; ++k)
          for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) pm[p][k][i][j] = 0.0;
      Ref<K>::linearize(dim, degp1, nvc.data(), ixm.data(), pmb, pou, pfv, pb, pm);
    } else {
      Ref<K>::columns(dim, wrkdeg, nbrcol, nvcol.data(), ixcol.data(), pcc, pac, pin, pfv, pm);
      Ref<K>::define_rhs(dim, degp1, pmb, pfv, pb);
    }
    for (int p = 0; p < K; ++p)
      for (int k = 0; k < degp1; ++k)
        for (int i = 0; i < dim; ++i) px[p][k][i] = 0.0;
    int newtail = tailidx;
    Ref<K>::solve(dim, degp1, tailidx, pm, pb, px, pq, pr, pw, &zeroQ, &noqr, &up, &bs, &newtail);
    tailidx = newtail;
    for (int j = tailidx - 1; j < degp1; ++j)
      for (int i = 0; i < dim; ++i) {
        double *x[K];
        const double *dx[K];
        for (int p = 0; p < K; ++p) {
          x[p] = &pin[p][i][j];
          dx[p] = &px[p][j][i];
        }
        inc<K>(x, dx);
      }
    if (wdeg) wdeg[step] = wrkdeg;
    if (tails) tails[step] = tailidx;
    if (ups) ups[step] = up;
    if (bss) bss[step] = bs;
    if (history) std::memcpy(history + (size_t)step * K * dim * D, input, sizeof(double) * K * dim * D);
    ++stepcnt;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  if (Qo) std::memcpy(Qo, Q.data(), sizeof(double) * K * dim * dim);
  if (Ro) std::memcpy(Ro, R.data(), sizeof(double) * K * dim * dim);
  return stepcnt;
}

template <int K>
void cx_houseqr_k(int nrows, int ncols, double *A, double *Q, double *R) {
  Mats a(A, 2 * K, 1, nrows, ncols), q(Q, 2 * K, 1, nrows, nrows), r(R, 2 * K, 1, nrows, ncols);
  double **pa[2 * K], **pq[2 * K], **pr[2 * K];
  for (int p = 0; p < 2 * K; ++p) { pa[p] = a.mat(p); pq[p] = q.mat(p); pr[p] = r.mat(p); }
  CRef<K>::houseqr(nrows, ncols, pa, pq, pr);
}

template <int K>
void cx_solve_k(int dim, int degp1, int tail, double *mat, double *rhs, double *sol, double *Q,
                double *R, int *zeroQ, int *noqr, int *up, int *bs, int *nt) {
  const int P = 2 * K;
  Mats m(mat, P, degp1, dim, dim), b(rhs, P, 1, degp1, dim), x(sol, P, 1, degp1, dim);
  Mats q(Q, P, 1, dim, dim), r(R, P, 1, dim, dim);
  std::vector<double> wrk((size_t)P * dim);
  double ***pm[P], **pb[P], **px[P], **pq[P], **pr[P], *pw[P];
  for (int p = 0; p < P; ++p) {
    pm[p] = m.ser(p); pb[p] = b.mat(p); px[p] = x.mat(p); pq[p] = q.mat(p); pr[p] = r.mat(p);
    pw[p] = wrk.data() + (size_t)p * dim;
  }
  bool zq = *zeroQ != 0, nq = *noqr != 0;
  CRef<K>::solve(dim, degp1, tail, pm, pb, px, pq, pr, pw, &zq, &nq, up, bs, nt);
  *zeroQ = zq;
  *noqr = nq;
}

template <int K>
void cx_residue_k(int dim, int degp1, int tail, double *mat, double *rhs, double *sol,
                  double *resvec, double *resmax) {
  const int P = 2 * K;
  Mats m(mat, P, degp1, dim, dim), b(rhs, P, 1, degp1, dim), x(sol, P, 1, degp1, dim);
  Mats rv(resvec, P, 1, degp1, dim);
  double ***pm[P], **pb[P], **px[P], **pr[P], *rm[K];
  for (int p = 0; p < P; ++p) {
    pm[p] = m.ser(p); pb[p] = b.mat(p); px[p] = x.mat(p); pr[p] = rv.mat(p);
  }
  for (int p = 0; p < K; ++p) rm[p] = resmax + p;
  CRef<K>::residue(dim, degp1, tail, pm, pb, px, pr, rm);
}

// The host branch of the staggered complex row Newton test
// (test_cmplx2_row_newton, Newton/cmplx2_newton_method.cpp:1759, with
// cmplx2_row_newton_qrstep :565-790, cmplx2_update_newton_qrstep :163-328
// and cmplx2_update_series, Newton/dbl2_newton_testers.cpp:151-180).
template <int K>
int cx_row_newton_k(int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                    const int *idxptr, const int *idxflat, double *cst, double *cff,
                    double *input, int *wdeg, int *tails, int *ups, int *bss, double *history,
                    double *Qo, double *Ro) {
  const int D = deg + 1, nmono = monoptr[dim], P = 2 * K;
  std::vector<int> nv(nvr, nvr + (nmono > 0 ? nmono : 1));
  std::vector<int> idxc(idxflat, idxflat + idxptr[nmono]);
  std::vector<int *> idx(nmono > 0 ? nmono : 1);
  for (int m = 0; m < nmono; ++m) idx[m] = idxc.data() + idxptr[m];
  Mats in(input, P, 1, dim, D);
  Mats cf(cff, P, 1, nmono > 0 ? nmono : 1, D);
  std::vector<double> out((size_t)P * dim * (dim + 1) * D);
  Mats ou(out.data(), P, dim, dim + 1, D);
  std::vector<double> jac((size_t)P * D * dim * dim), urhs((size_t)P * D * dim),
      sol((size_t)P * D * dim), Q((size_t)P * dim * dim), R((size_t)P * dim * dim),
      wrk((size_t)P * dim);
  Mats jm(jac.data(), P, D, dim, dim), ub(urhs.data(), P, 1, D, dim), sx(sol.data(), P, 1, D, dim);
  Mats qm(Q.data(), P, 1, dim, dim), rm(R.data(), P, 1, dim, dim);
  double ***pm[P], **pb[P], **px[P], **pq[P], **pr[P], *pw[P], **pin[P];
  for (int p = 0; p < P; ++p) {
    pm[p] = jm.ser(p); pb[p] = ub.mat(p); px[p] = sx.mat(p); pq[p] = qm.mat(p);
    pr[p] = rm.mat(p); pw[p] = wrk.data() + (size_t)p * dim; pin[p] = in.mat(p);
  }
  bool zeroQ = true, noqr = false;
  int tailidx = 1, wrkdeg = 0, up = 0, bs = 0, stepcnt = 0;
  for (int step = 0; step < nbsteps; ++step) {
    const int degp1 = wrkdeg + 1;
    for (int i = 0; i < dim; ++i) {
      double *c[P], **f[P], **o[P];
      for (int p = 0; p < P; ++p) {
        c[p] = cst + ((size_t)p * dim + i) * D;
        f[p] = cf.mat(p) + monoptr[i];
        o[p] = ou.ser(p)[i];
      }
      CEval<K>::evaldiff(dim, monoptr[i + 1] - monoptr[i], wrkdeg, nv.data() + monoptr[i],
                         idx.data() + monoptr[i], c, f, pin, o);
    }
    for (int p = 0; p < P; ++p)
      for (int i = 0; i < dim; ++i)
        for (int k = 0; k < degp1; ++k) {
          urhs[((size_t)p * D + k) * dim + i] = -ou.ser(p)[i][dim][k];
          sol[((size_t)p * D + k) * dim + i] = 0.0;
          for (int j = 0; j < dim; ++j)
            jac[(((size_t)p * D + k) * dim + i) * dim + j] = ou.ser(p)[i][j][k];
        }
    int newtail = tailidx;
    CRef<K>::solve(dim, degp1, tailidx, pm, pb, px, pq, pr, pw, &zeroQ, &noqr, &up, &bs, &newtail);
    tailidx = newtail;
    for (int j = tailidx - 1; j < degp1; ++j)
      for (int i = 0; i < dim; ++i) {
        double *xr[K], *xi[K];
        const double *dr[K], *di[K];
        for (int p = 0; p < K; ++p) {
          xr[p] = &pin[p][i][j];
          xi[p] = &pin[K + p][i][j];
          dr[p] = &px[p][j][i];
          di[p] = &px[K + p][j][i];
        }
        inc<K>(xr, dr);
        inc<K>(xi, di);
      }
    if (wdeg) wdeg[step] = wrkdeg;
    if (tails) tails[step] = tailidx;
    if (ups) ups[step] = up;
    if (bss) bss[step] = bs;
    if (history) std::memcpy(history + (size_t)step * P * dim * D, input, sizeof(double) * P * dim * D);
    ++stepcnt;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  if (Qo) std::memcpy(Qo, Q.data(), sizeof(double) * P * dim * dim);
  if (Ro) std::memcpy(Ro, R.data(), sizeof(double) * P * dim * dim);
  return stepcnt;
}

}  // namespace

extern "C" {

int ref_cx_houseqr(int K, int nrows, int ncols, double *A, double *Q, double *R) {
  switch (K) {
    case 2: cx_houseqr_k<2>(nrows, ncols, A, Q, R); return 0;
    case 4: cx_houseqr_k<4>(nrows, ncols, A, Q, R); return 0;
    case 8: cx_houseqr_k<8>(nrows, ncols, A, Q, R); return 0;
  }
  return -1;
}

int ref_cx_qrbs_solve(int K, int dim, int degp1, int tailidx, double *mat, double *rhs,
                      double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                      int *bsidx, int *newtail) {
  switch (K) {
#define CASE(KK) \
    case KK: cx_solve_k<KK>(dim, degp1, tailidx, mat, rhs, sol, Q, R, zeroQ, noqr, upidx, bsidx, newtail); return 0;
    CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

int ref_cx_linear_residue(int K, int dim, int degp1, int tailidx, double *mat, double *rhs,
                          double *sol, double *resvec, double *resmax) {
  switch (K) {
#define CASE(KK) case KK: cx_residue_k<KK>(dim, degp1, tailidx, mat, rhs, sol, resvec, resmax); return 0;
    CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

int ref_cx_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                      const int *idxptr, const int *idx, double *cst, double *cff, double *input,
                      int *wdeg, int *tail, int *up, int *bs, double *history, double *Q,
                      double *R) {
  if (K == 2)
    return cx_row_newton_k<2>(dim, deg, nbsteps, monoptr, nvr, idxptr, idx, cst, cff, input, wdeg,
                              tail, up, bs, history, Q, R);
  if (K == 4)
    return cx_row_newton_k<4>(dim, deg, nbsteps, monoptr, nvr, idxptr, idx, cst, cff, input, wdeg,
                              tail, up, bs, history, Q, R);
  return -1;
}


int ref_houseqr(int K, int nrows, int ncols, double *A, double *Q, double *R) {
  switch (K) {
    case 1: houseqr_k<1>(nrows, ncols, A, Q, R); return 0;
    case 2: houseqr_k<2>(nrows, ncols, A, Q, R); return 0;
    case 4: houseqr_k<4>(nrows, ncols, A, Q, R); return 0;
    case 8: houseqr_k<8>(nrows, ncols, A, Q, R); return 0;
  }
  return -1;
}

int ref_qrbs_solve(int K, int dim, int degp1, int tailidx, double *mat, double *rhs,
                   double *sol, double *Q, double *R, int *zeroQ, int *noqr, int *upidx,
                   int *bsidx, int *newtail) {
  switch (K) {
#define CASE(KK) \
    case KK: solve_k<KK>(dim, degp1, tailidx, mat, rhs, sol, Q, R, zeroQ, noqr, upidx, bsidx, newtail); return 0;
    CASE(1) CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

int ref_linear_residue(int K, int dim, int degp1, int tailidx, double *mat, double *rhs,
                       double *sol, double *resvec, double *resmax) {
  switch (K) {
#define CASE(KK) case KK: residue_k<KK>(dim, degp1, tailidx, mat, rhs, sol, resvec, resmax); return 0;
    CASE(1) CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

int ref_row_newton(int K, int dim, int deg, int nbsteps, const int *monoptr, const int *nvr,
                   const int *idxptr, const int *idx, const int *exps, double *cst, double *cff,
                   double *input, int *wdeg, int *tail, int *up, int *bs, double *history,
                   double *Q, double *R) {
  switch (K) {
#define CASE(KK)                                                                              \
    case KK:                                                                                  \
      return row_newton_k<KK>(dim, deg, nbsteps, monoptr, nvr, idxptr, idx, exps, cst, cff,   \
                              input, wdeg, tail, up, bs, history, Q, R);
    CASE(1) CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

// The reference's real double-double row Newton test system
// (dbl2_test_rows, Newton/test_dbl2_newton.cpp:191-212, with
// test_dbl2_row_newton :1289-1301 and dbl2_row_setup :823-991, mode 1):
// cyclic supports from make_polynomial_indices, random exponential
// coefficient series with random leading coefficients, a random exponential
// test solution with the constants adjusted so that it is a root, and the
// start series = leading coefficients of the test solution + 1e-16.
// Sizes: monoptr[dim+1], nvr/idxptr[nmono] (+1 for idxptr), idx[nidx];
// call with cst == nullptr to get nmono (return) and nidx (*nidx_out).
int ref_dbl2_cyclic_row_system(unsigned seed, int dim, int deg, int *monoptr, int *nvr,
                               int *idxptr, int *idxflat, int *nidx_out, double *cst,
                               double *cff, double *testsol, double *start) {
  const int D = deg + 1;
  std::vector<int> nbr(dim);
  std::vector<int *> nv(dim);
  std::vector<int **> ix(dim);
  make_polynomial_indices(dim, nbr.data(), nv.data(), ix.data());
  int nmono = 0, nidx = 0;
  for (int i = 0; i < dim; ++i) {
    nmono += nbr[i];
    for (int j = 0; j < nbr[i]; ++j) nidx += nv[i][j];
  }
  if (nidx_out) *nidx_out = nidx;
  if (cst == nullptr) return nmono;
  int m = 0, o = 0;
  monoptr[0] = 0;
  for (int i = 0; i < dim; ++i) {
    for (int j = 0; j < nbr[i]; ++j, ++m) {
      nvr[m] = nv[i][j];
      idxptr[m] = o;
      for (int k = 0; k < nv[i][j]; ++k) idxflat[o++] = ix[i][j][k];
    }
    monoptr[i + 1] = m;
  }
  idxptr[m] = o;
  srand(seed);
  std::vector<double *> csthi(dim), cstlo(dim);
  std::vector<double **> cffhi(dim), cfflo(dim);
  dbl2_make_coefficients(dim, deg, nbr.data(), nv.data(), ix.data(), csthi.data(),
                         cstlo.data(), cffhi.data(), cfflo.data(), 0);
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < nbr[i]; ++j) random_double_double(&cffhi[i][j][0], &cfflo[i][j][0]);
  std::vector<double *> shi(dim), slo(dim);
  for (int i = 0; i < dim; ++i) {
    shi[i] = new double[D];
    slo[i] = new double[D];
  }
  make_real2_exponentials(dim, deg, shi.data(), slo.data());
  std::vector<double> ob((size_t)2 * (dim + 1) * D);
  Mats ou(ob.data(), 2, 1, dim + 1, D);
  for (int i = 0; i < dim; ++i) {
    double el = 0.0;
    CPU_dbl2_poly_evaldiff(dim, nbr[i], deg, nv[i], ix[i], csthi[i], cstlo[i], cffhi[i],
                           cfflo[i], shi.data(), slo.data(), ou.mat(0), ou.mat(1), &el, 0);
    for (int j = 0; j <= deg; ++j)
      ddf_dec(&csthi[i][j], &cstlo[i][j], ou.mat(0)[dim][j], ou.mat(1)[dim][j]);
  }
  m = 0;
  for (int i = 0; i < dim; ++i) {
    for (int k = 0; k < D; ++k) {
      cst[(size_t)i * D + k] = csthi[i][k];
      cst[((size_t)dim + i) * D + k] = cstlo[i][k];
      testsol[(size_t)i * D + k] = shi[i][k];
      testsol[((size_t)dim + i) * D + k] = slo[i][k];
      start[(size_t)i * D + k] = 0.0;
      start[((size_t)dim + i) * D + k] = 0.0;
    }
    // real2_start_series_vector (Newton/dbl2_newton_testers.cpp:25-40)
    double hi = shi[i][0], lo = slo[i][0];
    ddf_inc(&hi, &lo, 1.0e-16, 0.0);
    start[(size_t)i * D] = hi;
    start[((size_t)dim + i) * D] = lo;
    for (int j = 0; j < nbr[i]; ++j, ++m)
      for (int k = 0; k < D; ++k) {
        cff[(size_t)m * D + k] = cffhi[i][j][k];
        cff[((size_t)nmono + m) * D + k] = cfflo[i][j][k];
      }
  }
  return nmono;
}

int ref_column_newton(int K, int dim, int nbrcol, int deg, int nbsteps, const int *nvr,
                      const int *idxptr, const int *idx, const int *exps, double *cff,
                      double *mb, double *input, int *wdeg, int *tail, int *up, int *bs,
                      double *history, double *Q, double *R) {
  switch (K) {
#define CASE(KK)                                                                              \
    case KK:                                                                                  \
      return column_newton_k<KK>(dim, nbrcol, deg, nbsteps, nvr, idxptr, idx, exps, cff, mb,  \
                                 input, wdeg, tail, up, bs, history, Q, R);
    CASE(1) CASE(2) CASE(4) CASE(8)
#undef CASE
  }
  return -1;
}

// The reference's monomial test system (make_monomial_system,
// Newton/unimodular_matrices.cpp:118-186, as dbl2_test_columns calls it,
// Newton/test_dbl2_newton.cpp:108-118, positive values): after srand(seed)
// the exponent matrix rowsA [dim][dim] (nbritr > 0: that many random
// unimodular row operations with multipliers below size; -1: lower
// triangular ones; -2: two-variable monomials), and per row its nonzero
// entries as nvr[dim], idx and exp [dim*dim] (row i from i*dim).
int ref_monomial_system(unsigned seed, int dim, int size, int nbritr, int *rowsA, int *nvr,
                        int *idx, int *exp) {
  std::vector<int *> A(dim), ix(dim), ex(dim), ef(dim);
  std::vector<int> nv(dim), nf(dim);
  srand(seed);
  make_monomial_system(dim, size, 1, nbritr, nv.data(), ix.data(), ex.data(), nf.data(),
                       ef.data(), A.data(), 0);
  for (int i = 0; i < dim; ++i) {
    nvr[i] = nv[i];
    for (int j = 0; j < dim; ++j) rowsA[(size_t)i * dim + j] = A[i][j];
    for (int j = 0; j < nv[i]; ++j) {
      idx[(size_t)i * dim + j] = ix[i][j];
      exp[(size_t)i * dim + j] = ex[i][j];
    }
  }
  return 0;
}

// The reference's double-double column setup (dbl2_column_setup,
// Newton/dbl2_newton_method.cpp:766-821, mode 1, after srand(seed)):
// the exponential test solution (make_real2_exponentials), for nbrcol > 1
// the random coefficients of make_real2_coefficients (drawn first, as
// test_dbl2_column_newton :1057-1058 does) on the cyclic columns
// (make_cyclic_variables/_columns, Newton/cyclic_columns.cpp:10-23), the
// right-hand side mb as the system at the test solution
// (evaluate_real2_monomials / evaluate_real2_columns) and the start series
// (real2_start_series_vector).  nbrcol = 1 takes rowsA [dim][dim]; nbrcol = 2
// takes, nbrcol > 2 writes (the cyclic columns) nvr [nbrcol*dim] and idx
// [nbrcol*dim*dim] (monomial m from m*dim).
// cff [2][nbrcol*dim][D], testsol, mb, start [2][dim][D].
int ref_dbl2_column_setup(unsigned seed, int dim, int deg, int nbrcol, const int *rowsA,
                          int *nvr, int *idx, double *cff, double *testsol, double *mb,
                          double *start) {
  const int D = deg + 1;
  std::vector<int *> A(dim);
  std::vector<int> Ab((size_t)dim * dim, 0);
  for (int i = 0; i < dim; ++i) A[i] = Ab.data() + (size_t)i * dim;
  if (rowsA) std::copy(rowsA, rowsA + (size_t)dim * dim, Ab.begin());
  std::vector<int> nvb((size_t)nbrcol * dim, 0), ixb((size_t)nbrcol * dim * dim, 0);
  std::vector<int *> nvc(nbrcol);
  std::vector<int **> ixc(nbrcol);
  std::vector<int *> ixm((size_t)nbrcol * dim);
  for (int c = 0; c < nbrcol; ++c) {
    nvc[c] = nvb.data() + (size_t)c * dim;
    ixc[c] = ixm.data() + (size_t)c * dim;
    for (int j = 0; j < dim; ++j) ixm[(size_t)c * dim + j] = ixb.data() + ((size_t)c * dim + j) * dim;
  }
  srand(seed);
  std::vector<double> cb((size_t)2 * nbrcol * dim * D, 0.0);
  Mats cm(cb.data(), 2, nbrcol, dim, D);
  if (nbrcol != 1) make_real2_coefficients(nbrcol, dim, cm.ser(0), cm.ser(1));
  if (nbrcol == 2) {  // the caller's two columns
    std::copy(nvr, nvr + nvb.size(), nvb.begin());
    std::copy(idx, idx + ixb.size(), ixb.begin());
  } else if (nbrcol > 2) {
    make_cyclic_variables(nbrcol, dim, nvc.data());
    make_cyclic_columns(nbrcol, dim, nvc.data(), ixc.data());
  }
  std::vector<double *> shi(dim), slo(dim), mhi(dim), mlo(dim);
  std::vector<double> sb((size_t)2 * dim * D), mbb((size_t)2 * dim * D);
  for (int i = 0; i < dim; ++i) {
    shi[i] = sb.data() + (size_t)i * D; slo[i] = sb.data() + ((size_t)dim + i) * D;
    mhi[i] = mbb.data() + (size_t)i * D; mlo[i] = mbb.data() + ((size_t)dim + i) * D;
  }
  make_real2_exponentials(dim, deg, shi.data(), slo.data());
  for (int i = 0; i < dim; ++i) {
    mhi[i][0] = 1.0; mlo[i][0] = 0.0;
    for (int k = 1; k < D; ++k) { mhi[i][k] = 0.0; mlo[i][k] = 0.0; }
  }
  if (nbrcol == 1)
    evaluate_real2_monomials(dim, deg, A.data(), shi.data(), slo.data(), mhi.data(), mlo.data());
  else
    evaluate_real2_columns(dim, deg, nbrcol, nvc.data(), ixc.data(), A.data(), cm.ser(0),
                           cm.ser(1), shi.data(), slo.data(), mhi.data(), mlo.data(), 0);
  std::copy(sb.begin(), sb.end(), testsol);
  std::copy(mbb.begin(), mbb.end(), mb);
  if (cff) std::copy(cb.begin(), cb.end(), cff);
  for (size_t t = 0; t < sb.size(); ++t) start[t] = 0.0;
  for (int i = 0; i < dim; ++i) {
    // real2_start_series_vector (Newton/dbl2_newton_testers.cpp:25-40)
    double hi = shi[i][0], lo = slo[i][0];
    ddf_inc(&hi, &lo, 1.0e-16, 0.0);
    start[(size_t)i * D] = hi;
    start[((size_t)dim + i) * D] = lo;
  }
  if (nbrcol > 2) {
    for (size_t t = 0; t < nvb.size(); ++t) nvr[t] = nvb[t];
    for (size_t t = 0; t < ixb.size(); ++t) idx[t] = ixb[t];
  }
  return 0;
}

}  // extern "C"
