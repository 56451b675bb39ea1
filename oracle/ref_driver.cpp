// Flat extern "C" wrappers around the reference's own CPU functions, compiled
// together with the reference sources (read in place under the reference
// checkout) into oracle/_ref/libphcref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and to write
// golden vectors.  Never linked into the product.
//
// Array convention (same as the oracle and the product C-ABI): a number with
// K limbs is stored limb-planar; complex numbers store the K real limbs then
// the K imaginary limbs, so an array of m numbers has P = K (real) or 2K
// (complex) planes of m doubles each, plane p at offset p*m.
#include <vector>
#include <cstdlib>
#include "double_double_functions.h"
#include "quad_double_functions.h"
#include "octo_double_functions.h"
#include "dbl2_convolutions_host.h"
#include "dbl4_convolutions_host.h"
#include "dbl8_convolutions_host.h"
#include "dbl_polynomials_host.h"
#include "dbl2_polynomials_host.h"
#include "dbl4_polynomials_host.h"
#include "dbl8_polynomials_host.h"

namespace {

// pointer rows: plane p, row r of length len in a flat [P][rows][len] array
struct Rows {
  std::vector<std::vector<double *>> r;
  Rows(double *base, int P, int rows, int len) : r(P) {
    for (int p = 0; p < P; ++p)
      for (int i = 0; i < rows; ++i) r[p].push_back(base + ((size_t)p * rows + i) * len);
  }
  double **operator[](int p) { return r[p].data(); }
};

}  // namespace

extern "C" {

// op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(a).  limb-planar arrays of n numbers
void ref_md_batch(int op, int K, int n, const double *a, const double *b, double *c) {
  for (int j = 0; j < n; ++j) {
    double x[8], y[8], z[8];
    for (int i = 0; i < K; ++i) { x[i] = a[i * n + j]; y[i] = b ? b[i * n + j] : 0.0; }
    if (K == 2) {
      if (op == 0) ddf_add(x[0], x[1], y[0], y[1], &z[0], &z[1]);
      else if (op == 1) ddf_sub(x[0], x[1], y[0], y[1], &z[0], &z[1]);
      else if (op == 2) ddf_mul(x[0], x[1], y[0], y[1], &z[0], &z[1]);
      else if (op == 3) ddf_div(x[0], x[1], y[0], y[1], &z[0], &z[1]);
      else ddf_sqrt(x[0], x[1], &z[0], &z[1]);
    } else if (K == 4) {
#define Q4(v) v[0], v[1], v[2], v[3]
#define Q4P(v) &v[0], &v[1], &v[2], &v[3]
      if (op == 0) qdf_add(Q4(x), Q4(y), Q4P(z));
      else if (op == 1) qdf_sub(Q4(x), Q4(y), Q4P(z));
      else if (op == 2) qdf_mul(Q4(x), Q4(y), Q4P(z));
      else if (op == 3) qdf_div(Q4(x), Q4(y), Q4P(z));
      else qdf_sqrt(Q4(x), Q4P(z));
    } else {
#define O8(v) v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]
#define O8P(v) &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7]
      if (op == 0) odf_add(O8(x), O8(y), O8P(z));
      else if (op == 1) odf_sub(O8(x), O8(y), O8P(z));
      else if (op == 2) odf_mul(O8(x), O8(y), O8P(z));
      else if (op == 3) odf_div(O8(x), O8(y), O8P(z));
      else odf_sqrt(O8(x), O8P(z));
    }
    for (int i = 0; i < K; ++i) c[i * n + j] = z[i];
  }
}

// z = x*y truncated at degree deg; x, y, z are [P][deg+1]
void ref_product(int K, int cplx, int deg, double *x, double *y, double *z) {
  const int D = deg + 1;
#define PL(a, p) ((a) + (size_t)(p) * D)
  if (K == 2 && !cplx)
    CPU_dbl2_product(deg, PL(x, 0), PL(x, 1), PL(y, 0), PL(y, 1), PL(z, 0), PL(z, 1));
  else if (K == 2)
    CPU_cmplx2_product(deg, PL(x, 0), PL(x, 1), PL(x, 2), PL(x, 3), PL(y, 0), PL(y, 1),
                       PL(y, 2), PL(y, 3), PL(z, 0), PL(z, 1), PL(z, 2), PL(z, 3));
  else if (K == 4 && !cplx)
    CPU_dbl4_product(deg, PL(x, 0), PL(x, 1), PL(x, 2), PL(x, 3), PL(y, 0), PL(y, 1),
                     PL(y, 2), PL(y, 3), PL(z, 0), PL(z, 1), PL(z, 2), PL(z, 3));
  else if (K == 4)
    CPU_cmplx4_product(deg, PL(x, 0), PL(x, 1), PL(x, 2), PL(x, 3), PL(x, 4), PL(x, 5),
                       PL(x, 6), PL(x, 7), PL(y, 0), PL(y, 1), PL(y, 2), PL(y, 3),
                       PL(y, 4), PL(y, 5), PL(y, 6), PL(y, 7), PL(z, 0), PL(z, 1),
                       PL(z, 2), PL(z, 3), PL(z, 4), PL(z, 5), PL(z, 6), PL(z, 7));
  else if (!cplx)
    CPU_dbl8_product(deg, PL(x, 0), PL(x, 1), PL(x, 2), PL(x, 3), PL(x, 4), PL(x, 5),
                     PL(x, 6), PL(x, 7), PL(y, 0), PL(y, 1), PL(y, 2), PL(y, 3),
                     PL(y, 4), PL(y, 5), PL(y, 6), PL(y, 7), PL(z, 0), PL(z, 1),
                     PL(z, 2), PL(z, 3), PL(z, 4), PL(z, 5), PL(z, 6), PL(z, 7));
  else
    CPU_cmplx8_product(
        deg, PL(x, 0), PL(x, 1), PL(x, 2), PL(x, 3), PL(x, 4), PL(x, 5), PL(x, 6), PL(x, 7),
        PL(x, 8), PL(x, 9), PL(x, 10), PL(x, 11), PL(x, 12), PL(x, 13), PL(x, 14), PL(x, 15),
        PL(y, 0), PL(y, 1), PL(y, 2), PL(y, 3), PL(y, 4), PL(y, 5), PL(y, 6), PL(y, 7),
        PL(y, 8), PL(y, 9), PL(y, 10), PL(y, 11), PL(y, 12), PL(y, 13), PL(y, 14), PL(y, 15),
        PL(z, 0), PL(z, 1), PL(z, 2), PL(z, 3), PL(z, 4), PL(z, 5), PL(z, 6), PL(z, 7),
        PL(z, 8), PL(z, 9), PL(z, 10), PL(z, 11), PL(z, 12), PL(z, 13), PL(z, 14), PL(z, 15));
#undef PL
}

// Polynomial value and gradient at a power series (CPU_*_poly_evaldiff).
// nvr[nbr], idx CSR (idxptr[nbr+1]); cst [P][D]; cff [P][nbr][D];
// input [P][dim][D]; output [P][dim+1][D] with output[dim] the value.
void ref_poly_evaldiff(int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                       const int *idxptr, const int *idxflat, double *cst, double *cff,
                       double *input, double *output) {
  const int D = deg + 1;
  const int P = cplx ? 2 * K : K;
  std::vector<int *> idx(nbr);
  std::vector<int> idxcopy(idxflat, idxflat + idxptr[nbr]);
  for (int i = 0; i < nbr; ++i) idx[i] = idxcopy.data() + idxptr[i];
  std::vector<int> nv(nvr, nvr + nbr);
  Rows C(cff, P, nbr, D), X(input, P, dim, D), Y(output, P, dim + 1, D);
  double el = 0.0;
#define CS(p) (cst + (size_t)(p) * D)
  if (K == 1 && !cplx)
    CPU_dbl_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), C[0], X[0], Y[0], &el, 0);
  else if (K == 2 && !cplx)
    CPU_dbl2_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), C[0], C[1],
                           X[0], X[1], Y[0], Y[1], &el, 0);
  else if (K == 2)
    CPU_cmplx2_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), CS(2),
                             CS(3), C[0], C[1], C[2], C[3], X[0], X[1], X[2], X[3], Y[0],
                             Y[1], Y[2], Y[3], &el, 0);
  else if (K == 4 && !cplx)
    CPU_dbl4_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), CS(2), CS(3),
                           C[0], C[1], C[2], C[3], X[0], X[1], X[2], X[3], Y[0], Y[1], Y[2],
                           Y[3], &el, 0);
  else if (K == 4)
    CPU_cmplx4_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), CS(2),
                             CS(3), CS(4), CS(5), CS(6), CS(7), C[0], C[1], C[2], C[3],
                             C[4], C[5], C[6], C[7], X[0], X[1], X[2], X[3], X[4], X[5],
                             X[6], X[7], Y[0], Y[1], Y[2], Y[3], Y[4], Y[5], Y[6], Y[7],
                             &el, 0);
  else if (!cplx)
    CPU_dbl8_poly_evaldiff(dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), CS(2), CS(3),
                           CS(4), CS(5), CS(6), CS(7), C[0], C[1], C[2], C[3], C[4], C[5],
                           C[6], C[7], X[0], X[1], X[2], X[3], X[4], X[5], X[6], X[7], Y[0],
                           Y[1], Y[2], Y[3], Y[4], Y[5], Y[6], Y[7], &el, 0);
  else
    CPU_cmplx8_poly_evaldiff(
        dim, nbr, deg, nv.data(), idx.data(), CS(0), CS(1), CS(2), CS(3), CS(4), CS(5),
        CS(6), CS(7), CS(8), CS(9), CS(10), CS(11), CS(12), CS(13), CS(14), CS(15), C[0],
        C[1], C[2], C[3], C[4], C[5], C[6], C[7], C[8], C[9], C[10], C[11], C[12], C[13],
        C[14], C[15], X[0], X[1], X[2], X[3], X[4], X[5], X[6], X[7], X[8], X[9], X[10],
        X[11], X[12], X[13], X[14], X[15], Y[0], Y[1], Y[2], Y[3], Y[4], Y[5], Y[6], Y[7],
        Y[8], Y[9], Y[10], Y[11], Y[12], Y[13], Y[14], Y[15], &el, 0);
#undef CS
}

}  // extern "C"
