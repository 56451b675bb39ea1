// Parity of the reference-signature entry points of libphcmd.so
// (include/phcmd_compat.h) with the CPU oracle (oracle/liboracle.so: this
// repository's C restatement of the reference's host code, bit-exact with
// the compiled reference, which the CPU tests check).  Written like the
// reference's testers (Convolutions/test_dbl*_convolutions.cpp,
// Polynomials/test_dbl*_polynomials.cpp, Matrices/test_dbl4_baqr.cpp,
// test_dbl4_tabs.cpp, Newton/test_dbl*_bals.cpp): random input, the CPU
// function and the GPU function on the same data, the largest difference.
// The caller side uses the reference's argument conventions throughout (one
// array per limb and per real/imaginary part, matrices as row pointers,
// job schedules built by ConvolutionJobs / AdditionJobs).
//
// Series products, polynomial evaluation, the block Toeplitz solve and the
// linear residue must agree bit for bit; the blocked QR, the tiled solvers
// and Q^H b split long sums across lanes, so they must agree within
// 64 * dim * eps^K, and the least-squares solve must recover a planted x0.
// Exit status = number of failed checks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/phcmd_compat.h"
#include "baqr.h"
#include "md.h"
#include "newton.h"
#include "series.h"

static int failures = 0;

static void check(bool ok, const char *what, double err, double tol) {
  std::printf("%-52s max diff %.3e  tol %.3e  %s\n", what, err, tol, ok ? "ok" : "FAILED");
  if (!ok) ++failures;
}

// uniform doubles in [-1, 1] from a 64-bit LCG (the top 53 bits)
static uint64_t rng_state = 20261202ULL;
static double uniform() {
  rng_state = rng_state * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(rng_state >> 11) * 0x1p-52 - 1.0;
}

// a K-limb number: a random leading limb plus random lower-order parts,
// renormalized by an md addition
static void random_md(int K, double *x) {
  double a[8] = {uniform(), 0, 0, 0, 0, 0, 0, 0}, b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double s = 0x1p-53;
  for (int l = 1; l < K; ++l, s *= 0x1p-53) b[l - 1] = uniform() * s;
  if (K == 1) x[0] = a[0];
  else omd_add(K, a, b, x);
}

// |leading limb of a - b| in K-limb arithmetic
static double md_diff(int K, const double *a, const double *b) {
  double c[8];
  if (K == 1) return std::fabs(a[0] - b[0]);
  omd_sub(K, a, b, c);
  return std::fabs(c[0]);
}

// P planes of rows x cols doubles, contiguous ([P][rows][cols]), with row
// pointers per plane for the reference-style calls
struct Planes {
  int P, rows, cols;
  std::vector<double> v;
  std::vector<std::vector<double *>> ptr;
  Planes(int P_, int r, int c) : P(P_), rows(r), cols(c), v((size_t)P_ * r * c, 0.0), ptr(P_) {
    for (int q = 0; q < P; ++q) {
      ptr[q].resize(r);
      for (int i = 0; i < r; ++i) ptr[q][i] = &v[((size_t)q * r + i) * c];
    }
  }
  Planes(const Planes &o) : Planes(o.P, o.rows, o.cols) { v = o.v; }
  Planes &operator=(const Planes &) = delete;
  double **p(int q) { return ptr[q].data(); }   // row pointers of plane q
  double *row(int q, int i) { return ptr[q][i]; }
  double *data() { return v.data(); }
  // K-limb number (plane offset off) at (i, j)
  void get(int K, int off, int i, int j, double *x) const {
    for (int l = 0; l < K; ++l) x[l] = v[((size_t)(off + l) * rows + i) * cols + j];
  }
  void set(int K, int off, int i, int j, const double *x) {
    for (int l = 0; l < K; ++l) v[((size_t)(off + l) * rows + i) * cols + j] = x[l];
  }
  // every entry random; K limbs per number, P / K numbers per entry
  void randomize(int K) {
    double x[8];
    for (int g = 0; g < P / K; ++g)
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
          random_md(K, x);
          set(K, g * K, i, j, x);
        }
  }
  // the largest |leading limb of a - b| over all entries (K-limb numbers)
  double diff(int K, const Planes &o) const {
    double m = 0.0, a[8], b[8];
    for (int g = 0; g < P / K; ++g)
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
          get(K, g * K, i, j, a);
          o.get(K, g * K, i, j, b);
          m = std::fmax(m, md_diff(K, a, b));
        }
    return m;
  }
  bool same(const Planes &o) const { return v == o.v; }
};

// ------------------------------------------------------------ products

// z = x*y through the GPU_* entry point and osr_product, K limbs, real and
// complex: bit for bit
static void test_products(int K, int deg) {
  const int D = deg + 1;
  for (int cplx = 0; cplx <= 1; ++cplx) {
    const int P = cplx ? 2 * K : K;
    Planes x(P, 1, D), y(P, 1, D), zc(P, 1, D), zg(P, 1, D);
    x.randomize(K);
    y.randomize(K);
    osr_product(K, cplx, deg, x.data(), y.data(), zc.data());
    double *a[16], *b[16], *c[16];
    for (int q = 0; q < P; ++q) { a[q] = x.row(q, 0); b[q] = y.row(q, 0); c[q] = zg.row(q, 0); }
    const char *name = "";
    if (K == 1 && !cplx) { GPU_dbl_product(a[0], b[0], c[0], deg, 1, D, 0); name = "GPU_dbl_product"; }
    if (K == 1 && cplx) {
      GPU_cmplx_product(a[0], a[1], b[0], b[1], c[0], c[1], deg, 1, D, 0);
      name = "GPU_cmplx_product";
    }
    if (K == 2 && !cplx) {
      GPU_dbl2_product(a[0], a[1], b[0], b[1], c[0], c[1], deg, 1, D, 1);
      name = "GPU_dbl2_product";
    }
    if (K == 2 && cplx) {
      GPU_cmplx2_product(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], c[0], c[1], c[2], c[3],
                         deg, 1, D, 0);
      name = "GPU_cmplx2_product";
    }
    if (K == 4 && !cplx) {
      GPU_dbl4_product(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], c[0], c[1], c[2], c[3],
                       deg, 1, D, 1);
      name = "GPU_dbl4_product";
    }
    if (K == 4 && cplx) {
      GPU_cmplx4_product(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], b[0], b[1], b[2], b[3],
                         b[4], b[5], b[6], b[7], c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
                         deg, 1, D, 0);
      name = "GPU_cmplx4_product";
    }
    if (K == 8 && !cplx) {
      GPU_dbl8_product(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], b[0], b[1], b[2], b[3],
                       b[4], b[5], b[6], b[7], c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
                       deg, 1, D, 1);
      name = "GPU_dbl8_product";
    }
    if (K == 8 && cplx) {
      GPU_cmplx8_product(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10],
                         a[11], a[12], a[13], a[14], a[15], b[0], b[1], b[2], b[3], b[4], b[5],
                         b[6], b[7], b[8], b[9], b[10], b[11], b[12], b[13], b[14], b[15], c[0],
                         c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], c[10], c[11],
                         c[12], c[13], c[14], c[15], deg, 1, D, 0);
      name = "GPU_cmplx8_product";
    }
    char what[96];
    std::snprintf(what, sizeof what, "%s deg %d (bitwise)", name, deg);
    check(zg.same(zc), what, zg.diff(K, zc), 0.0);
  }
}

// --------------------------------------------------- polynomial evaluation

// a polynomial in dim variables with nbr monomials of sorted distinct
// variables (sizes 1..maxnvr) through GPU_*_poly_evaldiff, with the job
// schedules built by the caller as the reference's testers do, against
// osr_poly_evaldiff (CPU_*_poly_evaldiff restated): bit for bit
static void test_poly_evaldiff(int K, int cplx, int dim, int nbr, int maxnvr, int deg) {
  const int D = deg + 1, P = cplx ? 2 * K : K;
  std::vector<int> nvr(nbr), idxptr(nbr + 1, 0), flat;
  std::vector<std::vector<int>> sup(nbr);
  for (int k = 0; k < nbr; ++k) {
    nvr[k] = 1 + (int)((uniform() + 1.0) * 0.5 * maxnvr) % maxnvr;
    std::vector<int> pick(dim);
    for (int i = 0; i < dim; ++i) pick[i] = i;
    for (int i = 0; i < nvr[k]; ++i) {
      const int j = i + (int)((uniform() + 1.0) * 0.5 * (dim - i)) % (dim - i);
      std::swap(pick[i], pick[j]);
    }
    sup[k].assign(pick.begin(), pick.begin() + nvr[k]);
    std::sort(sup[k].begin(), sup[k].end());
    for (int v : sup[k]) flat.push_back(v);
    idxptr[k + 1] = (int)flat.size();
  }
  std::vector<int *> idx(nbr);
  for (int k = 0; k < nbr; ++k) idx[k] = sup[k].data();
  Planes cst(P, 1, D), cff(P, nbr, D), input(P, dim, D), outc(P, dim + 1, D), outg(P, dim + 1, D);
  cst.randomize(K);
  cff.randomize(K);
  input.randomize(K);
  osr_poly_evaldiff(K, cplx, dim, nbr, deg, nvr.data(), idxptr.data(), flat.data(), cst.data(),
                    cff.data(), input.data(), outc.data());
  ConvolutionJobs cnvjobs(dim);
  cnvjobs.make(nbr, nvr.data(), idx.data(), false);
  AdditionJobs addjobs(dim, nbr);
  addjobs.make(nbr, nvr.data(), idx.data(), false);
  double *c[16], **f[16], **x[16], **y[16];
  for (int q = 0; q < P; ++q) {
    c[q] = cst.row(q, 0);
    f[q] = cff.p(q);
    x[q] = input.p(q);
    y[q] = outg.p(q);
  }
  double cnv = -1.0, add = -1.0, ela = -1.0, wall = -1.0;
  const char *name = "";
  if (K == 1 && !cplx) {
    GPU_dbl_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], f[0], x[0], y[0],
                          cnvjobs, addjobs, &cnv, &add, &ela, &wall, 0);
    name = "GPU_dbl_poly_evaldiff";
  } else if (K == 2 && !cplx) {
    GPU_dbl2_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], f[0], f[1],
                           x[0], x[1], y[0], y[1], cnvjobs, addjobs, &cnv, &add, &ela, &wall, 0);
    name = "GPU_dbl2_poly_evaldiff";
  } else if (K == 4 && !cplx) {
    GPU_dbl4_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2], c[3],
                           f[0], f[1], f[2], f[3], x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3],
                           cnvjobs, addjobs, &cnv, &add, &ela, &wall, 0);
    name = "GPU_dbl4_poly_evaldiff";
  } else if (K == 8 && !cplx) {
    GPU_dbl8_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2], c[3],
                           c[4], c[5], c[6], c[7], f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7],
                           x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], y[0], y[1], y[2], y[3],
                           y[4], y[5], y[6], y[7], cnvjobs, addjobs, &cnv, &add, &ela, &wall, 0);
    name = "GPU_dbl8_poly_evaldiff";
  } else if (K == 1 && cplx) {
    GPU_cmplx_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], f[0], f[1],
                            x[0], x[1], y[0], y[1], cnvjobs, addjobs, &cnv, &add, &ela, &wall,
                            0);
    name = "GPU_cmplx_poly_evaldiff";
  } else if (K == 2 && cplx) {
    GPU_cmplx2_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2], c[3],
                             f[0], f[1], f[2], f[3], x[0], x[1], x[2], x[3], y[0], y[1], y[2],
                             y[3], cnvjobs, addjobs, &cnv, &add, &ela, &wall, 0);
    name = "GPU_cmplx2_poly_evaldiff";
  } else {
    GPU_cmplx4_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2], c[3],
                             c[4], c[5], c[6], c[7], f[0], f[1], f[2], f[3], f[4], f[5], f[6],
                             f[7], x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], y[0], y[1],
                             y[2], y[3], y[4], y[5], y[6], y[7], cnvjobs, addjobs, &cnv, &add,
                             &ela, &wall, 0);
    name = "GPU_cmplx4_poly_evaldiff";
  }
  char what[96];
  std::snprintf(what, sizeof what, "%s dim %d nbr %d deg %d (bitwise)", name, dim, nbr, deg);
  check(outg.same(outc), what, outg.diff(K, outc), 0.0);
  if (!(cnv >= 0.0 && add == 0.0 && ela == cnv && wall > 0.0)) {
    std::printf("%s: timers not set\n", name);
    ++failures;
  }
  if (!cplx) return;
  // the vectorized complex evaluation, with the complex schedules
  ComplexConvolutionJobs ccnv(dim);
  ccnv.make(nbr, nvr.data(), idx.data(), false);
  ComplexIncrementJobs cinc(ccnv, false);
  ComplexAdditionJobs cadd(dim, nbr);
  cadd.make(nbr, nvr.data(), idx.data(), false);
  Planes outv(P, dim + 1, D);
  for (int q = 0; q < P; ++q) y[q] = outv.p(q);
  cnv = add = ela = wall = -1.0;
  if (K == 1) {
    GPU_cmplxvectorized_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], f[0],
                                      f[1], x[0], x[1], y[0], y[1], ccnv, cinc, cadd, &cnv, &add,
                                      &ela, &wall, 0);
    name = "GPU_cmplxvectorized_poly_evaldiff";
  } else if (K == 2) {
    GPU_cmplx2vectorized_poly_evaldiff(D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2],
                                       c[3], f[0], f[1], f[2], f[3], x[0], x[1], x[2], x[3], y[0],
                                       y[1], y[2], y[3], ccnv, cinc, cadd, &cnv, &add, &ela, &wall,
                                       0);
    name = "GPU_cmplx2vectorized_poly_evaldiff";
  } else {
    GPU_cmplx4vectorized_poly_evaldiff(
        D, dim, nbr, deg, nvr.data(), idx.data(), c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
        f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], x[0], x[1], x[2], x[3], x[4], x[5], x[6],
        x[7], y[0], y[1], y[2], y[3], y[4], y[5], y[6], y[7], ccnv, cinc, cadd, &cnv, &add, &ela,
        &wall, 0);
    name = "GPU_cmplx4vectorized_poly_evaldiff";
  }
  std::snprintf(what, sizeof what, "%s dim %d nbr %d deg %d (bitwise)", name, dim, nbr, deg);
  check(outv.same(outc), what, outv.diff(K, outc), 0.0);
  if (!(cnv >= 0.0 && add == 0.0 && ela == cnv && wall > 0.0)) {
    std::printf("%s: timers not set\n", name);
    ++failures;
  }
}

// ------------------------------------------------------------ QR, solver

static void qr_gpu(int K, Planes &A, Planes &Q, Planes &R, int szt) {
  const int m = A.rows, n = A.cols;
  double t[10] = {0}, wall = 0.0;
  long long add = 0, mul = 0, div = 0, sq = 0;
  if (K == 2)
    GPU_cmplx2_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), Q.p(0), Q.p(1),
                               Q.p(2), Q.p(3), R.p(0), R.p(1), R.p(2), R.p(3), &t[0], &t[1],
                               &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8], &t[9], &wall,
                               &add, &mul, &div, &sq, false);
  else if (K == 8)
    GPU_cmplx8_blocked_houseqr(
        m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), A.p(4), A.p(5), A.p(6), A.p(7),
        A.p(8), A.p(9), A.p(10), A.p(11), A.p(12), A.p(13), A.p(14), A.p(15), Q.p(0), Q.p(1),
        Q.p(2), Q.p(3), Q.p(4), Q.p(5), Q.p(6), Q.p(7), Q.p(8), Q.p(9), Q.p(10), Q.p(11),
        Q.p(12), Q.p(13), Q.p(14), Q.p(15), R.p(0), R.p(1), R.p(2), R.p(3), R.p(4), R.p(5),
        R.p(6), R.p(7), R.p(8), R.p(9), R.p(10), R.p(11), R.p(12), R.p(13), R.p(14), R.p(15),
        &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8], &t[9], &wall, &add, &mul,
        &div, &sq, false);
  else
    GPU_cmplx4_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), A.p(4), A.p(5),
                               A.p(6), A.p(7), Q.p(0), Q.p(1), Q.p(2), Q.p(3), Q.p(4), Q.p(5),
                               Q.p(6), Q.p(7), R.p(0), R.p(1), R.p(2), R.p(3), R.p(4), R.p(5),
                               R.p(6), R.p(7), &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6],
                               &t[7], &t[8], &t[9], &wall, &add, &mul, &div, &sq, false);
  if (!(wall > 0.0) || add <= 0 || mul <= 0) {
    std::printf("GPU_cmplx%d_blocked_houseqr: wall time or counters not set\n", K);
    ++failures;
  }
}

static void test_qr(int K, int m, int n, int szt) {
  Planes A(2 * K, m, n), Qc(2 * K, m, m), Rc(2 * K, m, n), Qg(2 * K, m, m), Rg(2 * K, m, n);
  A.randomize(K);
  obq_blocked_houseqr(K, m, n, szt, n / szt, A.data(), Qc.data(), Rc.data());
  qr_gpu(K, A, Qg, Rg, szt);
  const double tol = 64.0 * m * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_cmplx%d_blocked_houseqr %dx%d szt %d: Q", K, m, n, szt);
  check(Qg.diff(K, Qc) < tol, what, Qg.diff(K, Qc), tol);
  std::snprintf(what, sizeof what, "GPU_cmplx%d_blocked_houseqr %dx%d szt %d: R", K, m, n, szt);
  check(Rg.diff(K, Rc) < tol * std::sqrt((double)m), what, Rg.diff(K, Rc),
        tol * std::sqrt((double)m));
}

// the real QR through GPU_dbl{2,4,8}_blocked_houseqr against the oracle's
// restatement of CPU_dbl{2,4,8}_blocked_houseqr
static void test_real_qr(int K, int m, int n, int szt) {
  Planes A(K, m, n), Qc(K, m, m), Rc(K, m, n), Qg(K, m, m), Rg(K, m, n);
  A.randomize(K);
  obq_real_blocked_houseqr(K, m, n, szt, n / szt, A.data(), Qc.data(), Rc.data());
  double t[10] = {0}, wall = 0.0;
  long long add = 0, mul = 0, div = 0, sq = 0;
  if (K == 2)
    GPU_dbl2_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), Qg.p(0), Qg.p(1), Rg.p(0),
                             Rg.p(1), &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6], &t[7],
                             &t[8], &t[9], &wall, &add, &mul, &div, &sq, false);
  else if (K == 4)
    GPU_dbl4_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), Qg.p(0),
                             Qg.p(1), Qg.p(2), Qg.p(3), Rg.p(0), Rg.p(1), Rg.p(2), Rg.p(3),
                             &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8],
                             &t[9], &wall, &add, &mul, &div, &sq, false);
  else
    GPU_dbl8_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), A.p(4), A.p(5),
                             A.p(6), A.p(7), Qg.p(0), Qg.p(1), Qg.p(2), Qg.p(3), Qg.p(4),
                             Qg.p(5), Qg.p(6), Qg.p(7), Rg.p(0), Rg.p(1), Rg.p(2), Rg.p(3),
                             Rg.p(4), Rg.p(5), Rg.p(6), Rg.p(7), &t[0], &t[1], &t[2], &t[3],
                             &t[4], &t[5], &t[6], &t[7], &t[8], &t[9], &wall, &add, &mul, &div,
                             &sq, false);
  if (!(wall > 0.0) || add <= 0 || mul <= 0) {
    std::printf("GPU_dbl%d_blocked_houseqr: wall time or counters not set\n", K);
    ++failures;
  }
  const double tol = 8.0 * m * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%d_blocked_houseqr %dx%d szt %d: Q", K, m, n, szt);
  check(Qg.diff(K, Qc) < tol, what, Qg.diff(K, Qc), tol);
  std::snprintf(what, sizeof what, "GPU_dbl%d_blocked_houseqr %dx%d szt %d: R", K, m, n, szt);
  check(Rg.diff(K, Rc) < tol * std::sqrt((double)m), what, Rg.diff(K, Rc),
        tol * std::sqrt((double)m));
}

static void solve_gpu(int K, Planes &U, Planes &b, Planes &x, int szt) {
  const int n = U.rows;
  double t[4] = {0}, wall = 0.0, over[3] = {0};
  long long add = 0, mul = 0, div = 0;
  double *bp[16], *xp[16];
  for (int q = 0; q < 2 * K; ++q) { bp[q] = b.row(q, 0); xp[q] = x.row(q, 0); }
  if (K == 8)
    GPU_cmplx8_upper_tiled_solver(
        n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), U.p(4), U.p(5), U.p(6), U.p(7), U.p(8),
        U.p(9), U.p(10), U.p(11), U.p(12), U.p(13), U.p(14), U.p(15), bp[0], bp[1], bp[2], bp[3],
        bp[4], bp[5], bp[6], bp[7], bp[8], bp[9], bp[10], bp[11], bp[12], bp[13], bp[14], bp[15],
        xp[0], xp[1], xp[2], xp[3], xp[4], xp[5], xp[6], xp[7], xp[8], xp[9], xp[10], xp[11],
        xp[12], xp[13], xp[14], xp[15], &t[0], &t[1], &t[2], &t[3], &wall, &add, &mul, &div);
  else if (K == 2)
    GPU_cmplx2_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), bp[0], bp[1],
                                  bp[2], bp[3], xp[0], xp[1], xp[2], xp[3], &t[0], &t[1], &t[2],
                                  &t[3], &wall, &add, &mul, &div);
  else
    GPU_cmplx4_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), U.p(4), U.p(5),
                                  U.p(6), U.p(7), bp[0], bp[1], bp[2], bp[3], bp[4], bp[5], bp[6],
                                  bp[7], xp[0], xp[1], xp[2], xp[3], xp[4], xp[5], xp[6], xp[7],
                                  &t[0], &t[1], &t[2], &t[3], &wall, &add, &over[0], &mul,
                                  &over[1], &div, &over[2]);
}

static void qhb_gpu(int K, Planes &Q, Planes &b) {
  const int n = Q.rows;
  double ms = 0.0;
  double *bp[8];
  for (int q = 0; q < 2 * K; ++q) bp[q] = b.row(q, 0);
  if (K == 2)
    GPU_cmplx2_bals_qhb(n, 1, n, Q.p(0), Q.p(1), Q.p(2), Q.p(3), bp[0], bp[1], bp[2], bp[3], &ms,
                        0);
  else
    GPU_cmplx4_bals_qhb(n, 1, n, Q.p(0), Q.p(1), Q.p(2), Q.p(3), Q.p(4), Q.p(5), Q.p(6), Q.p(7),
                        bp[0], bp[1], bp[2], bp[3], bp[4], bp[5], bp[6], bp[7], &ms, 0);
}

// an upper triangular system, diagonally dominant enough; P = 2K planes
// (complex) or K (real)
static void upper_system(int K, Planes &U, Planes &b) {
  U.randomize(K);
  b.randomize(K);
  const double zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int g = 0; g < U.P / K; ++g)
    for (int i = 0; i < U.rows; ++i)
      for (int j = 0; j < i; ++j) U.set(K, g * K, i, j, zero);
  for (int i = 0; i < U.rows; ++i) U.v[(size_t)i * U.cols + i] += 4.0;
}

static void test_tiled(int K, int n, int szt) {
  Planes U(2 * K, n, n), b(2 * K, 1, n), xg(2 * K, 1, n);
  upper_system(K, U, b);
  Planes Uc(U), bc(b), xc(2 * K, 1, n);
  obq_upper_tiled_solver(K, n, szt, n / szt, Uc.data(), bc.data(), xc.data());
  const Planes U0(U), b0(b);
  solve_gpu(K, U, b, xg, szt);
  const double tol = 64.0 * n * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_cmplx%d_upper_tiled_solver %d szt %d: x", K, n, szt);
  check(xg.diff(K, xc) < tol, what, xg.diff(K, xc), tol);
  std::snprintf(what, sizeof what, "GPU_cmplx%d_upper_tiled_solver: U, b kept", K);
  check(U.same(U0) && b.same(b0), what, std::fmax(U.diff(K, U0), b.diff(K, b0)), 0.0);
}

// b = A x0; Q R = A (GPU); y = Q^H b (GPU); R x = y (GPU): x = x0
static void test_least_squares(int K, int n, int szt) {
  Planes A(2 * K, n, n), x0(2 * K, 1, n), b(2 * K, 1, n), Q(2 * K, n, n), R(2 * K, n, n),
      x(2 * K, 1, n);
  A.randomize(K);
  x0.randomize(K);
  obq_matvec(K, n, n, A.data(), x0.data(), b.data());
  qr_gpu(K, A, Q, R, szt);
  qhb_gpu(K, Q, b);
  solve_gpu(K, R, b, x, szt);
  const double tol = 1e6 * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "least squares K=%d n=%d: x - x0", K, n);
  check(x.diff(K, x0) < tol, what, x.diff(K, x0), tol);
}

// a real upper triangular system through GPU_dblK_upper_tiled_solver
// against the complex oracle solver with zero imaginary parts
static void test_tiled_real(int K, int n, int szt) {
  Planes U(K, n, n), b(K, 1, n), xg(K, 1, n);
  upper_system(K, U, b);
  Planes Uc(2 * K, n, n), bc(2 * K, 1, n), xc(2 * K, 1, n);
  std::copy(U.v.begin(), U.v.end(), Uc.v.begin());
  std::copy(b.v.begin(), b.v.end(), bc.v.begin());
  obq_upper_tiled_solver(K, n, szt, n / szt, Uc.data(), bc.data(), xc.data());
  const Planes U0(U), b0(b);
  double t[4] = {0, 0, 0, 0}, over[3] = {0, 0, 0};
  long long cnt[3] = {0, 0, 0};
  if (K == 2) {
    GPU_dbl2_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), b.row(0, 0), b.row(1, 0),
                                xg.row(0, 0), xg.row(1, 0), &t[0], &t[1], &t[2], &t[3], &t[0],
                                &cnt[0], &cnt[1], &cnt[2]);
  } else {
    GPU_dbl4_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), b.row(0, 0),
                                b.row(1, 0), b.row(2, 0), b.row(3, 0), xg.row(0, 0),
                                xg.row(1, 0), xg.row(2, 0), xg.row(3, 0), &t[0], &t[1], &t[2],
                                &t[3], &t[0], &cnt[0], &over[0], &cnt[1], &over[1], &cnt[2],
                                &over[2]);
  }
  Planes xr(K, 1, n);
  std::copy(xc.v.begin(), xc.v.begin() + (size_t)K * n, xr.v.begin());
  const double tol = 64.0 * n * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%d_upper_tiled_solver %d szt %d: x", K, n, szt);
  check(xg.diff(K, xr) < tol, what, xg.diff(K, xr), tol);
  std::snprintf(what, sizeof what, "GPU_dbl%d_upper_tiled_solver: U, b kept", K);
  check(U.same(U0) && b.same(b0), what, U.same(U0) && b.same(b0) ? 0.0 : 1.0, 0.0);
}

// ------------------------------------------------------------ bals solve

// the matrix series [K][degp1][dim][dim] with row pointers mp[q][j][i]
struct MatSeries {
  int K, degp1, dim;
  Planes all;
  std::vector<std::vector<double **>> mp;
  std::vector<std::vector<std::vector<double *>>> rows;
  MatSeries(int K_, int d, int n)
      : K(K_), degp1(d), dim(n), all(K_, d * n, n), mp(K_, std::vector<double **>(d)),
        rows(K_, std::vector<std::vector<double *>>(d, std::vector<double *>(n))) {
    for (int q = 0; q < K; ++q)
      for (int j = 0; j < d; ++j) {
        for (int i = 0; i < n; ++i) rows[q][j][i] = all.row(q, j * n + i);
        mp[q][j] = rows[q][j].data();
      }
  }
};

// the same block Toeplitz system through onw_qrbs_solve (CPU_dblK_qrbs_solve
// restated) and GPU_dblK_bals_solve: every output bit for bit
static void test_bals_solve(int K, int dim, int degp1, int tailidx) {
  MatSeries mat(K, degp1, dim);
  mat.all.randomize(K);
  Planes rc(K, degp1, dim), xc(K, degp1, dim), Qc(K, dim, dim), Rc(K, dim, dim);
  rc.randomize(K);
  Planes rg(rc), xg(K, degp1, dim), Qg(K, dim, dim), Rg(K, dim, dim);
  int zc = 1, nc = 0, uc = 0, bc = 0, tc = 0;
  onw_qrbs_solve(K, dim, degp1, tailidx, mat.all.data(), rc.data(), xc.data(), Qc.data(),
                 Rc.data(), &zc, &nc, &uc, &bc, &tc);
  bool zg = true, ng = false;
  int ug = 0, bg = 0, tg = 0;
  double t[4] = {0, 0, 0, 0};
  if (K == 1) {
    GPU_dbl_bals_solve(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), Qg.p(0), Rg.p(0), rg.p(0),
                       xg.p(0), &zg, &ng, &ug, &bg, &tg, &t[0], &t[1], &t[2], &t[3], 0);
  } else if (K == 2) {
    GPU_dbl2_bals_solve(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), mat.mp[1].data(), Qg.p(0),
                        Qg.p(1), Rg.p(0), Rg.p(1), rg.p(0), rg.p(1), xg.p(0), xg.p(1), &zg, &ng,
                        &ug, &bg, &tg, &t[0], &t[1], &t[2], &t[3], 0);
  } else {
    GPU_dbl4_bals_solve(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), mat.mp[1].data(),
                        mat.mp[2].data(), mat.mp[3].data(), Qg.p(0), Qg.p(1), Qg.p(2), Qg.p(3),
                        Rg.p(0), Rg.p(1), Rg.p(2), Rg.p(3), rg.p(0), rg.p(1), rg.p(2), rg.p(3),
                        xg.p(0), xg.p(1), xg.p(2), xg.p(3), &zg, &ng, &ug, &bg, &tg, &t[0], &t[1],
                        &t[2], &t[3], 0);
  }
  const bool same = xc.same(xg) && rc.same(rg) && Qc.same(Qg) && Rc.same(Rg) &&
                    (zc != 0) == zg && (nc != 0) == ng && uc == ug && bc == bg && tc == tg;
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%s_bals_solve dim %d degp1 %d (bitwise)",
                K == 1 ? "" : (K == 2 ? "2" : "4"), dim, degp1);
  check(same, what, same ? 0.0 : 1.0, 0.0);
}

// the residual of a block Toeplitz system through onw_linear_residue
// (CPU_dblK_linear_residue restated) and GPU_dblK_linear_residue: bit for
// bit, and the rows below tailidx left as they were
static void test_linear_residue(int K, int dim, int degp1, int tailidx) {
  MatSeries mat(K, degp1, dim);
  mat.all.randomize(K);
  Planes b(K, degp1, dim), x(K, degp1, dim), rc(K, degp1, dim);
  b.randomize(K);
  x.randomize(K);
  rc.randomize(K);
  Planes rg(rc);
  double mc[4] = {0, 0, 0, 0}, mg[4] = {0, 0, 0, 0}, ms = 0.0;
  onw_linear_residue(K, dim, degp1, tailidx, mat.all.data(), b.data(), x.data(), rc.data(), mc);
  long long add = -1, mul = -1;
  if (K == 1) {
    GPU_dbl_linear_residue(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), b.p(0), x.p(0),
                           rg.p(0), &mg[0], &ms, &add, &mul, 0);
  } else if (K == 2) {
    GPU_dbl2_linear_residue(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), mat.mp[1].data(),
                            b.p(0), b.p(1), x.p(0), x.p(1), rg.p(0), rg.p(1), &mg[0], &mg[1], &ms,
                            &add, &mul, 0);
  } else {
    GPU_dbl4_linear_residue(dim, degp1, 1, dim, tailidx, mat.mp[0].data(), mat.mp[1].data(),
                            mat.mp[2].data(), mat.mp[3].data(), b.p(0), b.p(1), b.p(2), b.p(3),
                            x.p(0), x.p(1), x.p(2), x.p(3), rg.p(0), rg.p(1), rg.p(2), rg.p(3),
                            &mg[0], &mg[1], &mg[2], &mg[3], &ms, &add, &mul, 0);
  }
  long long prods = 0;
  for (int i = tailidx; i < degp1; ++i) prods += i - tailidx + 1;
  bool same = rc.same(rg) && add == prods * dim * dim && mul == add;
  for (int q = 0; q < K; ++q) same = same && mc[q] == mg[q];
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%s_linear_residue dim %d degp1 %d tail %d (bitwise)",
                K == 1 ? "" : (K == 2 ? "2" : "4"), dim, degp1, tailidx);
  check(same, what, same ? 0.0 : 1.0, 0.0);
}

int main() {
  for (int K : {1, 2, 4, 8}) {
    test_products(K, 15);
    test_products(K, 8);
  }
  // the evaluation shapes of the reference's polynomial testers: random
  // supports, degree 15 series; all real and complex precisions
  for (int K : {1, 2, 4, 8}) test_poly_evaldiff(K, 0, 8, 8, 3, 15);
  for (int K : {1, 2, 4}) test_poly_evaldiff(K, 1, 8, 8, 3, 15);
  test_poly_evaldiff(2, 0, 12, 40, 6, 8);
  test_poly_evaldiff(4, 1, 6, 17, 6, 4);
  test_qr(2, 64, 64, 16);
  test_qr(4, 64, 64, 16);
  test_qr(2, 96, 64, 32);
  test_qr(4, 48, 32, 8);
  test_qr(8, 32, 32, 8);
  test_real_qr(2, 64, 64, 16);
  test_real_qr(4, 96, 64, 32);
  test_real_qr(8, 32, 32, 8);
  test_tiled(2, 64, 16);
  test_tiled(4, 96, 32);
  test_tiled(8, 32, 8);
  test_tiled_real(2, 64, 16);
  test_tiled_real(4, 96, 32);
  test_least_squares(2, 64, 16);
  test_least_squares(4, 64, 32);
  test_bals_solve(1, 8, 8, 1);
  test_bals_solve(2, 8, 16, 1);
  test_bals_solve(4, 8, 8, 1);
  test_bals_solve(2, 16, 5, 1);
  test_linear_residue(1, 8, 8, 0);
  test_linear_residue(2, 8, 16, 3);
  test_linear_residue(4, 12, 8, 1);
  std::printf("%d failed\n", failures);
  return failures;
}
