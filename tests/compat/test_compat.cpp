// Parity of the reference-signature entry points of libphcmd.so
// (include/phcmd_compat.h) with the reference's own host functions, compiled
// from the reference sources into oracle/_ref/libphcref.so.  Written like the
// reference's testers (Convolutions/test_dbl*_convolutions.cpp,
// Matrices/test_dbl4_baqr.cpp, test_dbl4_tabs.cpp): random input, the CPU
// function and the GPU function on the same data, the largest difference.
//
// Series products must agree bit for bit; the blocked QR, the tiled solver
// and Q^H b split long sums across lanes, so they must agree within
// 64 * dim * eps^K, and the least-squares solve must recover a planted x0.
// Exit status = number of failed checks.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/phcmd_compat.h"

// the reference's host functions (oracle/_ref/libphcref.so)
double random_double();
void ddf_sub(double, double, double, double, double *, double *);
void ddf_add(double, double, double, double, double *, double *);
void ddf_mul(double, double, double, double, double *, double *);
void qdf_sub(double, double, double, double, double, double, double, double, double *, double *,
             double *, double *);
void qdf_add(double, double, double, double, double, double, double, double, double *, double *,
             double *, double *);
void qdf_mul(double, double, double, double, double, double, double, double, double *, double *,
             double *, double *);
void CPU_dbl_product(int, double *, double *, double *);
void CPU_cmplx_product(int, double *, double *, double *, double *, double *, double *);
void CPU_dbl2_product(int, double *, double *, double *, double *, double *, double *);
void CPU_cmplx2_product(int, double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *);
void CPU_dbl4_product(int, double *, double *, double *, double *, double *, double *, double *,
                      double *, double *, double *, double *, double *);
void CPU_cmplx4_product(int, double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *);
void CPU_cmplx2_blocked_houseqr(int, int, int, int, double **, double **, double **, double **,
                                double **, double **, double **, double **, double **, double **,
                                double **, double **, double *, bool);
void CPU_cmplx4_blocked_houseqr(int, int, int, int, double **, double **, double **, double **,
                                double **, double **, double **, double **, double **, double **,
                                double **, double **, double **, double **, double **, double **,
                                double **, double **, double **, double **, double **, double **,
                                double **, double **, double *, bool);
void CPU_cmplx2_upper_tiled_solver(int, int, int, double **, double **, double **, double **,
                                   double *, double *, double *, double *, double *, double *,
                                   double *, double *, double *);
void CPU_cmplx4_upper_tiled_solver(int, int, int, double **, double **, double **, double **,
                                   double **, double **, double **, double **, double *, double *,
                                   double *, double *, double *, double *, double *, double *,
                                   double *, double *, double *, double *, double *, double *,
                                   double *, double *, double *);

void CPU_dbl_qrbs_solve(int, int, int, double ***, double **, double **, double **, double **,
                        double *, bool *, bool *, int *, int *, int *, int);
void CPU_dbl2_qrbs_solve(int, int, int, double ***, double ***, double **, double **, double **,
                         double **, double **, double **, double **, double **, double *,
                         double *, bool *, bool *, int *, int *, int *, int);
void CPU_dbl4_qrbs_solve(int, int, int, double ***, double ***, double ***, double ***, double **,
                         double **, double **, double **, double **, double **, double **,
                         double **, double **, double **, double **, double **, double **,
                         double **, double **, double **, double *, double *, double *, double *,
                         bool *, bool *, int *, int *, int *, int);

void CPU_dbl_linear_residue(int, int, int, double ***, double **, double **, double **, double *,
                            int);
void CPU_dbl2_linear_residue(int, int, int, double ***, double ***, double **, double **,
                             double **, double **, double **, double **, double *, double *, int);
void CPU_dbl4_linear_residue(int, int, int, double ***, double ***, double ***, double ***,
                             double **, double **, double **, double **, double **, double **,
                             double **, double **, double **, double **, double **, double **,
                             double *, double *, double *, double *, int);

void CPU_dbl8_product(int, double *, double *, double *, double *, double *, double *, double *,
                      double *, double *, double *, double *, double *, double *, double *,
                      double *, double *, double *, double *, double *, double *, double *,
                      double *, double *, double *);
void CPU_cmplx8_product(int, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *,
                        double *, double *, double *, double *, double *, double *, double *);

void CPU_dbl2_upper_tiled_solver(int, int, int, double **, double **, double *, double *,
                                 double *, double *, double *);
void CPU_dbl4_upper_tiled_solver(int, int, int, double **, double **, double **, double **,
                                 double *, double *, double *, double *, double *, double *,
                                 double *, double *, double *);

static int failures = 0;

static void check(bool ok, const char *what, double err, double tol) {
  std::printf("%-44s max diff %.3e  tol %.3e  %s\n", what, err, tol, ok ? "ok" : "FAILED");
  if (!ok) ++failures;
}

// a K-limb number (K = 1, 2, 4) with random limbs below the leading one
static void random_md(int K, double *x) {
  double a[4] = {random_double(), 0.0, 0.0, 0.0};
  if (K == 2) {
    ddf_add(a[0], 0.0, random_double() * 0x1p-53, 0.0, &x[0], &x[1]);
  } else if (K == 4) {
    double t[4];
    qdf_add(a[0], 0.0, 0.0, 0.0, random_double() * 0x1p-53, random_double() * 0x1p-106,
            random_double() * 0x1p-159, 0.0, &t[0], &t[1], &t[2], &t[3]);
    for (int i = 0; i < 4; ++i) x[i] = t[i];
  } else {
    x[0] = a[0];
  }
}

// |leading limb of a - b| in K-limb arithmetic
static double md_diff(int K, const double *a, const double *b) {
  double c[4];
  if (K == 1) return std::fabs(a[0] - b[0]);
  if (K == 2) {
    ddf_sub(a[0], a[1], b[0], b[1], &c[0], &c[1]);
    return std::fabs(c[0]);
  }
  qdf_sub(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], &c[0], &c[1], &c[2], &c[3]);
  return std::fabs(c[0]);
}

// complex K-limb matrix, one row-pointer array per plane (K re, then K im)
struct CMat {
  int K, rows, cols;
  std::vector<std::vector<double>> store;
  std::vector<std::vector<double *>> ptr;
  CMat(int K_, int r, int c) : K(K_), rows(r), cols(c), store(2 * K_), ptr(2 * K_) {
    for (int q = 0; q < 2 * K; ++q) store[q].assign((size_t)r * c, 0.0);
    rows_of();
  }
  // a copy owns its numbers: its row pointers point into its own storage
  CMat(const CMat &o) : K(o.K), rows(o.rows), cols(o.cols), store(o.store), ptr(2 * o.K) {
    rows_of();
  }
  CMat &operator=(const CMat &) = delete;
  void rows_of() {
    for (int q = 0; q < 2 * K; ++q) {
      ptr[q].resize(rows);
      for (int i = 0; i < rows; ++i) ptr[q][i] = &store[q][(size_t)i * cols];
    }
  }
  double **p(int q) { return ptr[q].data(); }
  void get(int i, int j, double *re, double *im) const {
    for (int l = 0; l < K; ++l) {
      re[l] = store[l][(size_t)i * cols + j];
      im[l] = store[K + l][(size_t)i * cols + j];
    }
  }
  void set(int i, int j, const double *re, const double *im) {
    for (int l = 0; l < K; ++l) {
      store[l][(size_t)i * cols + j] = re[l];
      store[K + l][(size_t)i * cols + j] = im[l];
    }
  }
  void randomize() {
    double re[4], im[4];
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) {
        random_md(K, re);
        random_md(K, im);
        set(i, j, re, im);
      }
  }
  double diff(const CMat &o) const {
    double m = 0.0, a[4], b[4], c[4], d[4];
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) {
        get(i, j, a, b);
        o.get(i, j, c, d);
        m = std::fmax(m, std::fmax(md_diff(K, a, c), md_diff(K, b, d)));
      }
    return m;
  }
};

// ------------------------------------------------------------ md helpers

static void md_add(int K, const double *a, const double *b, double *c) {
  if (K == 2) ddf_add(a[0], a[1], b[0], b[1], &c[0], &c[1]);
  else qdf_add(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], &c[0], &c[1], &c[2], &c[3]);
}

static void md_mul(int K, const double *a, const double *b, double *c) {
  if (K == 2) ddf_mul(a[0], a[1], b[0], b[1], &c[0], &c[1]);
  else qdf_mul(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3], &c[0], &c[1], &c[2], &c[3]);
}

static void md_neg(int K, double *a) {
  for (int l = 0; l < K; ++l) a[l] = -a[l];
}

// b = A x (complex, K limbs), x and b as n x 1 matrices
static void matvec(const CMat &A, const CMat &x, CMat &b) {
  const int K = A.K;
  for (int i = 0; i < A.rows; ++i) {
    double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
    for (int j = 0; j < A.cols; ++j) {
      double ar[4], ai[4], xr[4], xi[4], t[4];
      A.get(i, j, ar, ai);
      x.get(j, 0, xr, xi);
      md_mul(K, ar, xr, t); md_add(K, sr, t, sr);
      md_mul(K, ai, xi, t); md_neg(K, t); md_add(K, sr, t, sr);
      md_mul(K, ar, xi, t); md_add(K, si, t, si);
      md_mul(K, ai, xr, t); md_add(K, si, t, si);
    }
    b.set(i, 0, sr, si);
  }
}

// ------------------------------------------------------------ products

static void test_products(int deg) {
  const int D = deg + 1;
  std::vector<double> v(64 * D);
  double *p[64];
  for (int i = 0; i < 64; ++i) p[i] = &v[(size_t)i * D];
  for (auto &e : v) e = random_double();
  // inputs x: p[0..7], y: p[8..15]; outputs CPU p[16..23], GPU p[24..31]
  auto same = [&](int planes) {
    double m = 0.0;
    for (int q = 0; q < planes; ++q)
      for (int k = 0; k < D; ++k) m = std::fmax(m, std::fabs(p[16 + q][k] - p[24 + q][k]));
    return m;
  };
  CPU_dbl_product(deg, p[0], p[8], p[16]);
  GPU_dbl_product(p[0], p[8], p[24], deg, 1, D, 0);
  check(same(1) == 0.0, "GPU_dbl_product vs CPU_dbl_product", same(1), 0.0);
  CPU_cmplx_product(deg, p[0], p[1], p[8], p[9], p[16], p[17]);
  GPU_cmplx_product(p[0], p[1], p[8], p[9], p[24], p[25], deg, 1, D, 0);
  check(same(2) == 0.0, "GPU_cmplx_product vs CPU_cmplx_product", same(2), 0.0);
  CPU_dbl2_product(deg, p[0], p[1], p[8], p[9], p[16], p[17]);
  GPU_dbl2_product(p[0], p[1], p[8], p[9], p[24], p[25], deg, 1, D, 1);
  check(same(2) == 0.0, "GPU_dbl2_product vs CPU_dbl2_product", same(2), 0.0);
  CPU_cmplx2_product(deg, p[0], p[1], p[2], p[3], p[8], p[9], p[10], p[11], p[16], p[17], p[18],
                     p[19]);
  GPU_cmplx2_product(p[0], p[1], p[2], p[3], p[8], p[9], p[10], p[11], p[24], p[25], p[26], p[27],
                     deg, 1, D, 0);
  check(same(4) == 0.0, "GPU_cmplx2_product vs CPU_cmplx2_product", same(4), 0.0);
  CPU_dbl4_product(deg, p[0], p[1], p[2], p[3], p[8], p[9], p[10], p[11], p[16], p[17], p[18],
                   p[19]);
  GPU_dbl4_product(p[0], p[1], p[2], p[3], p[8], p[9], p[10], p[11], p[24], p[25], p[26], p[27],
                   deg, 1, D, 1);
  check(same(4) == 0.0, "GPU_dbl4_product vs CPU_dbl4_product", same(4), 0.0);
  CPU_cmplx4_product(deg, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                     p[11], p[12], p[13], p[14], p[15], p[16], p[17], p[18], p[19], p[20], p[21],
                     p[22], p[23]);
  GPU_cmplx4_product(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                     p[12], p[13], p[14], p[15], p[24], p[25], p[26], p[27], p[28], p[29], p[30],
                     p[31], deg, 1, D, 0);
  check(same(8) == 0.0, "GPU_cmplx4_product vs CPU_cmplx4_product", same(8), 0.0);
}

// octo-double products: x in p[0..15], y in p[16..31], CPU z in p[32..47],
// GPU z in p[48..63] (real: the first 8 planes of each)
static void test_products8(int deg) {
  const int D = deg + 1;
  std::vector<double> v(64 * D);
  double *p[64];
  for (int i = 0; i < 64; ++i) p[i] = &v[(size_t)i * D];
  for (auto &e : v) e = random_double();
  auto diff = [&](int planes) {
    double m = 0.0;
    for (int q = 0; q < planes; ++q)
      for (int k = 0; k < D; ++k) m = std::fmax(m, std::fabs(p[32 + q][k] - p[48 + q][k]));
    return m;
  };
  CPU_dbl8_product(deg, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[16], p[17], p[18],
                   p[19], p[20], p[21], p[22], p[23], p[32], p[33], p[34], p[35], p[36], p[37],
                   p[38], p[39]);
  GPU_dbl8_product(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[16], p[17], p[18], p[19],
                   p[20], p[21], p[22], p[23], p[48], p[49], p[50], p[51], p[52], p[53], p[54],
                   p[55], deg, 1, D, 1);
  check(diff(8) == 0.0, "GPU_dbl8_product vs CPU_dbl8_product", diff(8), 0.0);
  CPU_cmplx8_product(deg, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                     p[11], p[12], p[13], p[14], p[15], p[16], p[17], p[18], p[19], p[20], p[21],
                     p[22], p[23], p[24], p[25], p[26], p[27], p[28], p[29], p[30], p[31], p[32],
                     p[33], p[34], p[35], p[36], p[37], p[38], p[39], p[40], p[41], p[42], p[43],
                     p[44], p[45], p[46], p[47]);
  GPU_cmplx8_product(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                     p[12], p[13], p[14], p[15], p[16], p[17], p[18], p[19], p[20], p[21], p[22],
                     p[23], p[24], p[25], p[26], p[27], p[28], p[29], p[30], p[31], p[48], p[49],
                     p[50], p[51], p[52], p[53], p[54], p[55], p[56], p[57], p[58], p[59], p[60],
                     p[61], p[62], p[63], deg, 1, D, 0);
  check(diff(16) == 0.0, "GPU_cmplx8_product vs CPU_cmplx8_product", diff(16), 0.0);
}

// ------------------------------------------------------------ QR, solver

static void qr_cpu(CMat &A, CMat &Q, CMat &R, int szt) {
  const int K = A.K, m = A.rows, n = A.cols;
  double sec = 0.0;
  if (K == 2)
    CPU_cmplx2_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), Q.p(0), Q.p(1),
                               Q.p(2), Q.p(3), R.p(0), R.p(1), R.p(2), R.p(3), &sec, false);
  else
    CPU_cmplx4_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), A.p(4), A.p(5),
                               A.p(6), A.p(7), Q.p(0), Q.p(1), Q.p(2), Q.p(3), Q.p(4), Q.p(5),
                               Q.p(6), Q.p(7), R.p(0), R.p(1), R.p(2), R.p(3), R.p(4), R.p(5),
                               R.p(6), R.p(7), &sec, false);
}

static void qr_gpu(CMat &A, CMat &Q, CMat &R, int szt) {
  const int K = A.K, m = A.rows, n = A.cols;
  double t[10] = {0}, wall = 0.0;
  long long add = 0, mul = 0, div = 0, sq = 0;
  if (K == 2)
    GPU_cmplx2_blocked_houseqr(m, n, szt, n / szt, A.p(0), A.p(1), A.p(2), A.p(3), Q.p(0), Q.p(1),
                               Q.p(2), Q.p(3), R.p(0), R.p(1), R.p(2), R.p(3), &t[0], &t[1],
                               &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8], &t[9], &wall,
                               &add, &mul, &div, &sq, false);
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
  CMat A(K, m, n), Qc(K, m, m), Rc(K, m, n), Qg(K, m, m), Rg(K, m, n);
  A.randomize();
  qr_cpu(A, Qc, Rc, szt);
  qr_gpu(A, Qg, Rg, szt);
  const double tol = 64.0 * m * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_cmplx%d_blocked_houseqr %dx%d szt %d: Q", K, m, n, szt);
  check(Qg.diff(Qc) < tol, what, Qg.diff(Qc), tol);
  std::snprintf(what, sizeof what, "GPU_cmplx%d_blocked_houseqr %dx%d szt %d: R", K, m, n, szt);
  check(Rg.diff(Rc) < tol * std::sqrt((double)m), what, Rg.diff(Rc), tol * std::sqrt((double)m));
}

static void solve_cpu(CMat U, CMat b, CMat &x, int szt) {  // U, b by value: the CPU version updates them
  const int K = U.K, n = U.rows;
  double sec = 0.0;
  std::vector<double *> bp(2 * K), xp(2 * K);
  for (int q = 0; q < 2 * K; ++q) { bp[q] = b.store[q].data(); xp[q] = x.store[q].data(); }
  if (K == 2)
    CPU_cmplx2_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), bp[0], bp[1],
                                  bp[2], bp[3], xp[0], xp[1], xp[2], xp[3], &sec);
  else
    CPU_cmplx4_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), U.p(4), U.p(5),
                                  U.p(6), U.p(7), bp[0], bp[1], bp[2], bp[3], bp[4], bp[5], bp[6],
                                  bp[7], xp[0], xp[1], xp[2], xp[3], xp[4], xp[5], xp[6], xp[7],
                                  &sec);
}

static void solve_gpu(CMat &U, CMat &b, CMat &x, int szt) {
  const int K = U.K, n = U.rows;
  double t[4] = {0}, wall = 0.0, over[3] = {0};
  long long add = 0, mul = 0, div = 0;
  std::vector<double *> bp(2 * K), xp(2 * K);
  for (int q = 0; q < 2 * K; ++q) { bp[q] = b.store[q].data(); xp[q] = x.store[q].data(); }
  if (K == 2)
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

static void qhb_gpu(CMat &Q, CMat &b) {
  const int K = Q.K, n = Q.rows;
  double ms = 0.0;
  std::vector<double *> bp(2 * K);
  for (int q = 0; q < 2 * K; ++q) bp[q] = b.store[q].data();
  if (K == 2)
    GPU_cmplx2_bals_qhb(n, 1, n, Q.p(0), Q.p(1), Q.p(2), Q.p(3), bp[0], bp[1], bp[2], bp[3], &ms,
                        0);
  else
    GPU_cmplx4_bals_qhb(n, 1, n, Q.p(0), Q.p(1), Q.p(2), Q.p(3), Q.p(4), Q.p(5), Q.p(6), Q.p(7),
                        bp[0], bp[1], bp[2], bp[3], bp[4], bp[5], bp[6], bp[7], &ms, 0);
}

static void test_tiled(int K, int n, int szt) {
  CMat U(K, n, n), b(K, n, 1), xc(K, n, 1), xg(K, n, 1);
  U.randomize();
  b.randomize();
  double re[4], im[4], zero[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) U.set(i, j, zero, zero);
  for (int i = 0; i < n; ++i) {  // diagonally dominant enough
    U.get(i, i, re, im);
    re[0] += 4.0;
    U.set(i, i, re, im);
  }
  const CMat U0 = U, b0 = b;
  solve_cpu(U, b, xc, szt);
  solve_gpu(U, b, xg, szt);
  const double tol = 64.0 * n * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_cmplx%d_upper_tiled_solver %d szt %d: x", K, n, szt);
  check(xg.diff(xc) < tol, what, xg.diff(xc), tol);
  std::snprintf(what, sizeof what, "GPU_cmplx%d_upper_tiled_solver: U, b kept", K);
  check(U.diff(U0) == 0.0 && b.diff(b0) == 0.0, what, std::fmax(U.diff(U0), b.diff(b0)), 0.0);
}

// b = A x0; Q R = A (GPU); y = Q^H b (GPU); R x = y (GPU): x = x0
static void test_least_squares(int K, int n, int szt) {
  CMat A(K, n, n), x0(K, n, 1), b(K, n, 1), Q(K, n, n), R(K, n, n), x(K, n, 1);
  A.randomize();
  x0.randomize();
  matvec(A, x0, b);
  qr_gpu(A, Q, R, szt);
  qhb_gpu(Q, b);
  solve_gpu(R, b, x, szt);
  const double tol = 1e6 * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "least squares K=%d n=%d: x - x0", K, n);
  check(x.diff(x0) < tol, what, x.diff(x0), tol);
}

// ------------------------------------------------------------ bals solve

// K limb planes of a real rows x cols matrix with row pointers
struct RMat {
  int K, rows, cols;
  std::vector<std::vector<double>> store;
  std::vector<std::vector<double *>> ptr;
  RMat(int K_, int r, int c) : K(K_), rows(r), cols(c), store(K_), ptr(K_) {
    for (int q = 0; q < K; ++q) {
      store[q].assign((size_t)r * c, 0.0);
      ptr[q].resize(r);
      for (int i = 0; i < r; ++i) ptr[q][i] = &store[q][(size_t)i * c];
    }
  }
  RMat(const RMat &) = delete;
  double **p(int q) { return ptr[q].data(); }
  void randomize() {
    double v[4];
    for (int i = 0; i < rows * cols; ++i) {
      random_md(K, v);
      for (int q = 0; q < K; ++q) store[q][i] = v[q];
    }
  }
  bool same(const RMat &o) const { return store == o.store; }
};

// the same block Toeplitz system through CPU_dblK_qrbs_solve (reference) and
// GPU_dblK_bals_solve (this library): every output bit for bit
static void test_bals_solve(int K, int dim, int degp1, int tailidx) {
  std::vector<RMat *> mat;
  for (int j = 0; j < degp1; ++j) {
    mat.push_back(new RMat(K, dim, dim));
    mat.back()->randomize();
  }
  std::vector<std::vector<double **>> mp(K, std::vector<double **>(degp1));
  for (int q = 0; q < K; ++q)
    for (int j = 0; j < degp1; ++j) mp[q][j] = mat[j]->p(q);
  RMat rc(K, degp1, dim), rg(K, degp1, dim), xc(K, degp1, dim), xg(K, degp1, dim);
  RMat Qc(K, dim, dim), Qg(K, dim, dim), Rc(K, dim, dim), Rg(K, dim, dim);
  rc.randomize();
  rg.store = rc.store;
  std::vector<std::vector<double>> wrk(K, std::vector<double>(dim));
  bool zc = true, nc = false, zg = true, ng = false;
  int uc = 0, bc = 0, tc = 0, ug = 0, bg = 0, tg = 0;
  double t[4] = {0, 0, 0, 0};
  if (K == 1) {
    CPU_dbl_qrbs_solve(dim, degp1, tailidx, mp[0].data(), rc.p(0), xc.p(0), Qc.p(0), Rc.p(0),
                       wrk[0].data(), &zc, &nc, &uc, &bc, &tc, 0);
    GPU_dbl_bals_solve(dim, degp1, 1, dim, tailidx, mp[0].data(), Qg.p(0), Rg.p(0), rg.p(0),
                       xg.p(0), &zg, &ng, &ug, &bg, &tg, &t[0], &t[1], &t[2], &t[3], 0);
  } else if (K == 2) {
    CPU_dbl2_qrbs_solve(dim, degp1, tailidx, mp[0].data(), mp[1].data(), rc.p(0), rc.p(1),
                        xc.p(0), xc.p(1), Qc.p(0), Qc.p(1), Rc.p(0), Rc.p(1), wrk[0].data(),
                        wrk[1].data(), &zc, &nc, &uc, &bc, &tc, 0);
    GPU_dbl2_bals_solve(dim, degp1, 1, dim, tailidx, mp[0].data(), mp[1].data(), Qg.p(0),
                        Qg.p(1), Rg.p(0), Rg.p(1), rg.p(0), rg.p(1), xg.p(0), xg.p(1), &zg, &ng,
                        &ug, &bg, &tg, &t[0], &t[1], &t[2], &t[3], 0);
  } else {
    CPU_dbl4_qrbs_solve(dim, degp1, tailidx, mp[0].data(), mp[1].data(), mp[2].data(),
                        mp[3].data(), rc.p(0), rc.p(1), rc.p(2), rc.p(3), xc.p(0), xc.p(1),
                        xc.p(2), xc.p(3), Qc.p(0), Qc.p(1), Qc.p(2), Qc.p(3), Rc.p(0), Rc.p(1),
                        Rc.p(2), Rc.p(3), wrk[0].data(), wrk[1].data(), wrk[2].data(),
                        wrk[3].data(), &zc, &nc, &uc, &bc, &tc, 0);
    GPU_dbl4_bals_solve(dim, degp1, 1, dim, tailidx, mp[0].data(), mp[1].data(), mp[2].data(),
                        mp[3].data(), Qg.p(0), Qg.p(1), Qg.p(2), Qg.p(3), Rg.p(0), Rg.p(1),
                        Rg.p(2), Rg.p(3), rg.p(0), rg.p(1), rg.p(2), rg.p(3), xg.p(0), xg.p(1),
                        xg.p(2), xg.p(3), &zg, &ng, &ug, &bg, &tg, &t[0], &t[1], &t[2], &t[3], 0);
  }
  const bool same = xc.same(xg) && rc.same(rg) && Qc.same(Qg) && Rc.same(Rg) && zc == zg &&
                    nc == ng && uc == ug && bc == bg && tc == tg;
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%s_bals_solve dim %d degp1 %d (bitwise)",
                K == 1 ? "" : (K == 2 ? "2" : "4"), dim, degp1);
  check(same, what, same ? 0.0 : 1.0, 0.0);
  for (RMat *m : mat) delete m;
}

// the residual of a block Toeplitz system through CPU_dblK_linear_residue
// (reference) and GPU_dblK_linear_residue (this library): bit for bit, and
// the rows below tailidx left as they were
static void test_linear_residue(int K, int dim, int degp1, int tailidx) {
  std::vector<RMat *> mat;
  for (int j = 0; j < degp1; ++j) {
    mat.push_back(new RMat(K, dim, dim));
    mat.back()->randomize();
  }
  std::vector<std::vector<double **>> mp(K, std::vector<double **>(degp1));
  for (int q = 0; q < K; ++q)
    for (int j = 0; j < degp1; ++j) mp[q][j] = mat[j]->p(q);
  RMat b(K, degp1, dim), x(K, degp1, dim), rc(K, degp1, dim), rg(K, degp1, dim);
  b.randomize();
  x.randomize();
  rc.randomize();
  rg.store = rc.store;
  double mc[4] = {0, 0, 0, 0}, mg[4] = {0, 0, 0, 0}, ms = 0.0;
  long long add = -1, mul = -1;
  if (K == 1) {
    CPU_dbl_linear_residue(dim, degp1, tailidx, mp[0].data(), b.p(0), x.p(0), rc.p(0), &mc[0], 0);
    GPU_dbl_linear_residue(dim, degp1, 1, dim, tailidx, mp[0].data(), b.p(0), x.p(0), rg.p(0),
                           &mg[0], &ms, &add, &mul, 0);
  } else if (K == 2) {
    CPU_dbl2_linear_residue(dim, degp1, tailidx, mp[0].data(), mp[1].data(), b.p(0), b.p(1),
                            x.p(0), x.p(1), rc.p(0), rc.p(1), &mc[0], &mc[1], 0);
    GPU_dbl2_linear_residue(dim, degp1, 1, dim, tailidx, mp[0].data(), mp[1].data(), b.p(0),
                            b.p(1), x.p(0), x.p(1), rg.p(0), rg.p(1), &mg[0], &mg[1], &ms, &add,
                            &mul, 0);
  } else {
    CPU_dbl4_linear_residue(dim, degp1, tailidx, mp[0].data(), mp[1].data(), mp[2].data(),
                            mp[3].data(), b.p(0), b.p(1), b.p(2), b.p(3), x.p(0), x.p(1), x.p(2),
                            x.p(3), rc.p(0), rc.p(1), rc.p(2), rc.p(3), &mc[0], &mc[1], &mc[2],
                            &mc[3], 0);
    GPU_dbl4_linear_residue(dim, degp1, 1, dim, tailidx, mp[0].data(), mp[1].data(),
                            mp[2].data(), mp[3].data(), b.p(0), b.p(1), b.p(2), b.p(3), x.p(0),
                            x.p(1), x.p(2), x.p(3), rg.p(0), rg.p(1), rg.p(2), rg.p(3), &mg[0],
                            &mg[1], &mg[2], &mg[3], &ms, &add, &mul, 0);
  }
  long long prods = 0;
  for (int i = tailidx; i < degp1; ++i) prods += i - tailidx + 1;
  bool same = rc.same(rg) && add == prods * dim * dim && mul == add;
  for (int q = 0; q < K; ++q) same = same && mc[q] == mg[q];
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%s_linear_residue dim %d degp1 %d tail %d (bitwise)",
                K == 1 ? "" : (K == 2 ? "2" : "4"), dim, degp1, tailidx);
  check(same, what, same ? 0.0 : 1.0, 0.0);
  for (RMat *m : mat) delete m;
}

// a real upper triangular system through CPU_dblK_upper_tiled_solver
// (reference, on copies: it overwrites U and b) and GPU_dblK_upper_tiled_solver
static void test_tiled_real(int K, int n, int szt) {
  RMat U(K, n, n), Uc(K, n, n), b(K, 1, n), bc(K, 1, n), xc(K, 1, n), xg(K, 1, n);
  U.randomize();
  b.randomize();
  for (int q = 0; q < K; ++q)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) U.p(q)[i][j] = 0.0;
  for (int i = 0; i < n; ++i) U.p(0)[i][i] += 4.0;  // diagonally dominant enough
  Uc.store = U.store;
  bc.store = b.store;
  const std::vector<std::vector<double>> U0 = U.store, b0 = b.store;
  double t[4] = {0, 0, 0, 0};
  long long cnt[3] = {0, 0, 0};
  double over[3] = {0, 0, 0};
  if (K == 2) {
    CPU_dbl2_upper_tiled_solver(n, szt, n / szt, Uc.p(0), Uc.p(1), bc.p(0)[0], bc.p(1)[0],
                                xc.p(0)[0], xc.p(1)[0], &t[0]);
    GPU_dbl2_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), b.p(0)[0], b.p(1)[0],
                                xg.p(0)[0], xg.p(1)[0], &t[0], &t[1], &t[2], &t[3], &t[0],
                                &cnt[0], &cnt[1], &cnt[2]);
  } else {
    CPU_dbl4_upper_tiled_solver(n, szt, n / szt, Uc.p(0), Uc.p(1), Uc.p(2), Uc.p(3), bc.p(0)[0],
                                bc.p(1)[0], bc.p(2)[0], bc.p(3)[0], xc.p(0)[0], xc.p(1)[0],
                                xc.p(2)[0], xc.p(3)[0], &t[0]);
    GPU_dbl4_upper_tiled_solver(n, szt, n / szt, U.p(0), U.p(1), U.p(2), U.p(3), b.p(0)[0],
                                b.p(1)[0], b.p(2)[0], b.p(3)[0], xg.p(0)[0], xg.p(1)[0],
                                xg.p(2)[0], xg.p(3)[0], &t[0], &t[1], &t[2], &t[3], &t[0],
                                &cnt[0], &over[0], &cnt[1], &over[1], &cnt[2], &over[2]);
  }
  double err = 0.0;
  for (int i = 0; i < n; ++i) {
    double a[4], c[4];
    for (int q = 0; q < K; ++q) { a[q] = xg.p(q)[0][i]; c[q] = xc.p(q)[0][i]; }
    err = std::fmax(err, md_diff(K, a, c));
  }
  const double tol = 64.0 * n * std::pow(0x1p-52, K);
  char what[96];
  std::snprintf(what, sizeof what, "GPU_dbl%d_upper_tiled_solver %d szt %d: x", K, n, szt);
  check(err < tol, what, err, tol);
  std::snprintf(what, sizeof what, "GPU_dbl%d_upper_tiled_solver: U, b kept", K);
  const bool kept = U.store == U0 && b.store == b0;
  check(kept, what, kept ? 0.0 : 1.0, 0.0);
}

int main() {
  std::srand(20261202);
  test_products(15);
  test_products(8);
  test_products8(15);
  test_products8(8);
  test_qr(2, 64, 64, 16);
  test_qr(4, 64, 64, 16);
  test_qr(2, 96, 64, 32);
  test_qr(4, 48, 32, 8);
  test_tiled(2, 64, 16);
  test_tiled(4, 96, 32);
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
