// Flat extern "C" wrappers around the reference's complex and real blocked
// Householder QR and tiled back substitution (host code, compiled in place
// into oracle/_ref/libphcref.so).  TEST INFRASTRUCTURE ONLY.
// Arrays are complex limb-planar [2K][rows][cols] as in oracle/baqr.h.
#include <cstddef>
#include <vector>
#include "dbl2_baqr_host.h"
#include "dbl4_baqr_host.h"
#include "dbl8_baqr_host.h"
#include "dbl2_tabs_host.h"
#include "dbl4_tabs_host.h"
#include "dbl8_tabs_host.h"

namespace {

struct Planes {
  std::vector<std::vector<double *>> rows;
  std::vector<double **> p;
  Planes(double *base, int P, int nr, int nc) : rows(P), p(P) {
    for (int q = 0; q < P; ++q) {
      for (int i = 0; i < nr; ++i) rows[q].push_back(base + ((std::size_t)q * nr + i) * nc);
      p[q] = rows[q].data();
    }
  }
};

#define X2(a) a[0], a[1]
#define X4(a) a[0], a[1], a[2], a[3]
#define X8(a) a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]
#define X16(a) X8(a), a[8], a[9], a[10], a[11], a[12], a[13], a[14], a[15]

}  // namespace

extern "C" {

int ref_cmplx_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, double *A,
                              double *Q, double *R) {
  const int P = 2 * K;
  Planes a(A, P, nrows, ncols), q(Q, P, nrows, nrows), r(R, P, nrows, ncols);
  double sec = 0.0;
  if (K == 2)
    CPU_cmplx2_blocked_houseqr(nrows, ncols, szt, nbt, X4(a.p), X4(q.p), X4(r.p), &sec, false);
  else if (K == 4)
    CPU_cmplx4_blocked_houseqr(nrows, ncols, szt, nbt, X8(a.p), X8(q.p), X8(r.p), &sec, false);
  else if (K == 8)
    CPU_cmplx8_blocked_houseqr(nrows, ncols, szt, nbt, X16(a.p), X16(q.p), X16(r.p), &sec, false);
  else
    return -1;
  return 0;
}

// real A, Q, R limb-planar [K][rows][cols]
int ref_real_blocked_houseqr(int K, int nrows, int ncols, int szt, int nbt, double *A, double *Q,
                             double *R) {
  Planes a(A, K, nrows, ncols), q(Q, K, nrows, nrows), r(R, K, nrows, ncols);
  double sec = 0.0;
  if (K == 2)
    CPU_dbl2_blocked_houseqr(nrows, ncols, szt, nbt, X2(a.p), X2(q.p), X2(r.p), &sec, false);
  else if (K == 4)
    CPU_dbl4_blocked_houseqr(nrows, ncols, szt, nbt, X4(a.p), X4(q.p), X4(r.p), &sec, false);
  else if (K == 8)
    CPU_dbl8_blocked_houseqr(nrows, ncols, szt, nbt, X8(a.p), X8(q.p), X8(r.p), &sec, false);
  else
    return -1;
  return 0;
}

int ref_cmplx_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b,
                                 double *x) {
  const int P = 2 * K;
  Planes u(U, P, dim, dim);
  double *pb[16], *px[16];
  for (int q = 0; q < P; ++q) {
    pb[q] = b + (std::size_t)q * dim;
    px[q] = x + (std::size_t)q * dim;
  }
  double sec = 0.0;
  if (K == 2)
    CPU_cmplx2_upper_tiled_solver(dim, szt, nbt, X4(u.p), X4(pb), X4(px), &sec);
  else if (K == 4)
    CPU_cmplx4_upper_tiled_solver(dim, szt, nbt, X8(u.p), X8(pb), X8(px), &sec);
  else if (K == 8)
    CPU_cmplx8_upper_tiled_solver(dim, szt, nbt, X16(u.p), X16(pb), X16(px), &sec);
  else
    return -1;
  return 0;
}

}  // extern "C"
