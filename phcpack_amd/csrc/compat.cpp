// The reference's C++ host entry points (declared in include/phcmd_compat.h
// with the reference's signatures) on top of the C-ABI of libphcmd.so.
//
// Each entry point packs the caller's per-limb host arrays into the
// limb-planar layout of the C-ABI ([P][rows][cols], P = K real planes then
// K imaginary planes), copies them to device buffers of its own, calls the
// C-ABI on the default stream, and unpacks the results, as the reference's
// GPU_* functions do with their own layouts.
#include <chrono>
#include <cstdio>
#include <vector>

#include "../../include/phcmd.h"
#include "../../include/phcmd_compat.h"

namespace {

typedef std::chrono::steady_clock Clock;

bool ok(int rc, const char *fn) {
  if (rc == 0) return true;
  std::fprintf(stderr, "%s: %s\n", fn, phc_last_error());
  return false;
}

// a device buffer of doubles, freed on scope exit
struct Buf {
  void *p = nullptr;
  size_t n = 0;
  explicit Buf(size_t count) : n(count) {
    if (phc_malloc(&p, sizeof(double) * (count > 0 ? count : 1)) != 0) p = nullptr;
  }
  ~Buf() {
    if (p) phc_free(p);
  }
  double *d() const { return static_cast<double *>(p); }
  bool up(const std::vector<double> &h) { return phc_memcpy_htod(p, h.data(), sizeof(double) * n) == 0; }
  bool down(std::vector<double> &h) {
    h.resize(n);
    return phc_memcpy_dtoh(h.data(), p, sizeof(double) * n) == 0;
  }
};

// planes[q] (rows x cols, row pointers) -> h[q][rows][cols]
void pack(double **const *planes, int P, int rows, int cols, std::vector<double> &h) {
  h.resize((size_t)P * rows * cols);
  for (int q = 0; q < P; ++q)
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) h[((size_t)q * rows + i) * cols + j] = planes[q][i][j];
}

void unpack(const std::vector<double> &h, int P, int rows, int cols, double **const *planes) {
  for (int q = 0; q < P; ++q)
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j) planes[q][i][j] = h[((size_t)q * rows + i) * cols + j];
}

void pack_vec(double *const *v, int P, int n, std::vector<double> &h) {
  h.resize((size_t)P * n);
  for (int q = 0; q < P; ++q)
    for (int i = 0; i < n; ++i) h[(size_t)q * n + i] = v[q][i];
}

void unpack_vec(const std::vector<double> &h, int P, int n, double *const *v) {
  for (int q = 0; q < P; ++q)
    for (int i = 0; i < n; ++i) v[q][i] = h[(size_t)q * n + i];
}

// z = x*y for one series with K limbs (P = K or 2K planes of deg+1 doubles)
void product(int K, int cplx, double *const *x, double *const *y, double *const *z, int deg,
             int freq, const char *fn) {
  const int P = cplx ? 2 * K : K, D = deg + 1;
  std::vector<double> hx, hy, hz;
  pack_vec(x, P, D, hx);
  pack_vec(y, P, D, hy);
  Buf dx(hx.size()), dy(hy.size()), dz(hx.size());
  if (!dx.p || !dy.p || !dz.p) { ok(-1, fn); return; }
  if (!dx.up(hx) || !dy.up(hy)) { ok(-1, fn); return; }
  for (int r = 0; r < (freq > 0 ? freq : 1); ++r)
    if (!ok(phc_series_product(K, cplx, deg, 1, dx.d(), dy.d(), dz.d(), nullptr), fn)) return;
  if (!ok(phc_synchronize(), fn) || !dz.down(hz)) { ok(-1, fn); return; }
  unpack_vec(hz, P, D, z);
}

// complex multiply-adds of the blocked QR as this build runs it (panel dot
// products and updates, Gram matrix, S and W, the thin trailing products)
long long baqr_macs(long long m, long long n, long long szt) {
  long long macs = 0;
  const long long nbt = n / szt;
  for (long long k = 0; k < nbt; ++k) {
    const long long rowoff = k * szt, rowdim = m - rowoff;
    for (long long L = 0; L < szt; ++L) {
      const long long nr = m - (rowoff + L), nc = szt - L;
      macs += 2 * nr * nc + nr;
    }
    macs += szt * szt * rowdim + szt * szt * szt / 6 + rowdim * szt * szt;
    if (k < nbt - 1) macs += 2 * szt * (n - (k + 1) * szt) * rowdim;
    macs += 2 * m * szt * rowdim;
  }
  return macs;
}

// real: K planes per matrix and phc_real_blocked_houseqr
void blocked_qr(int K, int nrows, int ncols, int szt, int nbt, double **const *A,
                double **const *Q, double **const *R, double *walltimesec, long long *addcnt,
                long long *mulcnt, long long *divcnt, long long *sqrtcnt, const char *fn,
                bool real = false) {
  const Clock::time_point t0 = Clock::now();
  if (szt * nbt != ncols) {
    std::fprintf(stderr, "%s: ncols must equal szt*nbt\n", fn);
    return;
  }
  const int P = real ? K : 2 * K;
  std::vector<double> hA, hQ, hR;
  pack(A, P, nrows, ncols, hA);
  phc_baqr *plan = nullptr;
  if (!ok(phc_baqr_create(&plan, K, nrows, ncols, szt), fn)) return;
  {
    Buf dA(hA.size()), dQ((size_t)P * nrows * nrows), dR(hA.size());
    if (!dA.p || !dQ.p || !dR.p || !dA.up(hA)) {
      ok(-1, fn);
    } else if (ok(real ? phc_real_blocked_houseqr(plan, dA.d(), dQ.d(), dR.d(), nullptr)
                       : phc_cmplx_blocked_houseqr(plan, dA.d(), dQ.d(), dR.d(), nullptr),
                  fn) &&
               ok(phc_synchronize(), fn) && dQ.down(hQ) && dR.down(hR)) {
      unpack(hQ, P, nrows, nrows, Q);
      unpack(hR, P, nrows, ncols, R);
      const long long macs = baqr_macs(nrows, ncols, szt), per = real ? 1 : 4;
      if (addcnt) *addcnt += per * macs;
      if (mulcnt) *mulcnt += per * macs;
      if (divcnt) *divcnt += (real ? 1LL : 4LL) * ncols;
      if (sqrtcnt) *sqrtcnt += (real ? 1LL : 2LL) * ncols;
    }
  }
  phc_baqr_destroy(plan);
  if (walltimesec) *walltimesec += std::chrono::duration<double>(Clock::now() - t0).count();
}

void tiled_solver(int K, int dim, int szt, int nbt, double **const *U, double *const *b,
                  double *const *x, double *totlapms, double *walltimesec, long long *addcnt,
                  long long *mulcnt, long long *divcnt, const char *fn) {
  const Clock::time_point t0 = Clock::now();
  const int P = 2 * K;
  std::vector<double> hU, hb, hx;
  pack(U, P, dim, dim, hU);
  pack_vec(b, P, dim, hb);
  Buf dU(hU.size()), db(hb.size()), dx(hb.size());
  if (!dU.p || !db.p || !dx.p || !dU.up(hU) || !db.up(hb)) { ok(-1, fn); return; }
  void *e0 = nullptr, *e1 = nullptr;
  phc_event_create(&e0);
  phc_event_create(&e1);
  phc_event_record(e0, nullptr);
  const bool run = ok(phc_cmplx_upper_tiled_solver(K, dim, szt, nbt, dU.d(), db.d(), dx.d(), nullptr), fn);
  phc_event_record(e1, nullptr);
  if (run && ok(phc_synchronize(), fn) && dx.down(hx)) {
    unpack_vec(hx, P, dim, x);
    float ms = 0.0f;
    if (totlapms && phc_event_elapsed_ms(e0, e1, &ms) == 0) *totlapms += ms;
    // tile inverses szt^3/6 per tile, stage products dim^2/2, one complex
    // division per inverse entry of a diagonal
    const long long macs = (long long)nbt * szt * szt * szt / 6 + (long long)dim * dim / 2;
    if (addcnt) *addcnt += 4 * macs;
    if (mulcnt) *mulcnt += 4 * macs;
    if (divcnt) *divcnt += (long long)nbt * szt * szt / 2;
  }
  phc_event_destroy(e0);
  phc_event_destroy(e1);
  if (walltimesec) *walltimesec += std::chrono::duration<double>(Clock::now() - t0).count();
}

void qhb(int K, int n, double **const *Q, double *const *b, double *totlapms, const char *fn) {
  const int P = 2 * K;
  std::vector<double> hQ, hb, hy;
  pack(Q, P, n, n, hQ);
  pack_vec(b, P, n, hb);
  Buf dQ(hQ.size()), db(hb.size()), dy(hb.size());
  if (!dQ.p || !db.p || !dy.p || !dQ.up(hQ) || !db.up(hb)) { ok(-1, fn); return; }
  void *e0 = nullptr, *e1 = nullptr;
  phc_event_create(&e0);
  phc_event_create(&e1);
  phc_event_record(e0, nullptr);
  // y(i) = sum_k conj(Q(k, i)) b(k)
  const bool run = ok(phc_cmplx_gemm(K, 1, 0, n, 1, n, dQ.d(), (size_t)n * n, 1, n, db.d(),
                                     (size_t)n, 1, 0, dy.d(), (size_t)n, 1, 0, 0, nullptr), fn);
  phc_event_record(e1, nullptr);
  if (run && ok(phc_synchronize(), fn) && dy.down(hy)) {
    unpack_vec(hy, P, n, b);
    float ms = 0.0f;
    if (totlapms && phc_event_elapsed_ms(e0, e1, &ms) == 0) *totlapms += ms;
  }
  phc_event_destroy(e0);
  phc_event_destroy(e1);
}

// the block Toeplitz solve of one system (K limbs, real): planes of the
// degp1 matrices mat[q][j][r][c], Q, R, rhs, sol
void bals_solve(int K, int dim, int degp1, int tailidx, double ***const *mat,
                double **const *Q, double **const *R, double **const *rhs, double **const *sol,
                bool *zeroQ, bool *noqr, int *upidx, int *bsidx, int *newtail, double *totms,
                const char *fn) {
  const int n = dim;
  std::vector<double> hm((size_t)K * degp1 * n * n), hq, hr, hb, hx;
  for (int q = 0; q < K; ++q)
    for (int j = 0; j < degp1; ++j)
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) hm[(((size_t)q * degp1 + j) * n + r) * n + c] = mat[q][j][r][c];
  pack(Q, K, n, n, hq);
  pack(R, K, n, n, hr);
  pack(rhs, K, degp1, n, hb);
  pack(sol, K, degp1, n, hx);
  int hf[5] = {*zeroQ ? 1 : 0, *noqr ? 1 : 0, *upidx, *bsidx, *newtail};
  Buf dm(hm.size()), dq(hq.size()), dr(hr.size()), db(hb.size()), dx(hx.size());
  void *df = nullptr;
  if (!dm.p || !dq.p || !dr.p || !db.p || !dx.p || phc_malloc(&df, sizeof hf) != 0 ||
      !dm.up(hm) || !dq.up(hq) || !dr.up(hr) || !db.up(hb) || !dx.up(hx) ||
      phc_memcpy_htod(df, hf, sizeof hf) != 0) {
    ok(-1, fn);
    if (df) phc_free(df);
    return;
  }
  void *e0 = nullptr, *e1 = nullptr;
  phc_event_create(&e0);
  phc_event_create(&e1);
  phc_event_record(e0, nullptr);
  const bool run = ok(phc_qrbs_solve_batch(K, n, degp1, tailidx, 1, dm.d(), db.d(), dx.d(), dq.d(),
                                           dr.d(), static_cast<int *>(df), nullptr), fn);
  phc_event_record(e1, nullptr);
  if (run && ok(phc_synchronize(), fn) && dq.down(hq) && dr.down(hr) && db.down(hb) &&
      dx.down(hx) && phc_memcpy_dtoh(hf, df, sizeof hf) == 0) {
    unpack(hq, K, n, n, Q);
    unpack(hr, K, n, n, R);
    unpack(hb, K, degp1, n, rhs);
    unpack(hx, K, degp1, n, sol);
    *zeroQ = hf[0] != 0;
    *noqr = hf[1] != 0;
    *upidx = hf[2];
    *bsidx = hf[3];
    *newtail = hf[4];
    float ms = 0.0f;
    if (totms && phc_event_elapsed_ms(e0, e1, &ms) == 0) *totms += ms;
  }
  phc_event_destroy(e0);
  phc_event_destroy(e1);
  phc_free(df);
}

// r_i = b_i - sum_{j <= i-tailidx} A_j x_{i-j}, i = tailidx..degp1-1, of one
// system (K limbs, real); rows of resvec below tailidx are left as they are
void linear_residue(int K, int dim, int degp1, int tailidx, double ***const *mat,
                    double **const *rhs, double **const *sol, double **const *resvec,
                    double *const *resmax, double *lapms, long long *add, long long *mul,
                    const char *fn) {
  const int n = dim;
  std::vector<double> hm((size_t)K * degp1 * n * n), hb, hx, hr, hmax(K);
  for (int q = 0; q < K; ++q)
    for (int j = 0; j < degp1; ++j)
      for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) hm[(((size_t)q * degp1 + j) * n + r) * n + c] = mat[q][j][r][c];
  pack(rhs, K, degp1, n, hb);
  pack(sol, K, degp1, n, hx);
  pack(resvec, K, degp1, n, hr);
  Buf dm(hm.size()), db(hb.size()), dx(hx.size()), dr(hr.size()), dmax(K);
  if (!dm.p || !db.p || !dx.p || !dr.p || !dmax.p || !dm.up(hm) || !db.up(hb) || !dx.up(hx) ||
      !dr.up(hr)) {
    ok(-1, fn);
    return;
  }
  void *e0 = nullptr, *e1 = nullptr;
  phc_event_create(&e0);
  phc_event_create(&e1);
  phc_event_record(e0, nullptr);
  const bool run = ok(phc_linear_residue_batch(K, n, degp1, tailidx, 1, dm.d(), db.d(), dx.d(),
                                               dr.d(), dmax.d(), nullptr), fn);
  phc_event_record(e1, nullptr);
  if (run && ok(phc_synchronize(), fn) && dr.down(hr) && dmax.down(hmax)) {
    unpack(hr, K, degp1, n, resvec);
    for (int q = 0; q < K; ++q) *resmax[q] = hmax[q];
    // the reference's count: dim*dim additions and multiplications per
    // matrix-vector product, (i - tailidx + 1) products for residual i
    long long prods = 0;
    for (int i = tailidx; i < degp1; ++i) prods += i - tailidx + 1;
    *add = *mul = prods * (long long)dim * dim;
    float ms = 0.0f;
    if (lapms && phc_event_elapsed_ms(e0, e1, &ms) == 0) *lapms += ms;
  }
  phc_event_destroy(e0);
  phc_event_destroy(e1);
}

// value and gradient of one polynomial at series (P = K or 2K planes):
// cst[q] has deg+1 doubles, cff[q][k] the deg+1 coefficients of monomial k,
// input[q][i] the series of variable i, output[q][i] (i <= dim) receives
// the derivative to x_i, output[q][dim] the value
void poly_evaldiff(int K, int cplx, int dim, int nbr, int deg, const int *nvr, int **idx,
                   double *const *cst, double **const *cff, double **const *input,
                   double **const *output, double *cnvlapms, double *addlapms,
                   double *elapsedms, double *walltimesec, const char *fn) {
  const Clock::time_point t0 = Clock::now();
  const int P = cplx ? 2 * K : K, D = deg + 1;
  std::vector<int> flat;
  for (int k = 0; k < nbr; ++k)
    for (int i = 0; i < nvr[k]; ++i) flat.push_back(idx[k][i]);
  std::vector<double> hcst((size_t)P * D), hcff((size_t)P * nbr * D), hin, hout;
  for (int q = 0; q < P; ++q)
    for (int j = 0; j < D; ++j) hcst[(size_t)q * D + j] = cst[q][j];
  for (int q = 0; q < P; ++q)
    for (int k = 0; k < nbr; ++k)
      for (int j = 0; j < D; ++j) hcff[((size_t)q * nbr + k) * D + j] = cff[q][k][j];
  pack(input, P, dim, D, hin);
  phc_poly *plan = nullptr;
  if (!ok(phc_poly_create(&plan, K, cplx, dim, nbr, deg, nvr, flat.data(), hcst.data(),
                          hcff.data()), fn))
    return;
  float ms = 0.0f;
  {
    Buf din(hin.size()), dout((size_t)P * (dim + 1) * D);
    void *e0 = nullptr, *e1 = nullptr;
    if (!din.p || !dout.p || !din.up(hin)) {
      ok(-1, fn);
    } else if (ok(phc_event_create(&e0), fn) && ok(phc_event_create(&e1), fn) &&
               ok(phc_event_record(e0, nullptr), fn) &&
               ok(phc_poly_evaldiff(plan, 1, din.d(), dout.d(), nullptr), fn) &&
               ok(phc_event_record(e1, nullptr), fn) && ok(phc_synchronize(), fn) &&
               ok(phc_event_elapsed_ms(e0, e1, &ms), fn) && dout.down(hout)) {
      unpack(hout, P, dim + 1, D, output);
    }
    if (e0) phc_event_destroy(e0);
    if (e1) phc_event_destroy(e1);
  }
  phc_poly_destroy(plan);
  if (cnvlapms) *cnvlapms = ms;
  if (addlapms) *addlapms = 0.0;
  if (elapsedms) *elapsedms = ms;
  if (walltimesec) *walltimesec = std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace

// ------------------------------------------------------------ block Toeplitz solve

// the whole solve runs as one kernel: its time goes to *totbslapsedms, the
// QR, Q^T b and update timers are left unchanged
void GPU_dbl_bals_solve(int dim, int degp1, int, int, int tailidx, double ***mat, double **Q,
                        double **R, double **rhs, double **sol, bool *zeroQ, bool *noqr,
                        int *upidx, int *bsidx, int *newtail, double *, double *,
                        double *totbslapsedms, double *, int) {
  double ***m[] = {mat};
  double **q[] = {Q}, **r[] = {R}, **b[] = {rhs}, **x[] = {sol};
  bals_solve(1, dim, degp1, tailidx, m, q, r, b, x, zeroQ, noqr, upidx, bsidx, newtail,
             totbslapsedms, "GPU_dbl_bals_solve");
}

void GPU_dbl2_bals_solve(int dim, int degp1, int, int, int tailidx, double ***mathi,
                         double ***matlo, double **Qhi, double **Qlo, double **Rhi, double **Rlo,
                         double **rhshi, double **rhslo, double **solhi, double **sollo,
                         bool *zeroQ, bool *noqr, int *upidx, int *bsidx, int *newtail, double *,
                         double *, double *totbslapsedms, double *, int) {
  double ***m[] = {mathi, matlo};
  double **q[] = {Qhi, Qlo}, **r[] = {Rhi, Rlo}, **b[] = {rhshi, rhslo}, **x[] = {solhi, sollo};
  bals_solve(2, dim, degp1, tailidx, m, q, r, b, x, zeroQ, noqr, upidx, bsidx, newtail,
             totbslapsedms, "GPU_dbl2_bals_solve");
}

void GPU_dbl4_bals_solve(int dim, int degp1, int, int, int tailidx, double ***mathihi,
                         double ***matlohi, double ***mathilo, double ***matlolo,
                         double **Qhihi, double **Qlohi, double **Qhilo, double **Qlolo,
                         double **Rhihi, double **Rlohi, double **Rhilo, double **Rlolo,
                         double **rhshihi, double **rhslohi, double **rhshilo, double **rhslolo,
                         double **solhihi, double **sollohi, double **solhilo, double **sollolo,
                         bool *zeroQ, bool *noqr, int *upidx, int *bsidx, int *newtail, double *,
                         double *, double *totbslapsedms, double *, int) {
  double ***m[] = {mathihi, matlohi, mathilo, matlolo};
  double **q[] = {Qhihi, Qlohi, Qhilo, Qlolo}, **r[] = {Rhihi, Rlohi, Rhilo, Rlolo};
  double **b[] = {rhshihi, rhslohi, rhshilo, rhslolo}, **x[] = {solhihi, sollohi, solhilo, sollolo};
  bals_solve(4, dim, degp1, tailidx, m, q, r, b, x, zeroQ, noqr, upidx, bsidx, newtail,
             totbslapsedms, "GPU_dbl4_bals_solve");
}

// ------------------------------------------------------------ products

void GPU_dbl_product(double *x_h, double *y_h, double *z_h, int deg, int freq, int, int) {
  double *x[] = {x_h}, *y[] = {y_h}, *z[] = {z_h};
  product(1, 0, x, y, z, deg, freq, "GPU_dbl_product");
}

void GPU_cmplx_product(double *xre, double *xim, double *yre, double *yim, double *zre,
                       double *zim, int deg, int freq, int, int) {
  double *x[] = {xre, xim}, *y[] = {yre, yim}, *z[] = {zre, zim};
  product(1, 1, x, y, z, deg, freq, "GPU_cmplx_product");
}

void GPU_dbl2_product(double *xhi_h, double *xlo_h, double *yhi_h, double *ylo_h, double *zhi_h,
                      double *zlo_h, int deg, int freq, int, int) {
  double *x[] = {xhi_h, xlo_h}, *y[] = {yhi_h, ylo_h}, *z[] = {zhi_h, zlo_h};
  product(2, 0, x, y, z, deg, freq, "GPU_dbl2_product");
}

void GPU_cmplx2_product(double *xrehi_h, double *xrelo_h, double *ximhi_h, double *ximlo_h,
                        double *yrehi_h, double *yrelo_h, double *yimhi_h, double *yimlo_h,
                        double *zrehi_h, double *zrelo_h, double *zimhi_h, double *zimlo_h,
                        int deg, int freq, int, int) {
  double *x[] = {xrehi_h, xrelo_h, ximhi_h, ximlo_h};
  double *y[] = {yrehi_h, yrelo_h, yimhi_h, yimlo_h};
  double *z[] = {zrehi_h, zrelo_h, zimhi_h, zimlo_h};
  product(2, 1, x, y, z, deg, freq, "GPU_cmplx2_product");
}

void GPU_dbl4_product(double *xhihi_h, double *xlohi_h, double *xhilo_h, double *xlolo_h,
                      double *yhihi_h, double *ylohi_h, double *yhilo_h, double *ylolo_h,
                      double *zhihi_h, double *zlohi_h, double *zhilo_h, double *zlolo_h,
                      int deg, int freq, int, int) {
  double *x[] = {xhihi_h, xlohi_h, xhilo_h, xlolo_h};
  double *y[] = {yhihi_h, ylohi_h, yhilo_h, ylolo_h};
  double *z[] = {zhihi_h, zlohi_h, zhilo_h, zlolo_h};
  product(4, 0, x, y, z, deg, freq, "GPU_dbl4_product");
}

void GPU_cmplx4_product(double *xrehihi_h, double *xrelohi_h, double *xrehilo_h,
                        double *xrelolo_h, double *ximhihi_h, double *ximlohi_h,
                        double *ximhilo_h, double *ximlolo_h, double *yrehihi_h,
                        double *yrelohi_h, double *yrehilo_h, double *yrelolo_h,
                        double *yimhihi_h, double *yimlohi_h, double *yimhilo_h,
                        double *yimlolo_h, double *zrehihi_h, double *zrelohi_h,
                        double *zrehilo_h, double *zrelolo_h, double *zimhihi_h,
                        double *zimlohi_h, double *zimhilo_h, double *zimlolo_h, int deg,
                        int freq, int, int) {
  double *x[] = {xrehihi_h, xrelohi_h, xrehilo_h, xrelolo_h,
                 ximhihi_h, ximlohi_h, ximhilo_h, ximlolo_h};
  double *y[] = {yrehihi_h, yrelohi_h, yrehilo_h, yrelolo_h,
                 yimhihi_h, yimlohi_h, yimhilo_h, yimlolo_h};
  double *z[] = {zrehihi_h, zrelohi_h, zrehilo_h, zrelolo_h,
                 zimhihi_h, zimlohi_h, zimhilo_h, zimlolo_h};
  product(4, 1, x, y, z, deg, freq, "GPU_cmplx4_product");
}

// ------------------------------------------------------------ blocked QR

void GPU_dbl2_blocked_houseqr(int nrows, int ncols, int szt, int nbt, double **Ahi, double **Alo,
                              double **Qhi, double **Qlo, double **Rhi, double **Rlo, double *,
                              double *, double *, double *, double *, double *, double *,
                              double *, double *, double *, double *walltimesec,
                              long long int *addcnt, long long int *mulcnt,
                              long long int *divcnt, long long int *sqrtcnt, bool) {
  double **A[] = {Ahi, Alo};
  double **Q[] = {Qhi, Qlo};
  double **R[] = {Rhi, Rlo};
  blocked_qr(2, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_dbl2_blocked_houseqr", true);
}

void GPU_dbl4_blocked_houseqr(int nrows, int ncols, int szt, int nbt, double **Ahihi,
                              double **Alohi, double **Ahilo, double **Alolo, double **Qhihi,
                              double **Qlohi, double **Qhilo, double **Qlolo, double **Rhihi,
                              double **Rlohi, double **Rhilo, double **Rlolo, double *, double *,
                              double *, double *, double *, double *, double *, double *,
                              double *, double *, double *walltimesec, long long int *addcnt,
                              long long int *mulcnt, long long int *divcnt,
                              long long int *sqrtcnt, bool) {
  double **A[] = {Ahihi, Alohi, Ahilo, Alolo};
  double **Q[] = {Qhihi, Qlohi, Qhilo, Qlolo};
  double **R[] = {Rhihi, Rlohi, Rhilo, Rlolo};
  blocked_qr(4, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_dbl4_blocked_houseqr", true);
}

void GPU_dbl8_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Ahihihi, double **Alohihi, double **Ahilohi,
    double **Alolohi, double **Ahihilo, double **Alohilo, double **Ahilolo, double **Alololo,
    double **Qhihihi, double **Qlohihi, double **Qhilohi, double **Qlolohi, double **Qhihilo,
    double **Qlohilo, double **Qhilolo, double **Qlololo, double **Rhihihi, double **Rlohihi,
    double **Rhilohi, double **Rlolohi, double **Rhihilo, double **Rlohilo, double **Rhilolo,
    double **Rlololo, double *, double *, double *, double *, double *, double *, double *,
    double *, double *, double *, double *walltimesec, long long int *addcnt,
    long long int *mulcnt, long long int *divcnt, long long int *sqrtcnt, bool) {
  double **A[] = {Ahihihi, Alohihi, Ahilohi, Alolohi, Ahihilo, Alohilo, Ahilolo, Alololo};
  double **Q[] = {Qhihihi, Qlohihi, Qhilohi, Qlolohi, Qhihilo, Qlohilo, Qhilolo, Qlololo};
  double **R[] = {Rhihihi, Rlohihi, Rhilohi, Rlolohi, Rhihilo, Rlohilo, Rhilolo, Rlololo};
  blocked_qr(8, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_dbl8_blocked_houseqr", true);
}

void GPU_cmplx2_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Arehi, double **Arelo, double **Aimhi,
    double **Aimlo, double **Qrehi, double **Qrelo, double **Qimhi, double **Qimlo,
    double **Rrehi, double **Rrelo, double **Rimhi, double **Rimlo, double *, double *,
    double *, double *, double *, double *, double *, double *, double *, double *,
    double *walltimesec, long long int *addcnt, long long int *mulcnt, long long int *divcnt,
    long long int *sqrtcnt, bool) {
  double **A[] = {Arehi, Arelo, Aimhi, Aimlo};
  double **Q[] = {Qrehi, Qrelo, Qimhi, Qimlo};
  double **R[] = {Rrehi, Rrelo, Rimhi, Rimlo};
  blocked_qr(2, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_cmplx2_blocked_houseqr");
}

void GPU_cmplx4_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Arehihi, double **Arelohi,
    double **Arehilo, double **Arelolo, double **Aimhihi, double **Aimlohi, double **Aimhilo,
    double **Aimlolo, double **Qrehihi, double **Qrelohi, double **Qrehilo, double **Qrelolo,
    double **Qimhihi, double **Qimlohi, double **Qimhilo, double **Qimlolo, double **Rrehihi,
    double **Rrelohi, double **Rrehilo, double **Rrelolo, double **Rimhihi, double **Rimlohi,
    double **Rimhilo, double **Rimlolo, double *, double *, double *, double *, double *,
    double *, double *, double *, double *, double *, double *walltimesec,
    long long int *addcnt, long long int *mulcnt, long long int *divcnt,
    long long int *sqrtcnt, bool) {
  double **A[] = {Arehihi, Arelohi, Arehilo, Arelolo, Aimhihi, Aimlohi, Aimhilo, Aimlolo};
  double **Q[] = {Qrehihi, Qrelohi, Qrehilo, Qrelolo, Qimhihi, Qimlohi, Qimhilo, Qimlolo};
  double **R[] = {Rrehihi, Rrelohi, Rrehilo, Rrelolo, Rimhihi, Rimlohi, Rimhilo, Rimlolo};
  blocked_qr(4, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_cmplx4_blocked_houseqr");
}

void GPU_cmplx8_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Arehihihi, double **Arelohihi,
    double **Arehilohi, double **Arelolohi, double **Arehihilo, double **Arelohilo,
    double **Arehilolo, double **Arelololo, double **Aimhihihi, double **Aimlohihi,
    double **Aimhilohi, double **Aimlolohi, double **Aimhihilo, double **Aimlohilo,
    double **Aimhilolo, double **Aimlololo, double **Qrehihihi, double **Qrelohihi,
    double **Qrehilohi, double **Qrelolohi, double **Qrehihilo, double **Qrelohilo,
    double **Qrehilolo, double **Qrelololo, double **Qimhihihi, double **Qimlohihi,
    double **Qimhilohi, double **Qimlolohi, double **Qimhihilo, double **Qimlohilo,
    double **Qimhilolo, double **Qimlololo, double **Rrehihihi, double **Rrelohihi,
    double **Rrehilohi, double **Rrelolohi, double **Rrehihilo, double **Rrelohilo,
    double **Rrehilolo, double **Rrelololo, double **Rimhihihi, double **Rimlohihi,
    double **Rimhilohi, double **Rimlolohi, double **Rimhihilo, double **Rimlohilo,
    double **Rimhilolo, double **Rimlololo, double *, double *, double *, double *, double *,
    double *, double *, double *, double *, double *, double *walltimesec,
    long long int *addcnt, long long int *mulcnt, long long int *divcnt,
    long long int *sqrtcnt, bool) {
  double **A[] = {Arehihihi, Arelohihi, Arehilohi, Arelolohi, Arehihilo, Arelohilo,
                  Arehilolo, Arelololo, Aimhihihi, Aimlohihi, Aimhilohi, Aimlolohi,
                  Aimhihilo, Aimlohilo, Aimhilolo, Aimlololo};
  double **Q[] = {Qrehihihi, Qrelohihi, Qrehilohi, Qrelolohi, Qrehihilo, Qrelohilo,
                  Qrehilolo, Qrelololo, Qimhihihi, Qimlohihi, Qimhilohi, Qimlolohi,
                  Qimhihilo, Qimlohilo, Qimhilolo, Qimlololo};
  double **R[] = {Rrehihihi, Rrelohihi, Rrehilohi, Rrelolohi, Rrehihilo, Rrelohilo,
                  Rrehilolo, Rrelololo, Rimhihihi, Rimlohihi, Rimhilohi, Rimlolohi,
                  Rimhihilo, Rimlohilo, Rimhilolo, Rimlololo};
  blocked_qr(8, nrows, ncols, szt, nbt, A, Q, R, walltimesec, addcnt, mulcnt, divcnt, sqrtcnt,
             "GPU_cmplx8_blocked_houseqr");
}

// ------------------------------------------------------------ tiled solver

void GPU_cmplx2_upper_tiled_solver(int dim, int szt, int nbt, double **Urehi, double **Urelo,
                                   double **Uimhi, double **Uimlo, double *brehi, double *brelo,
                                   double *bimhi, double *bimlo, double *xrehi, double *xrelo,
                                   double *ximhi, double *ximlo, double *, double *, double *,
                                   double *totlapms, double *walltimesec, long long int *addcnt,
                                   long long int *mulcnt, long long int *divcnt) {
  double **U[] = {Urehi, Urelo, Uimhi, Uimlo};
  double *b[] = {brehi, brelo, bimhi, bimlo};
  double *x[] = {xrehi, xrelo, ximhi, ximlo};
  tiled_solver(2, dim, szt, nbt, U, b, x, totlapms, walltimesec, addcnt, mulcnt, divcnt,
               "GPU_cmplx2_upper_tiled_solver");
}

void GPU_cmplx4_upper_tiled_solver(
    int dim, int szt, int nbt, double **Urehihi, double **Urelohi, double **Urehilo,
    double **Urelolo, double **Uimhihi, double **Uimlohi, double **Uimhilo, double **Uimlolo,
    double *brehihi, double *brelohi, double *brehilo, double *brelolo, double *bimhihi,
    double *bimlohi, double *bimhilo, double *bimlolo, double *xrehihi, double *xrelohi,
    double *xrehilo, double *xrelolo, double *ximhihi, double *ximlohi, double *ximhilo,
    double *ximlolo, double *, double *, double *, double *totlapms, double *walltimesec,
    long long int *addcnt, double *, long long int *mulcnt, double *, long long int *divcnt,
    double *) {
  double **U[] = {Urehihi, Urelohi, Urehilo, Urelolo, Uimhihi, Uimlohi, Uimhilo, Uimlolo};
  double *b[] = {brehihi, brelohi, brehilo, brelolo, bimhihi, bimlohi, bimhilo, bimlolo};
  double *x[] = {xrehihi, xrelohi, xrehilo, xrelolo, ximhihi, ximlohi, ximhilo, ximlolo};
  tiled_solver(4, dim, szt, nbt, U, b, x, totlapms, walltimesec, addcnt, mulcnt, divcnt,
               "GPU_cmplx4_upper_tiled_solver");
}

void GPU_cmplx8_upper_tiled_solver(
    int dim, int szt, int nbt, double **Urehihihi, double **Urelohihi, double **Urehilohi,
    double **Urelolohi, double **Urehihilo, double **Urelohilo, double **Urehilolo,
    double **Urelololo, double **Uimhihihi, double **Uimlohihi, double **Uimhilohi,
    double **Uimlolohi, double **Uimhihilo, double **Uimlohilo, double **Uimhilolo,
    double **Uimlololo, double *brehihihi, double *brelohihi, double *brehilohi,
    double *brelolohi, double *brehihilo, double *brelohilo, double *brehilolo,
    double *brelololo, double *bimhihihi, double *bimlohihi, double *bimhilohi,
    double *bimlolohi, double *bimhihilo, double *bimlohilo, double *bimhilolo,
    double *bimlololo, double *xrehihihi, double *xrelohihi, double *xrehilohi,
    double *xrelolohi, double *xrehihilo, double *xrelohilo, double *xrehilolo,
    double *xrelololo, double *ximhihihi, double *ximlohihi, double *ximhilohi,
    double *ximlolohi, double *ximhihilo, double *ximlohilo, double *ximhilolo,
    double *ximlololo, double *, double *, double *, double *totlapms, double *walltimesec,
    long long int *addcnt, long long int *mulcnt, long long int *divcnt) {
  double **U[] = {Urehihihi, Urelohihi, Urehilohi, Urelolohi, Urehihilo, Urelohilo,
                  Urehilolo, Urelololo, Uimhihihi, Uimlohihi, Uimhilohi, Uimlolohi,
                  Uimhihilo, Uimlohilo, Uimhilolo, Uimlololo};
  double *b[] = {brehihihi, brelohihi, brehilohi, brelolohi, brehihilo, brelohilo,
                 brehilolo, brelololo, bimhihihi, bimlohihi, bimhilohi, bimlolohi,
                 bimhihilo, bimlohilo, bimhilolo, bimlololo};
  double *x[] = {xrehihihi, xrelohihi, xrehilohi, xrelolohi, xrehihilo, xrelohilo,
                 xrehilolo, xrelololo, ximhihihi, ximlohihi, ximhilohi, ximlolohi,
                 ximhihilo, ximlohilo, ximhilolo, ximlololo};
  tiled_solver(8, dim, szt, nbt, U, b, x, totlapms, walltimesec, addcnt, mulcnt, divcnt,
               "GPU_cmplx8_upper_tiled_solver");
}

// real U: the complex solver with zero imaginary parts (the imaginary part
// of x, zero, is discarded)
void real_tiled_solver(int K, int dim, int szt, int nbt, double **const *U, double *const *b,
                       double *const *x, double *totlapms, double *walltimesec, long long *addcnt,
                       long long *mulcnt, long long *divcnt, const char *fn) {
  std::vector<double> zeros((size_t)dim * dim, 0.0), xim((size_t)K * dim);
  std::vector<double *> zrows(dim);
  for (int i = 0; i < dim; ++i) zrows[i] = &zeros[(size_t)i * dim];
  double **Uc[8];
  double *bc[8], *xc[8];
  for (int q = 0; q < K; ++q) {
    Uc[q] = U[q];
    Uc[K + q] = zrows.data();
    bc[q] = b[q];
    bc[K + q] = zeros.data();
    xc[q] = x[q];
    xc[K + q] = &xim[(size_t)q * dim];
  }
  tiled_solver(K, dim, szt, nbt, Uc, bc, xc, totlapms, walltimesec, addcnt, mulcnt, divcnt, fn);
}

void GPU_dbl2_upper_tiled_solver(int dim, int szt, int nbt, double **Uhi, double **Ulo,
                                 double *bhi, double *blo, double *xhi, double *xlo, double *,
                                 double *, double *, double *totlapms, double *walltimesec,
                                 long long int *addcnt, long long int *mulcnt,
                                 long long int *divcnt) {
  double **U[] = {Uhi, Ulo};
  double *b[] = {bhi, blo}, *x[] = {xhi, xlo};
  real_tiled_solver(2, dim, szt, nbt, U, b, x, totlapms, walltimesec, addcnt, mulcnt, divcnt,
                    "GPU_dbl2_upper_tiled_solver");
}

void GPU_dbl4_upper_tiled_solver(int dim, int szt, int nbt, double **Uhihi, double **Ulohi,
                                 double **Uhilo, double **Ulolo, double *bhihi, double *blohi,
                                 double *bhilo, double *blolo, double *xhihi, double *xlohi,
                                 double *xhilo, double *xlolo, double *, double *, double *,
                                 double *totlapms, double *walltimesec, long long int *addcnt,
                                 double *, long long int *mulcnt, double *,
                                 long long int *divcnt, double *) {
  double **U[] = {Uhihi, Ulohi, Uhilo, Ulolo};
  double *b[] = {bhihi, blohi, bhilo, blolo}, *x[] = {xhihi, xlohi, xhilo, xlolo};
  real_tiled_solver(4, dim, szt, nbt, U, b, x, totlapms, walltimesec, addcnt, mulcnt, divcnt,
                    "GPU_dbl4_upper_tiled_solver");
}

// ------------------------------------------------------------ Q^H b

void GPU_cmplx2_bals_qhb(int ncols, int, int, double **Qrehi, double **Qrelo, double **Qimhi,
                         double **Qimlo, double *brehi, double *brelo, double *bimhi,
                         double *bimlo, double *totqtblapsedms, int) {
  double **Q[] = {Qrehi, Qrelo, Qimhi, Qimlo};
  double *b[] = {brehi, brelo, bimhi, bimlo};
  qhb(2, ncols, Q, b, totqtblapsedms, "GPU_cmplx2_bals_qhb");
}

void GPU_cmplx4_bals_qhb(int ncols, int, int, double **Qrehihi, double **Qrelohi,
                         double **Qrehilo, double **Qrelolo, double **Qimhihi, double **Qimlohi,
                         double **Qimhilo, double **Qimlolo, double *brehihi, double *brelohi,
                         double *brehilo, double *brelolo, double *bimhihi, double *bimlohi,
                         double *bimhilo, double *bimlolo, double *totqtblapsedms, int) {
  double **Q[] = {Qrehihi, Qrelohi, Qrehilo, Qrelolo, Qimhihi, Qimlohi, Qimhilo, Qimlolo};
  double *b[] = {brehihi, brelohi, brehilo, brelolo, bimhihi, bimlohi, bimhilo, bimlolo};
  qhb(4, ncols, Q, b, totqtblapsedms, "GPU_cmplx4_bals_qhb");
}

void GPU_dbl8_product(double *xhihihi_h, double *xlohihi_h, double *xhilohi_h,
                      double *xlolohi_h, double *xhihilo_h, double *xlohilo_h,
                      double *xhilolo_h, double *xlololo_h, double *yhihihi_h,
                      double *ylohihi_h, double *yhilohi_h, double *ylolohi_h,
                      double *yhihilo_h, double *ylohilo_h, double *yhilolo_h,
                      double *ylololo_h, double *zhihihi_h, double *zlohihi_h,
                      double *zhilohi_h, double *zlolohi_h, double *zhihilo_h,
                      double *zlohilo_h, double *zhilolo_h, double *zlololo_h, int deg, int freq,
                      int, int) {
  double *x[] = {xhihihi_h, xlohihi_h, xhilohi_h, xlolohi_h,
                 xhihilo_h, xlohilo_h, xhilolo_h, xlololo_h};
  double *y[] = {yhihihi_h, ylohihi_h, yhilohi_h, ylolohi_h,
                 yhihilo_h, ylohilo_h, yhilolo_h, ylololo_h};
  double *z[] = {zhihihi_h, zlohihi_h, zhilohi_h, zlolohi_h,
                 zhihilo_h, zlohilo_h, zhilolo_h, zlololo_h};
  product(8, 0, x, y, z, deg, freq, "GPU_dbl8_product");
}

void GPU_cmplx8_product(
    double *xrehihihi_h, double *xrelohihi_h, double *xrehilohi_h, double *xrelolohi_h,
    double *xrehihilo_h, double *xrelohilo_h, double *xrehilolo_h, double *xrelololo_h,
    double *ximhihihi_h, double *ximlohihi_h, double *ximhilohi_h, double *ximlolohi_h,
    double *ximhihilo_h, double *ximlohilo_h, double *ximhilolo_h, double *ximlololo_h,
    double *yrehihihi_h, double *yrelohihi_h, double *yrehilohi_h, double *yrelolohi_h,
    double *yrehihilo_h, double *yrelohilo_h, double *yrehilolo_h, double *yrelololo_h,
    double *yimhihihi_h, double *yimlohihi_h, double *yimhilohi_h, double *yimlolohi_h,
    double *yimhihilo_h, double *yimlohilo_h, double *yimhilolo_h, double *yimlololo_h,
    double *zrehihihi_h, double *zrelohihi_h, double *zrehilohi_h, double *zrelolohi_h,
    double *zrehihilo_h, double *zrelohilo_h, double *zrehilolo_h, double *zrelololo_h,
    double *zimhihihi_h, double *zimlohihi_h, double *zimhilohi_h, double *zimlolohi_h,
    double *zimhihilo_h, double *zimlohilo_h, double *zimhilolo_h, double *zimlololo_h, int deg,
    int freq, int, int) {
  double *x[] = {xrehihihi_h, xrelohihi_h, xrehilohi_h, xrelolohi_h,
                 xrehihilo_h, xrelohilo_h, xrehilolo_h, xrelololo_h,
                 ximhihihi_h, ximlohihi_h, ximhilohi_h, ximlolohi_h,
                 ximhihilo_h, ximlohilo_h, ximhilolo_h, ximlololo_h};
  double *y[] = {yrehihihi_h, yrelohihi_h, yrehilohi_h, yrelolohi_h,
                 yrehihilo_h, yrelohilo_h, yrehilolo_h, yrelololo_h,
                 yimhihihi_h, yimlohihi_h, yimhilohi_h, yimlolohi_h,
                 yimhihilo_h, yimlohilo_h, yimhilolo_h, yimlololo_h};
  double *z[] = {zrehihihi_h, zrelohihi_h, zrehilohi_h, zrelolohi_h,
                 zrehihilo_h, zrelohilo_h, zrehilolo_h, zrelololo_h,
                 zimhihihi_h, zimlohihi_h, zimhilohi_h, zimlolohi_h,
                 zimhihilo_h, zimlohilo_h, zimhilolo_h, zimlololo_h};
  product(8, 1, x, y, z, deg, freq, "GPU_cmplx8_product");
}

// ------------------------------------------------------------ linear residue

void GPU_dbl_linear_residue(int dim, int degp1, int, int, int tailidx, double ***mat,
                            double **rhs, double **sol, double **resvec, double *resmax,
                            double *totreslapsedms, long long int *add, long long int *mul, int) {
  double ***m[] = {mat};
  double **b[] = {rhs}, **x[] = {sol}, **r[] = {resvec};
  double *mx[] = {resmax};
  linear_residue(1, dim, degp1, tailidx, m, b, x, r, mx, totreslapsedms, add, mul,
                 "GPU_dbl_linear_residue");
}

void GPU_dbl2_linear_residue(int dim, int degp1, int, int, int tailidx, double ***mathi,
                             double ***matlo, double **rhshi, double **rhslo, double **solhi,
                             double **sollo, double **resvechi, double **resveclo,
                             double *resmaxhi, double *resmaxlo, double *lapms,
                             long long int *add, long long int *mul, int) {
  double ***m[] = {mathi, matlo};
  double **b[] = {rhshi, rhslo}, **x[] = {solhi, sollo}, **r[] = {resvechi, resveclo};
  double *mx[] = {resmaxhi, resmaxlo};
  linear_residue(2, dim, degp1, tailidx, m, b, x, r, mx, lapms, add, mul,
                 "GPU_dbl2_linear_residue");
}

void GPU_dbl4_linear_residue(int dim, int degp1, int, int, int tailidx, double ***mathihi,
                             double ***matlohi, double ***mathilo, double ***matlolo,
                             double **rhshihi, double **rhslohi, double **rhshilo,
                             double **rhslolo, double **solhihi, double **sollohi,
                             double **solhilo, double **sollolo, double **resvechihi,
                             double **resveclohi, double **resvechilo, double **resveclolo,
                             double *resmaxhihi, double *resmaxlohi, double *resmaxhilo,
                             double *resmaxlolo, double *lapms, long long int *add,
                             long long int *mul, int) {
  double ***m[] = {mathihi, matlohi, mathilo, matlolo};
  double **b[] = {rhshihi, rhslohi, rhshilo, rhslolo};
  double **x[] = {solhihi, sollohi, solhilo, sollolo};
  double **r[] = {resvechihi, resveclohi, resvechilo, resveclolo};
  double *mx[] = {resmaxhihi, resmaxlohi, resmaxhilo, resmaxlolo};
  linear_residue(4, dim, degp1, tailidx, m, b, x, r, mx, lapms, add, mul,
                 "GPU_dbl4_linear_residue");
}

// ------------------------------------------------ polynomial evaluation

void GPU_dbl_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx, double *cst,
                           double **cff, double **input, double **output, ConvolutionJobs,
                           AdditionJobs, double *cnvlapms, double *addlapms, double *elapsedms,
                           double *walltimesec, int) {
  double *c[] = {cst};
  double **f[] = {cff}, **x[] = {input}, **y[] = {output};
  poly_evaldiff(1, 0, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_dbl_poly_evaldiff");
}

void GPU_dbl2_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx, double *csthi,
                            double *cstlo, double **cffhi, double **cfflo, double **inputhi,
                            double **inputlo, double **outputhi, double **outputlo,
                            ConvolutionJobs, AdditionJobs, double *cnvlapms, double *addlapms,
                            double *elapsedms, double *walltimesec, int) {
  double *c[] = {csthi, cstlo};
  double **f[] = {cffhi, cfflo}, **x[] = {inputhi, inputlo}, **y[] = {outputhi, outputlo};
  poly_evaldiff(2, 0, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_dbl2_poly_evaldiff");
}

void GPU_dbl4_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx,
                            double *csthihi, double *cstlohi, double *csthilo, double *cstlolo,
                            double **cffhihi, double **cfflohi, double **cffhilo,
                            double **cfflolo, double **inputhihi, double **inputlohi,
                            double **inputhilo, double **inputlolo, double **outputhihi,
                            double **outputlohi, double **outputhilo, double **outputlolo,
                            ConvolutionJobs, AdditionJobs, double *cnvlapms, double *addlapms,
                            double *elapsedms, double *walltimesec, int) {
  double *c[] = {csthihi, cstlohi, csthilo, cstlolo};
  double **f[] = {cffhihi, cfflohi, cffhilo, cfflolo};
  double **x[] = {inputhihi, inputlohi, inputhilo, inputlolo};
  double **y[] = {outputhihi, outputlohi, outputhilo, outputlolo};
  poly_evaldiff(4, 0, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_dbl4_poly_evaldiff");
}

void GPU_dbl8_poly_evaldiff(
    int, int dim, int nbr, int deg, int *nvr, int **idx, double *csthihihi, double *cstlohihi,
    double *csthilohi, double *cstlolohi, double *csthihilo, double *cstlohilo,
    double *csthilolo, double *cstlololo, double **cffhihihi, double **cfflohihi,
    double **cffhilohi, double **cfflolohi, double **cffhihilo, double **cfflohilo,
    double **cffhilolo, double **cfflololo, double **inputhihihi, double **inputlohihi,
    double **inputhilohi, double **inputlolohi, double **inputhihilo, double **inputlohilo,
    double **inputhilolo, double **inputlololo, double **outputhihihi, double **outputlohihi,
    double **outputhilohi, double **outputlolohi, double **outputhihilo, double **outputlohilo,
    double **outputhilolo, double **outputlololo, ConvolutionJobs, AdditionJobs,
    double *cnvlapms, double *addlapms, double *elapsedms, double *walltimesec, int) {
  double *c[] = {csthihihi, cstlohihi, csthilohi, cstlolohi,
                 csthihilo, cstlohilo, csthilolo, cstlololo};
  double **f[] = {cffhihihi, cfflohihi, cffhilohi, cfflolohi,
                  cffhihilo, cfflohilo, cffhilolo, cfflololo};
  double **x[] = {inputhihihi, inputlohihi, inputhilohi, inputlolohi,
                  inputhihilo, inputlohilo, inputhilolo, inputlololo};
  double **y[] = {outputhihihi, outputlohihi, outputhilohi, outputlolohi,
                  outputhihilo, outputlohilo, outputhilolo, outputlololo};
  poly_evaldiff(8, 0, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_dbl8_poly_evaldiff");
}

void GPU_cmplx_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx, double *cstre,
                             double *cstim, double **cffre, double **cffim, double **inputre,
                             double **inputim, double **outputre, double **outputim,
                             ConvolutionJobs, AdditionJobs, double *cnvlapms, double *addlapms,
                             double *elapsedms, double *walltimesec, int) {
  double *c[] = {cstre, cstim};
  double **f[] = {cffre, cffim}, **x[] = {inputre, inputim}, **y[] = {outputre, outputim};
  poly_evaldiff(1, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplx_poly_evaldiff");
}

void GPU_cmplx2_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx,
                              double *cstrehi, double *cstrelo, double *cstimhi,
                              double *cstimlo, double **cffrehi, double **cffrelo,
                              double **cffimhi, double **cffimlo, double **inputrehi,
                              double **inputrelo, double **inputimhi, double **inputimlo,
                              double **outputrehi, double **outputrelo, double **outputimhi,
                              double **outputimlo, ConvolutionJobs, AdditionJobs,
                              double *cnvlapms, double *addlapms, double *elapsedms,
                              double *walltimesec, int) {
  double *c[] = {cstrehi, cstrelo, cstimhi, cstimlo};
  double **f[] = {cffrehi, cffrelo, cffimhi, cffimlo};
  double **x[] = {inputrehi, inputrelo, inputimhi, inputimlo};
  double **y[] = {outputrehi, outputrelo, outputimhi, outputimlo};
  poly_evaldiff(2, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplx2_poly_evaldiff");
}

void GPU_cmplx4_poly_evaldiff(
    int, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehihi, double *cstrelohi,
    double *cstrehilo, double *cstrelolo, double *cstimhihi, double *cstimlohi,
    double *cstimhilo, double *cstimlolo, double **cffrehihi, double **cffrelohi,
    double **cffrehilo, double **cffrelolo, double **cffimhihi, double **cffimlohi,
    double **cffimhilo, double **cffimlolo, double **inputrehihi, double **inputrelohi,
    double **inputrehilo, double **inputrelolo, double **inputimhihi, double **inputimlohi,
    double **inputimhilo, double **inputimlolo, double **outputrehihi, double **outputrelohi,
    double **outputrehilo, double **outputrelolo, double **outputimhihi, double **outputimlohi,
    double **outputimhilo, double **outputimlolo, ConvolutionJobs, AdditionJobs,
    double *cnvlapms, double *addlapms, double *elapsedms, double *walltimesec, int) {
  double *c[] = {cstrehihi, cstrelohi, cstrehilo, cstrelolo,
                 cstimhihi, cstimlohi, cstimhilo, cstimlolo};
  double **f[] = {cffrehihi, cffrelohi, cffrehilo, cffrelolo,
                  cffimhihi, cffimlohi, cffimhilo, cffimlolo};
  double **x[] = {inputrehihi, inputrelohi, inputrehilo, inputrelolo,
                  inputimhihi, inputimlohi, inputimhilo, inputimlolo};
  double **y[] = {outputrehihi, outputrelohi, outputrehilo, outputrelolo,
                  outputimhihi, outputimlohi, outputimhilo, outputimlolo};
  poly_evaldiff(4, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplx4_poly_evaldiff");
}

// the vectorized complex evaluation (Polynomials/dbl_polynomials_kernels.h:735,
// dbl2_polynomials_kernels.h:960, dbl4_polynomials_kernels.h:1397): the same
// value and derivatives, with CPU_cmplx*_poly_evaldiff's operation order
void GPU_cmplxvectorized_poly_evaldiff(int, int dim, int nbr, int deg, int *nvr, int **idx,
                                       double *cstre, double *cstim, double **cffre,
                                       double **cffim, double **inputre, double **inputim,
                                       double **outputre, double **outputim,
                                       ComplexConvolutionJobs, ComplexIncrementJobs,
                                       ComplexAdditionJobs, double *cnvlapms, double *addlapms,
                                       double *elapsedms, double *walltimesec, int) {
  double *c[] = {cstre, cstim};
  double **f[] = {cffre, cffim}, **x[] = {inputre, inputim}, **y[] = {outputre, outputim};
  poly_evaldiff(1, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplxvectorized_poly_evaldiff");
}

void GPU_cmplx2vectorized_poly_evaldiff(
    int, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehi, double *cstrelo,
    double *cstimhi, double *cstimlo, double **cffrehi, double **cffrelo, double **cffimhi,
    double **cffimlo, double **inputrehi, double **inputrelo, double **inputimhi,
    double **inputimlo, double **outputrehi, double **outputrelo, double **outputimhi,
    double **outputimlo, ComplexConvolutionJobs, ComplexIncrementJobs, ComplexAdditionJobs,
    double *cnvlapms, double *addlapms, double *elapsedms, double *walltimesec, int) {
  double *c[] = {cstrehi, cstrelo, cstimhi, cstimlo};
  double **f[] = {cffrehi, cffrelo, cffimhi, cffimlo};
  double **x[] = {inputrehi, inputrelo, inputimhi, inputimlo};
  double **y[] = {outputrehi, outputrelo, outputimhi, outputimlo};
  poly_evaldiff(2, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplx2vectorized_poly_evaldiff");
}

void GPU_cmplx4vectorized_poly_evaldiff(
    int, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehihi, double *cstrelohi,
    double *cstrehilo, double *cstrelolo, double *cstimhihi, double *cstimlohi,
    double *cstimhilo, double *cstimlolo, double **cffrehihi, double **cffrelohi,
    double **cffrehilo, double **cffrelolo, double **cffimhihi, double **cffimlohi,
    double **cffimhilo, double **cffimlolo, double **inputrehihi, double **inputrelohi,
    double **inputrehilo, double **inputrelolo, double **inputimhihi, double **inputimlohi,
    double **inputimhilo, double **inputimlolo, double **outputrehihi, double **outputrelohi,
    double **outputrehilo, double **outputrelolo, double **outputimhihi, double **outputimlohi,
    double **outputimhilo, double **outputimlolo, ComplexConvolutionJobs, ComplexIncrementJobs,
    ComplexAdditionJobs, double *cnvlapms, double *addlapms, double *elapsedms,
    double *walltimesec, int) {
  double *c[] = {cstrehihi, cstrelohi, cstrehilo, cstrelolo,
                 cstimhihi, cstimlohi, cstimhilo, cstimlolo};
  double **f[] = {cffrehihi, cffrelohi, cffrehilo, cffrelolo,
                  cffimhihi, cffimlohi, cffimhilo, cffimlolo};
  double **x[] = {inputrehihi, inputrelohi, inputrehilo, inputrelolo,
                  inputimhihi, inputimlohi, inputimhilo, inputimlolo};
  double **y[] = {outputrehihi, outputrelohi, outputrehilo, outputrelolo,
                  outputimhihi, outputimlohi, outputimhilo, outputimlolo};
  poly_evaldiff(4, 1, dim, nbr, deg, nvr, idx, c, f, x, y, cnvlapms, addlapms, elapsedms,
                walltimesec, "GPU_cmplx4vectorized_poly_evaldiff");
}
