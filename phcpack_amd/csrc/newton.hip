// Linearized power-series Newton on real and complex multiple-double series,
// batched: one wavefront runs the whole staggered Newton method of one system
// in LDS.  The complex twins (CX) follow CPU_cmplx2_factors_houseqr,
// CPU_cmplx2_factors_qrbs, CPU_cmplx2_qrbs_head/tail/solve,
// CPU_cmplx2_linear_residue and cmplx2_row_newton_qrstep
// (Newton/cmplx2_newton_method.cpp:565) with their operation orders.
//
//   reference (host)                          here
//   CPU_dbl2_poly_evaldiff per polynomial     eval_system (slots of lanes, one
//     (Polynomials/dbl2_polynomials_host.cpp)   polynomial per slot, lane = coefficient)
//   dbl2_map_evaldiff_output, rhs = -f        read in place from the output rows
//     (Newton/dbl2_systems_host.cpp:709)
//   CPU_dbl2_factors_houseqr (Matrices/       houseqr: lane per column for the
//     dbl2_factorizations.cpp:723)              reflector products, lane per entry for updates
//   CPU_dbl2_factors_qrbs / _backward         qrbs: lane per row for Q^T b, column sweep
//     (:823, :121)                              for the back substitution
//   CPU_dbl2_qrbs_head/tail/solve             qrbs_solve: same skip rules and flags
//     (Newton/dbl2_bals_host.cpp:15,228,479)
//   CPU_dbl2_linear_residue (:548)            linear_residue
//   dbl2_row_newton_qrstep (Newton/           row_newton_kernel: the staggered loop of
//     dbl2_newton_method.cpp:429) and the       test_dbl2_row_newton (:1447-1494) in-kernel
//     staggered loop
//
// Every number is computed by the same sequence of multiple-double operations
// as the reference's host code (each lane keeps the reference's accumulation
// order for the entries it owns; lanes split only independent entries), so
// the results are bit-identical to the CPU oracle.  Lanes of one wavefront
// exchange data through LDS; within a wavefront LDS operations complete in
// program order, and wsync() keeps the compiler from moving them.
#include <vector>

#include "common.hpp"
#include "mdtypes.hpp"
#include "wave.hpp"

using namespace md;
using phc::grid_for;

namespace {

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ double wave_fmax(double v) { return wave::max64(v); }

constexpr size_t LDS_LIMIT = 160 * 1024;

// skip thresholds of the tail solve (Newton/dbl_bals_host.cpp:199,
// dbl2_bals_host.cpp:43, dbl4_bals_host.cpp:51, dbl8_bals_host.cpp:70)
template <int K> __host__ __device__ constexpr double threshold() {
  return K == 1 ? 1.0e-15 : K == 2 ? 1.0e-28 : K == 4 ? 1.0e-56 : 1.0e-120;
}

// Real (CX = false) or complex (CX = true, K real planes then K imaginary
// planes) multiple-double numbers: the operations the Newton method needs,
// each in the order of the reference's host code of that type.
template <int K, bool CX> struct NT;
template <int K> struct NT<K, false> {
  typedef typename RealT<K>::type V;
  static constexpr int P = K;
  __device__ static V ld(const double *b, size_t m, size_t j) { return R<K>::load(b, m, j); }
  __device__ static void st(double *b, size_t m, size_t j, const V &v) { R<K>::store(b, m, j, v); }
  __device__ static V zero() { return R<K>::zero(); }
  __device__ static V one() { return R<K>::from(1.0); }
  __device__ static V add(const V &a, const V &b) { return md::add(a, b); }
  __device__ static V neg(const V &a) { return md::neg(a); }
  __device__ static V mul(const V &a, const V &b) { return md::mul(a, b); }
  __device__ static V mac(const V &c, const V &a, const V &b) { return md::add(c, md::mul(a, b)); }
  // b - a x, the update of the tail solve and the residual
  __device__ static V msub(const V &b, const V &a, const V &x) { return md::sub(b, md::mul(a, x)); }
  // the magnitude CPU_dbl_onenorm takes: |leading limb|
  __device__ static double mag(const V &a) { return fabs(R<K>::hi(a)); }
};
template <int K> struct NT<K, true> {
  typedef cplx<K> V;
  static constexpr int P = 2 * K;
  __device__ static V ld(const double *b, size_t m, size_t j) { return C<K>::load(b, m, j); }
  __device__ static void st(double *b, size_t m, size_t j, const V &v) { C<K>::store(b, m, j, v); }
  __device__ static V zero() { return C<K>::zero(); }
  __device__ static V one() { return C<K>::from(1.0, 0.0); }
  __device__ static V add(const V &a, const V &b) { return C<K>::add(a, b); }
  __device__ static V neg(const V &a) { return C<K>::neg(a); }
  // the first term and the inner loop of CPU_cmplx2_product
  __device__ static V mul(const V &a, const V &b) { return C<K>::mul(a, b); }
  __device__ static V mac(const V &c, const V &a, const V &b) {
    V z = c;
    C<K>::mac(z, a, b);
    return z;
  }
  // b - a x as CPU_cmplx2_qrbs_tail and CPU_cmplx2_linear_residue subtract
  // it: re -= a.re x.re, re += a.im x.im, im -= a.re x.im, im -= a.im x.re
  __device__ static V msub(const V &b, const V &a, const V &x) {
    V z;
    z.re = md::sub(b.re, md::mul(a.re, x.re));
    z.re = md::add(z.re, md::mul(a.im, x.im));
    z.im = md::sub(b.im, md::mul(a.re, x.im));
    z.im = md::sub(z.im, md::mul(a.im, x.re));
    return z;
  }
  // CPU_cmplx_onenorm's |re| + |im| of the leading limbs
  __device__ static double mag(const V &a) { return fabs(R<K>::hi(a.re)) + fabs(R<K>::hi(a.im)); }
};

// a planar LDS (or global) array of m numbers with K limbs (2K planes when
// complex)
template <int K, bool CX = false> struct Arr {
  typedef typename NT<K, CX>::V T;
  double *b;
  int m;
  __device__ T ld(int j) const { return NT<K, CX>::ld(b, m, j); }
  __device__ void st(int j, const T &v) const { NT<K, CX>::st(b, m, j, v); }
};

__device__ __forceinline__ bool all_zero(double a) { return a == 0.0; }
template <int K> __device__ __forceinline__ bool all_zero(const mdr<K> &a) {
  bool z = true;
#pragma unroll
  for (int i = 0; i < K; ++i) z = z && (a.x[i] == 0.0);
  return z;
}
__device__ __forceinline__ double limb(double a, int) { return a; }
template <int K> __device__ __forceinline__ double limb(const mdr<K> &a, int i) { return a.x[i]; }

// carve a planar array of m numbers out of LDS
template <int K, bool CX = false> __device__ Arr<K, CX> carve(double *&p, int m) {
  Arr<K, CX> a{p, m};
  p += (size_t)NT<K, CX>::P * m;
  return a;
}

// largest |leading limb| of v[0..n-1], as CPU_dbl_onenorm computes it
// (Newton/dbl_onenorms_host.cpp:9-17): NaN when v[0] is NaN, later NaNs
// skipped.  Uniform over the wavefront; every lane must call it.
template <int K, bool CX = false> __device__ double maxnorm(const Arr<K, CX> &v, int off, int n) {
  const int l = lane_id();
  double r = 0.0, v0 = 0.0;
  for (int i = l; i < n; i += 64) {
    const double a = NT<K, CX>::mag(v.ld(off + i));
    if (!(a <= r)) r = (a != a) ? r : a;
  }
  if (l == 0) v0 = NT<K, CX>::mag(v.ld(off));
  r = wave_fmax(r);
  v0 = __shfl(v0, 0, 64);
  return (v0 != v0) ? v0 : r;
}

// ------------------------------------------------ Householder QR (square)

// per-wave work arrays of the linear algebra; sc[0] = beta, sc[1] = v0
// (complex: 1/v0), sc[2] = 1 when sigma = 0
template <int K, bool CX = false> struct LinWork {
  Arr<K, CX> q, r, v, w1, w2, wrk, sc;
};

// Q R = A for the n x n matrix given by the accessor a(r, c)
// (CPU_dbl2_factors_houseqr with CPU_dbl2_factors_house, leftRupdate and
// rightQupdate, Matrices/dbl2_factorizations.cpp:374-440, 534-566, 629-660)
template <int K, class A> __device__ void houseqr(int n, const A &a, const LinWork<K> &w) {
  typedef typename RealT<K>::type T;
  const int l = lane_id();
  for (int t = l; t < n * n; t += 64) {
    const int i = t / n, j = t % n;
    w.q.st(t, R<K>::from(i == j ? 1.0 : 0.0));
    w.r.st(t, a(i, j));
  }
  wsync();
  for (int k = 0; k < n; ++k) {
    const int m = n - k;
    // Householder vector of x = R[k..n-1][k]: sigma, mu, v0, beta in lane 0
    if (l == 0) {
      T sigma = R<K>::zero();
      for (int i = 1; i < m; ++i) sigma = add(sigma, sqr(w.r.ld((k + i) * n + k)));
      const bool zero = all_zero(sigma);
      w.sc.st(2, R<K>::from(zero ? 1.0 : 0.0));
      if (zero) {
        w.sc.st(0, R<K>::zero());
        w.sc.st(1, R<K>::from(1.0));
      } else {
        const T x0 = w.r.ld(k * n + k);
        T acc = sqr(x0);
        acc = add(acc, sigma);
        const T mu = R<K>::sqrt(acc);
        T v0;
        if (R<K>::hi(x0) <= 0.0) {
          v0 = sub(x0, mu);
        } else {
          acc = add(x0, mu);
          v0 = neg(div(sigma, acc));
        }
        const T v0p2 = sqr(v0);
        acc = add(sigma, v0p2);
        T beta = div(v0p2, acc);
        beta = mul_d(beta, 2.0);
        w.sc.st(0, beta);
        w.sc.st(1, v0);
      }
    }
    wsync();
    {
      const T v0 = w.sc.ld(1);
      const bool szero = R<K>::hi(w.sc.ld(2)) != 0.0;
      // v[i] = x[i] / v0 (x[i] unchanged when sigma = 0), v[0] = 1
      for (int i = l; i < m; i += 64) {
        if (i == 0) w.v.st(0, R<K>::from(1.0));
        else {
          const T xi = w.r.ld((k + i) * n + k);
          w.v.st(i, szero ? xi : div(xi, v0));
        }
      }
    }
    wsync();
    const T beta = w.sc.ld(0);
    // w1[j-k] = beta * sum_i R[i][j] v[i-k]
    for (int j = k + l; j < n; j += 64) {
      T s = R<K>::zero();
      for (int i = k; i < n; ++i) s = add(s, mul(w.r.ld(i * n + j), w.v.ld(i - k)));
      w.w1.st(j - k, mul(beta, s));
    }
    // w2[i] = beta * sum_{j>=k} Q[i][j] v[j-k]
    for (int i = l; i < n; i += 64) {
      T s = R<K>::zero();
      for (int j = k; j < n; ++j) s = add(s, mul(w.q.ld(i * n + j), w.v.ld(j - k)));
      w.w2.st(i, mul(beta, s));
    }
    wsync();
    // R[i][j] -= v[i-k] w1[j-k];  Q[i][j] -= w2[i] v[j-k]
    for (int t = l; t < m * m; t += 64) {
      const int i = k + t / m, j = k + t % m;
      w.r.st(i * n + j, sub(w.r.ld(i * n + j), mul(w.v.ld(i - k), w.w1.ld(j - k))));
    }
    for (int t = l; t < n * m; t += 64) {
      const int i = t / m, j = k + t % m;
      w.q.st(i * n + j, sub(w.q.ld(i * n + j), mul(w.w2.ld(i), w.v.ld(j - k))));
    }
    wsync();
  }
}

// x = R^{-1} Q^T b (CPU_dbl2_factors_qrbs and _backward): b = bv[bo..],
// x = xv[xo..]; b and x may be the same row.
template <int K>
__device__ void qrbs(int n, const LinWork<K> &w, const Arr<K> &bv, int bo, const Arr<K> &xv,
                     int xo) {
  typedef typename RealT<K>::type T;
  const int l = lane_id();
  for (int i = l; i < n; i += 64) {
    T s = R<K>::zero();
    for (int j = 0; j < n; ++j) s = add(s, mul(w.q.ld(j * n + i), bv.ld(bo + j)));
    w.wrk.st(i, s);
  }
  wsync();
  for (int i = l; i < n; i += 64) xv.st(xo + i, w.wrk.ld(i));
  wsync();
  // column sweep: row i receives the terms j = n-1 .. i+1 in that order
  for (int j = n - 1; j >= 0; --j) {
    if (l == (j & 63)) xv.st(xo + j, div(xv.ld(xo + j), w.r.ld(j * n + j)));
    wsync();
    const T xj = xv.ld(xo + j);
    for (int i = l; i < j; i += 64) xv.st(xo + i, sub(xv.ld(xo + i), mul(w.r.ld(i * n + j), xj)));
    wsync();
  }
}

// Complex Q R = A (CPU_cmplx2_factors_houseqr with CPU_cmplx2_factors_house,
// leftRupdate and rightQupdate, Matrices/dbl2_factorizations.cpp:441-533,
// 567-627, 661-722, 766-822): lane 0 forms the Householder vector's scalars
// (sigma = sum |x_i|^2, mu = sqrt(|x0|^2 + sigma) / |x0|, v0 = x0 - mu x0,
// beta = 2 |v0|^2 / (sigma + |v0|^2), 1/v0 = conj(v0) / |v0|^2), the lanes
// v_i = x_i / v0, then w1 = beta R^H v, w2 = beta Q v and the rank-one
// updates R -= v w1^H, Q -= w2 v^H, every entry in the reference's order.
template <int K, class A> __device__ void cx_houseqr(int n, const A &a, const LinWork<K, true> &w) {
  typedef typename RealT<K>::type T;
  typedef cplx<K> Z;
  const int l = lane_id();
  for (int t = l; t < n * n; t += 64) {
    const int i = t / n, j = t % n;
    w.q.st(t, C<K>::from(i == j ? 1.0 : 0.0, 0.0));
    w.r.st(t, a(i, j));
  }
  wsync();
  for (int k = 0; k < n; ++k) {
    const int m = n - k;
    if (l == 0) {
      T sigma = R<K>::zero();
      for (int i = 1; i < m; ++i) {
        const Z x = w.r.ld((k + i) * n + k);
        sigma = add(sigma, sqr(x.re));
        sigma = add(sigma, sqr(x.im));
      }
      const bool zero = all_zero(sigma);
      Z b;
      b.re = R<K>::zero();
      b.im = R<K>::from(zero ? 1.0 : 0.0);
      w.sc.st(2, b);
      if (zero) {
        w.sc.st(0, C<K>::zero());
        w.sc.st(1, C<K>::from(1.0, 0.0));
      } else {
        const Z x0 = w.r.ld(k * n + k);
        T sqrx0 = add(sqr(x0.re), sqr(x0.im));
        const T x0rad = R<K>::sqrt(sqrx0);
        sqrx0 = add(sqrx0, sigma);
        T mu = R<K>::sqrt(sqrx0);
        Z v0;
        if (all_zero(x0rad)) {
          v0.re = neg(mu);
          v0.im = R<K>::zero();
        } else {
          mu = div(mu, x0rad);
          v0.re = sub(x0.re, mul(mu, x0.re));
          v0.im = sub(x0.im, mul(mu, x0.im));
        }
        const T sqrv0 = add(sqr(v0.re), sqr(v0.im));
        sigma = add(sigma, sqrv0);
        T beta = div(sqrv0, sigma);
        beta = mul_d(beta, 2.0);
        Z inv;
        inv.re = div(v0.re, sqrv0);
        inv.im = neg(div(v0.im, sqrv0));
        Z bz;
        bz.re = beta;
        bz.im = R<K>::zero();
        w.sc.st(0, bz);
        w.sc.st(1, inv);
      }
    }
    wsync();
    {
      const Z inv = w.sc.ld(1);
      const bool szero = R<K>::hi(w.sc.ld(2).im) != 0.0;
      for (int i = l; i < m; i += 64) {
        if (i == 0) {
          w.v.st(0, C<K>::from(1.0, 0.0));
        } else {
          const Z x = w.r.ld((k + i) * n + k);
          if (szero) {
            w.v.st(i, x);
          } else {
            Z v;
            v.re = mul(x.re, inv.re);
            v.re = sub(v.re, mul(x.im, inv.im));
            v.im = mul(x.im, inv.re);
            v.im = add(v.im, mul(x.re, inv.im));
            w.v.st(i, v);
          }
        }
      }
    }
    wsync();
    const T beta = w.sc.ld(0).re;
    // w1[j-k] = beta * sum_i conj(R[i][j]) v[i-k]
    for (int j = k + l; j < n; j += 64) {
      Z s = C<K>::zero();
      for (int i = k; i < n; ++i) {
        const Z r = w.r.ld(i * n + j), v = w.v.ld(i - k);
        T zre = mul(r.re, v.re);
        zre = add(zre, mul(r.im, v.im));
        T zim = mul(r.re, v.im);
        zim = sub(zim, mul(r.im, v.re));
        s.re = add(s.re, zre);
        s.im = add(s.im, zim);
      }
      s.re = mul(beta, s.re);
      s.im = mul(beta, s.im);
      w.w1.st(j - k, s);
    }
    // w2[i] = beta * sum_{j>=k} Q[i][j] v[j-k]
    for (int i = l; i < n; i += 64) {
      Z s = C<K>::zero();
      for (int j = k; j < n; ++j) {
        const Z q = w.q.ld(i * n + j), v = w.v.ld(j - k);
        T zre = mul(q.re, v.re);
        zre = sub(zre, mul(q.im, v.im));
        T zim = mul(q.im, v.re);
        zim = add(zim, mul(q.re, v.im));
        s.re = add(s.re, zre);
        s.im = add(s.im, zim);
      }
      s.re = mul(beta, s.re);
      s.im = mul(beta, s.im);
      w.w2.st(i, s);
    }
    wsync();
    // R[i][j] -= v[i-k] conj(w1[j-k]);  Q[i][j] -= w2[i] conj(v[j-k])
    for (int t = l; t < m * m; t += 64) {
      const int i = k + t / m, j = k + t % m;
      const Z v = w.v.ld(i - k), x = w.w1.ld(j - k);
      Z r = w.r.ld(i * n + j);
      T zre = mul(v.re, x.re);
      zre = add(zre, mul(v.im, x.im));
      T zim = mul(v.im, x.re);
      zim = sub(zim, mul(v.re, x.im));
      r.re = sub(r.re, zre);
      r.im = sub(r.im, zim);
      w.r.st(i * n + j, r);
    }
    for (int t = l; t < n * m; t += 64) {
      const int i = t / m, j = k + t % m;
      const Z x = w.w2.ld(i), v = w.v.ld(j - k);
      Z q = w.q.ld(i * n + j);
      T zre = mul(x.re, v.re);
      zre = add(zre, mul(x.im, v.im));
      T zim = mul(x.im, v.re);
      zim = sub(zim, mul(x.re, v.im));
      q.re = sub(q.re, zre);
      q.im = sub(q.im, zim);
      w.q.st(i * n + j, q);
    }
    wsync();
  }
}

// x = R^{-1} Q^H b (CPU_cmplx2_factors_qrbs and _backward, Matrices/
// dbl2_factorizations.cpp:842-879, 142-200): b = bv[bo..], x = xv[xo..]
template <int K>
__device__ void cx_qrbs(int n, const LinWork<K, true> &w, const Arr<K, true> &bv, int bo,
                        const Arr<K, true> &xv, int xo) {
  typedef typename RealT<K>::type T;
  typedef cplx<K> Z;
  const int l = lane_id();
  for (int i = l; i < n; i += 64) {
    Z s = C<K>::zero();
    for (int j = 0; j < n; ++j) {
      const Z q = w.q.ld(j * n + i), b = bv.ld(bo + j);
      s.re = add(s.re, mul(q.re, b.re));
      s.re = add(s.re, mul(q.im, b.im));
      s.im = add(s.im, mul(q.re, b.im));
      s.im = sub(s.im, mul(q.im, b.re));
    }
    w.wrk.st(i, s);
  }
  wsync();
  for (int i = l; i < n; i += 64) xv.st(xo + i, w.wrk.ld(i));
  wsync();
  for (int j = n - 1; j >= 0; --j) {
    if (l == (j & 63)) {
      const Z u = w.r.ld(j * n + j), x = xv.ld(xo + j);
      const T den = add(mul(u.re, u.re), mul(u.im, u.im));
      const T a1re = div(u.re, den), a1im = neg(div(u.im, den));
      Z y;
      y.re = sub(mul(x.re, a1re), mul(x.im, a1im));
      y.im = add(mul(x.im, a1re), mul(x.re, a1im));
      xv.st(xo + j, y);
    }
    wsync();
    const Z xj = xv.ld(xo + j);
    for (int i = l; i < j; i += 64) {
      const Z u = w.r.ld(i * n + j);
      Z y = xv.ld(xo + i);
      const T a1re = mul(u.re, xj.re), a1im = mul(u.im, xj.im);
      const T a2re = mul(u.im, xj.re), a2im = mul(u.re, xj.im);
      y.re = sub(y.re, a1re);
      y.re = add(y.re, a1im);
      y.im = sub(y.im, a2re);
      y.im = sub(y.im, a2im);
      xv.st(xo + i, y);
    }
    wsync();
  }
}

// the real or complex factorization and solve
template <int K, bool CX, class A> __device__ void any_houseqr(int n, const A &a, const LinWork<K, CX> &w) {
  if constexpr (CX) cx_houseqr<K>(n, a, w);
  else houseqr<K>(n, a, w);
}
template <int K, bool CX>
__device__ void any_qrbs(int n, const LinWork<K, CX> &w, const Arr<K, CX> &bv, int bo,
                         const Arr<K, CX> &xv, int xo) {
  if constexpr (CX) cx_qrbs<K>(n, w, bv, bo, xv, xo);
  else qrbs<K>(n, w, bv, bo, xv, xo);
}

struct Flags {
  int zeroQ, noqr, up, bs, newtail;
};

// lower block-triangular Toeplitz solve sum_{j<=k} A_j x_{k-j} = b_k
// (CPU_dbl2_qrbs_solve with _head and _tail); mat(j, r, c) = A_j[r][c];
// rhs and sol rows k at offsets k*n.  Uniform control flow.
template <int K, bool CX, class M>
__device__ void qrbs_solve(int n, int degp1, int tailidx, const M &mat, const Arr<K, CX> &rhs,
                           const Arr<K, CX> &sol, const LinWork<K, CX> &w, Flags &f) {
  typedef typename NT<K, CX>::V T;
  const int l = lane_id();
  const double thr = threshold<K>();
  if (!(f.noqr && !f.zeroQ)) {
    const double nrm = maxnorm<K, CX>(rhs, 0, n);
    const bool small = (K == 1) ? (nrm + 1.0 == 1.0) : (nrm < thr);
    if (small) f.noqr = f.zeroQ ? 0 : 1;
    if (!f.noqr) {
      any_houseqr<K, CX>(n, [&](int r, int c) { return mat(0, r, c); }, w);
      f.zeroQ = 0;
      any_qrbs<K, CX>(n, w, rhs, 0, sol, 0);
    }
  }
  if (degp1 > 1) {
    int skipup = 0, skipbs = 0;
    bool firstbs = true;
    f.newtail = degp1;
    for (int i = tailidx; i < degp1; ++i) {
      if (maxnorm<K, CX>(sol, (i - 1) * n, n) < thr) {
        ++skipup;
      } else {
        const int npair = (degp1 - i) * n;
        for (int t = l; t < npair; t += 64) {
          const int j = i + t / n, k = t % n;
          T b = rhs.ld(j * n + k);
          for (int L = 0; L < n; ++L) b = NT<K, CX>::msub(b, mat(j - i + 1, k, L), sol.ld((i - 1) * n + L));
          rhs.st(j * n + k, b);
        }
      }
      wsync();
      const double nrm = maxnorm<K, CX>(rhs, i * n, n);
      if (nrm < thr || nrm > 1.0e+99) {
        ++skipbs;
        for (int j = l; j < n; j += 64) sol.st(i * n + j, NT<K, CX>::zero());
        wsync();
      } else {
        if (firstbs) {
          f.newtail = i;
          firstbs = false;
        }
        any_qrbs<K, CX>(n, w, rhs, i * n, sol, i * n);
      }
    }
    f.up = skipup;
    f.bs = skipbs;
  }
}

// ---------------------------------------------- polynomial system values

// the supports of a row system (shared by every system of a batch)
struct RowSys {
  int dim, nmono, nvmax, nbrmax;
  const int *monoptr, *nvr, *idxptr, *idx;
  const int *exp;  // exponents parallel to idx, nullptr when all are 1
  // A system in the column representation (the column Newton method),
  // equation j holding monomial j of every column in column order: each
  // monomial is evaluated as CPU_dbl2_evaldiff does (a two-variable
  // monomial's value as (x_a x_b) c), the coefficient of row monomial m is
  // coefficient series cmap[m] of the caller's ncff per system (unit series
  // when the coefficients are null: a monomial system), the constant is -1
  // in the last equation when negone and 0 elsewhere, and the right-hand
  // side of the linearized system is -(f - mb).
  int column, negone, ncff;
  const int *cmap;
  // a monomial system (one monomial per equation): its value and
  // derivatives are stored, not added to zero, as
  // dbl2_linearize_evaldiff_output copies them (Newton/dbl2_systems_host.cpp:
  // 455-459, 506-513); an addition to zero renormalizes a quad double
  int assign;
};

// v * e, the derivative of a monomial times its exponent, with the
// multiplication by a double of each precision (ddf_mul_d_dd and its twins,
// as CPU_dbl2_evaluate_monomials does, Newton/dbl2_systems_host.cpp:103-110)
__device__ __forceinline__ double times_exp(double v, double e) { return e * v; }
template <int K> __device__ __forceinline__ mdr<K> times_exp(const mdr<K> &v, double e) {
  return mul_d(v, e);
}

// coefficient k of x*y: x_0 y_k, then += x_i y_{k-i} (CPU_dbl2_product,
// Convolutions/dbl2_convolutions_host.cpp:7-26)
template <int K, bool CX, class FX, class FY>
__device__ __forceinline__ typename NT<K, CX>::V cprod(const FX &x, const FY &y, int k) {
  typename NT<K, CX>::V acc = NT<K, CX>::mul(x(0), y(k));
  for (int i = 1; i <= k; ++i) acc = NT<K, CX>::mac(acc, x(i), y(k - i));
  return acc;
}

// value and gradient series of every polynomial at X (working degree Dw-1):
// out row (p*(dim+1)+j) = d p / d x_j, row (p*(dim+1)+dim) = p, as
// CPU_dbl2_poly_evaldiff with CPU_dbl2_poly_speel and CPU_dbl2_speel
// (Polynomials/dbl2_polynomials_host.cpp, Convolutions/dbl2_monomials_host.cpp:28-87).
// Slot s (Dp lanes) evaluates polynomials s, s + nslot, ...; lane k of a slot
// owns coefficient k.  cst/cff planes: cst + p*D, cff + m*D with plane strides.
// A monomial's coefficient is staged in the slot's row CF; with exponents
// it is first multiplied by the common factor prod_j x_j^(e_j - 1), one
// factor at a time in the order of idx (the first loop of
// CPU_dbl2_evaluate_monomials, Newton/dbl2_systems_host.cpp:18-45), and each
// derivative contribution is multiplied by its exponent before it is added.
// (complex: the products of CPU_cmplx2_poly_evaldiff; no exponents)
template <int K, bool CX>
__device__ void eval_system(const RowSys &S, int D, int Dw, int Dp, const double *cst,
                            size_t cstride, const double *cff, size_t fstride,
                            const Arr<K, CX> &X, const Arr<K, CX> &OUT, const Arr<K, CX> &FW,
                            const Arr<K, CX> &BW, const Arr<K, CX> &CR, const Arr<K, CX> &CF) {
  typedef NT<K, CX> N;
  typedef typename N::V T;
  const int l = lane_id(), nslot = 64 / Dp, s = l / Dp, k = l % Dp;
  const int dim = S.dim, D1 = dim + 1;
  const int eo = s * S.nvmax * D;  // this slot's work rows
  const int co = s * D;            // this slot's coefficient row
  auto xin = [&](int var) { return [&X, var, D](int j) { return X.ld(var * D + j); }; };
  auto fw = [&](int r) { return [&FW, eo, r, D](int j) { return FW.ld(eo + r * D + j); }; };
  auto bw = [&](int r) { return [&BW, eo, r, D](int j) { return BW.ld(eo + r * D + j); }; };
  auto cf = [&CF, co](int j) { return CF.ld(co + j); };
  for (int r0 = 0; r0 < dim; r0 += nslot) {
    const int p = r0 + s;
    const bool on = (p < dim) && (k < Dw);
    if (on) {
      if constexpr (CX) {
        OUT.st((p * D1 + dim) * D + k, N::ld(cst, cstride, (size_t)p * D + k));
      } else {
        // column systems (cst null): 0, or -1 in the last equation
        T c0 = N::zero();
        if (cst) c0 = N::ld(cst, cstride, (size_t)p * D + k);
        else if (S.negone && p == dim - 1 && k == 0) c0 = N::neg(N::one());
        OUT.st((p * D1 + dim) * D + k, c0);
      }
      for (int j = 0; j < dim; ++j) OUT.st((p * D1 + j) * D + k, N::zero());
    }
    const int nb = (p < dim) ? S.monoptr[p + 1] - S.monoptr[p] : 0;
    for (int mi = 0; mi < S.nbrmax; ++mi) {
      const bool act = on && mi < nb;
      const int mono = act ? S.monoptr[p] + mi : 0;
      const int nv = act ? S.nvr[mono] : 0;
      const int *ix = S.idx + (act ? S.idxptr[mono] : 0);
      const int *ex = (!CX && act && S.exp) ? S.exp + S.idxptr[mono] : nullptr;
      // v times the exponent of variable position j (positions with
      // exponent 1 are added as they are)
      auto addto = [&](int row, const T &v, int j) {
        const int o = (p * D1 + row) * D + k;
        if constexpr (CX) {
          OUT.st(o, N::add(OUT.ld(o), v));
        } else {
          const int e = (ex && j >= 0) ? ex[j] : 1;
          const T ve = e > 1 ? times_exp(v, (double)e) : v;
          OUT.st(o, S.assign ? ve : add(OUT.ld(o), ve));
        }
      };
      if constexpr (CX) {
        if (act) CF.st(co + k, N::ld(cff, fstride, (size_t)mono * D + k));
      } else if (act) {
        const size_t cm = S.cmap ? (size_t)S.cmap[mono] : (size_t)mono;
        CF.st(co + k, cff ? N::ld(cff, fstride, cm * D + k) : (k == 0 ? N::one() : N::zero()));
      }
      wsync();
      if (ex) {
        for (int j = 0; j < nv; ++j)
          for (int e = 1; e < ex[j]; ++e) {
            const T a = cprod<K, CX>(xin(ix[j]), cf, k);
            wsync();
            CF.st(co + k, a);
            wsync();
          }
      }
      if (nv == 1) {
        const T f = cprod<K, CX>(xin(ix[0]), cf, k);
        addto(dim, f, -1);
        addto(ix[0], cf(k), 0);
      } else if (nv == 2) {
        // the value: (c x_a) x_b in a polynomial, (x_a x_b) c in a column
        // system (real only)
        bool col = false;
        if constexpr (!CX) col = S.column != 0;
        const T f0 = cprod<K, CX>(cf, xin(ix[0]), k);
        if (col) FW.st(eo + k, cprod<K, CX>(xin(ix[0]), xin(ix[1]), k));
        else FW.st(eo + k, f0);
        addto(ix[1], f0, 1);
        const T b0 = cprod<K, CX>(cf, xin(ix[1]), k);
        addto(ix[0], b0, 0);
        wsync();
        if (col) addto(dim, cprod<K, CX>(fw(0), cf, k), -1);
        else addto(dim, cprod<K, CX>(fw(0), xin(ix[1]), k), -1);
      } else if (nv > 2) {
        // forward and backward chains advance together
        FW.st(eo + k, cprod<K, CX>(cf, xin(ix[0]), k));
        BW.st(eo + k, cprod<K, CX>(xin(ix[nv - 1]), xin(ix[nv - 2]), k));
        wsync();
        for (int j = 1; j < nv; ++j) {
          FW.st(eo + j * D + k, cprod<K, CX>(fw(j - 1), xin(ix[j]), k));
          if (j < nv - 2) BW.st(eo + j * D + k, cprod<K, CX>(bw(j - 1), xin(ix[nv - 2 - j]), k));
          wsync();
        }
        const T blast = cprod<K, CX>(bw(nv - 3), cf, k);
        wsync();
        BW.st(eo + (nv - 3) * D + k, blast);
        wsync();
        if (nv == 3) {
          CR.st(eo + k, cprod<K, CX>(fw(0), xin(ix[2]), k));
        } else {
          for (int j = 0; j < nv - 3; ++j) CR.st(eo + j * D + k, cprod<K, CX>(fw(j), bw(nv - 4 - j), k));
          CR.st(eo + (nv - 3) * D + k, cprod<K, CX>(fw(nv - 3), xin(ix[nv - 1]), k));
        }
        addto(dim, FW.ld(eo + (nv - 1) * D + k), -1);
        addto(ix[nv - 1], FW.ld(eo + (nv - 2) * D + k), nv - 1);
        addto(ix[0], BW.ld(eo + (nv - 3) * D + k), 0);
        for (int j = 1; j < nv - 1; ++j) addto(ix[j], CR.ld(eo + (j - 1) * D + k), j);
      }
      wsync();
    }
  }
  wsync();
}

// ------------------------------------------------------------ kernels

struct NewtonLayout {
  int X, OUT, RHS, SOL, Q, R, V, W1, W2, WRK, SC, FW, BW, CR, CF;  // sizes in numbers
  // P doubles per number: K (real) or 2K (complex)
  __host__ __device__ size_t bytes(int P) const {
    return sizeof(double) * (size_t)P *
           ((size_t)X + OUT + RHS + SOL + Q + R + V + W1 + W2 + WRK + SC + FW + BW + CR + CF);
  }
};

__host__ __device__ inline int pow2_at_least(int d) {
  int p = 1;
  while (p < d) p <<= 1;
  return p;
}

__host__ __device__ inline NewtonLayout newton_layout(const RowSys &S, int D) {
  const int n = S.dim, nslot = 64 / pow2_at_least(D), ev = nslot * (S.nvmax > 0 ? S.nvmax : 1) * D;
  NewtonLayout L;
  L.X = n * D; L.OUT = n * (n + 1) * D; L.RHS = D * n; L.SOL = D * n;
  L.Q = n * n; L.R = n * n; L.V = n; L.W1 = n; L.W2 = n; L.WRK = n; L.SC = 3;
  L.FW = ev; L.BW = ev; L.CR = ev; L.CF = nslot * D;
  return L;
}

template <int K, bool CX = false>
__device__ LinWork<K, CX> carve_lin(double *&p, int n) {
  LinWork<K, CX> w;
  w.q = carve<K, CX>(p, n * n); w.r = carve<K, CX>(p, n * n); w.v = carve<K, CX>(p, n);
  w.w1 = carve<K, CX>(p, n); w.w2 = carve<K, CX>(p, n); w.wrk = carve<K, CX>(p, n);
  w.sc = carve<K, CX>(p, 3);
  return w;
}

// x [K][nsys][dim][D] in/out; cst [K][nc][dim][D], cff [K][nf][ncff][D] with
// nc, nf = nsys or 1 (shared); logs [nsys][nbsteps]; hist
// [nbsteps][K][nsys][dim][D]; Q, R [K][nsys][dim][dim].  Column systems:
// cst null, cff null for unit coefficients, mb [K][nm][dim][D] (nm = nsys
// or 1); rows: mb null.
template <int K, bool CX>
__global__ void __launch_bounds__(512)
row_newton_kernel(RowSys S, int deg, int nbsteps, int nsys, const double *cst, int cshared,
                  const double *cff, int fshared, const double *mb, int mshared, double *x,
                  int *steps_out, int *wdeg_out, int *tail_out, int *up_out, int *bs_out,
                  double *hist, double *Qo, double *Ro) {
  typedef NT<K, CX> N;
  extern __shared__ double lds[];
  const int D = deg + 1, n = S.dim, Dp = pow2_at_least(D);
  const NewtonLayout L = newton_layout(S, D);
  const int wave = threadIdx.x >> 6, l = lane_id();
  const int sysid = blockIdx.x * (blockDim.x >> 6) + wave;
  if (sysid >= nsys) return;
  double *p = lds + (size_t)wave * (L.bytes(N::P) / sizeof(double));
  const Arr<K, CX> X = carve<K, CX>(p, L.X), OUT = carve<K, CX>(p, L.OUT),
                   RHS = carve<K, CX>(p, L.RHS), SOL = carve<K, CX>(p, L.SOL);
  const LinWork<K, CX> w = carve_lin<K, CX>(p, n);
  const Arr<K, CX> FW = carve<K, CX>(p, L.FW), BW = carve<K, CX>(p, L.BW),
                   CR = carve<K, CX>(p, L.CR), CF = carve<K, CX>(p, L.CF);
  const size_t xs = (size_t)nsys * n * D, xb = (size_t)sysid * n * D;
  const size_t cs = (size_t)(cshared ? 1 : nsys) * n * D;
  const size_t fs = (size_t)(fshared ? 1 : nsys) * S.ncff * D;
  const size_t ms = (size_t)(mshared ? 1 : nsys) * n * D, mo = mshared ? 0 : (size_t)sysid * n * D;
  const double *cbase = cst ? cst + (cshared ? 0 : (size_t)sysid * n * D) : nullptr;
  const double *fbase = cff ? cff + (fshared ? 0 : (size_t)sysid * S.ncff * D) : nullptr;
  for (int t = l; t < n * D; t += 64) X.st(t, N::ld(x, xs, xb + t));
  wsync();
  const int D1 = n + 1;
  auto jac = [&](int j, int r, int c) { return OUT.ld((r * D1 + c) * D + j); };
  Flags f{1, 0, 0, 0, 1};
  int tailidx = 1, wrkdeg = 0, stepcnt = 0;
  for (int step = 0; step < nbsteps; ++step) {
    const int Dw = wrkdeg + 1;
    eval_system<K, CX>(S, D, Dw, Dp, cbase, cs, fbase, fs, X, OUT, FW, BW, CR, CF);
    for (int t = l; t < Dw * n; t += 64) {
      const int kk = t / n, i = t % n;
      const typename N::V fv = OUT.ld((i * D1 + n) * D + kk);
      if constexpr (CX) {
        RHS.st(t, N::neg(fv));
      } else {
        // -(f - mb) as dbl2_define_rhs (Newton/dbl2_systems_host.cpp:626-659)
        RHS.st(t, mb ? N::neg(md::sub(fv, N::ld(mb, ms, mo + (size_t)i * D + kk))) : N::neg(fv));
      }
      SOL.st(t, N::zero());
    }
    wsync();
    f.newtail = tailidx;
    qrbs_solve<K, CX>(n, Dw, tailidx, jac, RHS, SOL, w, f);
    tailidx = f.newtail;
    for (int t = l; t < n * Dw; t += 64) {
      const int i = t / Dw, kk = t % Dw;
      if (kk >= tailidx - 1) X.st(i * D + kk, N::add(X.ld(i * D + kk), SOL.ld(kk * n + i)));
    }
    wsync();
    if (l == 0) {
      const size_t o = (size_t)sysid * nbsteps + step;
      if (wdeg_out) wdeg_out[o] = wrkdeg;
      if (tail_out) tail_out[o] = tailidx;
      if (up_out) up_out[o] = f.up;
      if (bs_out) bs_out[o] = f.bs;
    }
    if (hist) {
      double *h = hist + (size_t)step * N::P * xs;
      for (int t = l; t < n * D; t += 64) N::st(h, xs, xb + t, X.ld(t));
    }
    ++stepcnt;
    if (tailidx >= deg) break;
    wrkdeg = wrkdeg + 1 + wrkdeg / 2;
    if (wrkdeg > deg) wrkdeg = deg;
  }
  for (int t = l; t < n * D; t += 64) N::st(x, xs, xb + t, X.ld(t));
  const size_t qs = (size_t)nsys * n * n, qb = (size_t)sysid * n * n;
  for (int t = l; t < n * n; t += 64) {
    if (Qo) N::st(Qo, qs, qb + t, w.q.ld(t));
    if (Ro) N::st(Ro, qs, qb + t, w.r.ld(t));
  }
  if (l == 0 && steps_out) steps_out[sysid] = stepcnt;
}

// batched Householder QR of n x n matrices A [K][nb][n][n] -> Q, R
template <int K, bool CX>
__global__ void __launch_bounds__(512)
houseqr_kernel(int n, int nb, const double *A, double *Qo, double *Ro) {
  typedef NT<K, CX> N;
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6, l = lane_id();
  const int b = blockIdx.x * (blockDim.x >> 6) + wave;
  if (b >= nb) return;
  const size_t per = (size_t)N::P * (2 * n * n + 4 * n + 3);
  double *p = lds + (size_t)wave * per;
  const LinWork<K, CX> w = carve_lin<K, CX>(p, n);
  const size_t s = (size_t)nb * n * n, o = (size_t)b * n * n;
  any_houseqr<K, CX>(n, [&](int r, int c) { return N::ld(A, s, o + (size_t)r * n + c); }, w);
  for (int t = l; t < n * n; t += 64) {
    N::st(Qo, s, o + t, w.q.ld(t));
    N::st(Ro, s, o + t, w.r.ld(t));
  }
}

// batched Toeplitz solve: mat [K][nb][degp1][n][n], rhs/sol [K][nb][degp1][n]
// (in/out), Q/R [K][nb][n][n] (in/out), flags [nb][5] (zeroQ, noqr, up, bs,
// newtail; in/out)
template <int K, bool CX>
__global__ void __launch_bounds__(512)
qrbs_solve_kernel(int n, int degp1, int tailidx, int nb, const double *mat, double *rhs,
                  double *sol, double *Qio, double *Rio, int *flags) {
  typedef NT<K, CX> N;
  extern __shared__ double lds[];
  const int wave = threadIdx.x >> 6, l = lane_id();
  const int b = blockIdx.x * (blockDim.x >> 6) + wave;
  if (b >= nb) return;
  const size_t per = (size_t)N::P * (2 * n * n + 4 * n + 3 + 2 * degp1 * n);
  double *p = lds + (size_t)wave * per;
  const LinWork<K, CX> w = carve_lin<K, CX>(p, n);
  const Arr<K, CX> RHS = carve<K, CX>(p, degp1 * n), SOL = carve<K, CX>(p, degp1 * n);
  const size_t vs = (size_t)nb * degp1 * n, vo = (size_t)b * degp1 * n;
  const size_t qs = (size_t)nb * n * n, qo = (size_t)b * n * n;
  const size_t ms = (size_t)nb * degp1 * n * n, mo = (size_t)b * degp1 * n * n;
  for (int t = l; t < degp1 * n; t += 64) {
    RHS.st(t, N::ld(rhs, vs, vo + t));
    SOL.st(t, N::ld(sol, vs, vo + t));
  }
  for (int t = l; t < n * n; t += 64) {
    w.q.st(t, N::ld(Qio, qs, qo + t));
    w.r.st(t, N::ld(Rio, qs, qo + t));
  }
  wsync();
  Flags f{flags[b * 5], flags[b * 5 + 1], flags[b * 5 + 2], flags[b * 5 + 3], flags[b * 5 + 4]};
  auto m = [&](int j, int r, int c) { return N::ld(mat, ms, mo + ((size_t)j * n + r) * n + c); };
  qrbs_solve<K, CX>(n, degp1, tailidx, m, RHS, SOL, w, f);
  for (int t = l; t < degp1 * n; t += 64) {
    N::st(rhs, vs, vo + t, RHS.ld(t));
    N::st(sol, vs, vo + t, SOL.ld(t));
  }
  for (int t = l; t < n * n; t += 64) {
    N::st(Qio, qs, qo + t, w.q.ld(t));
    N::st(Rio, qs, qo + t, w.r.ld(t));
  }
  if (l == 0) {
    flags[b * 5] = f.zeroQ; flags[b * 5 + 1] = f.noqr; flags[b * 5 + 2] = f.up;
    flags[b * 5 + 3] = f.bs; flags[b * 5 + 4] = f.newtail;
  }
}

// r_i = b_i - sum_{j <= i-tailidx} A_j x_{i-j} for i >= tailidx
// (CPU_dbl2_linear_residue, Newton/dbl2_bals_host.cpp:548-600); one lane per
// residual entry; resmax [K][nb] = entry of largest |leading limb|, the first
// such entry in the reference's scan order.
template <int K, bool CX>
__global__ void __launch_bounds__(512)
linear_residue_kernel(int n, int degp1, int tailidx, int nb, const double *mat,
                      const double *rhs, const double *sol, double *res, double *resmax) {
  typedef NT<K, CX> N;
  typedef typename N::V T;
  const int wave = threadIdx.x >> 6, l = lane_id();
  const int b = blockIdx.x * (blockDim.x >> 6) + wave;
  if (b >= nb) return;
  const size_t vs = (size_t)nb * degp1 * n, vo = (size_t)b * degp1 * n;
  const size_t ms = (size_t)nb * degp1 * n * n, mo = (size_t)b * degp1 * n * n;
  double best = 0.0;
  int bidx = -1;
  for (int t = tailidx * n + l; t < degp1 * n; t += 64) {
    const int i = t / n, L = t % n;
    T r = N::ld(rhs, vs, vo + t);
    for (int j = 0; j <= i - tailidx; ++j)
      for (int k = 0; k < n; ++k)
        r = N::msub(r, N::ld(mat, ms, mo + ((size_t)j * n + L) * n + k),
                    N::ld(sol, vs, vo + (size_t)(i - j) * n + k));
    N::st(res, vs, vo + t, r);
    const double a = N::mag(r);
    if (a > best) { best = a; bidx = t; }
  }
  // the scan keeps the first strictly larger entry: largest value, then the
  // smallest index among equal values
  for (int m = 1; m < 64; m <<= 1) {
    const double ob = __shfl_xor(best, m, 64);
    const int oi = __shfl_xor(bidx, m, 64);
    if (ob > best || (ob == best && oi >= 0 && (bidx < 0 || oi < bidx))) { best = ob; bidx = oi; }
  }
  if (l == 0) {
    T r = N::zero();
    if (bidx >= 0) r = N::ld(res, vs, vo + bidx);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      // complex: |re| + |im| limb by limb, as CPU_cmplx2_linear_residue
      if constexpr (CX) resmax[(size_t)i * nb + b] = fabs(limb(r.re, i)) + fabs(limb(r.im, i));
      else resmax[(size_t)i * nb + b] = fabs(limb(r, i));
    }
  }
}

int waves_for(size_t per_wave, int *waves) {
  int dev = 0, ldsmax = 0;
  PHC_HIP(hipGetDevice(&dev));
  PHC_HIP(hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t cap = (size_t)ldsmax < LDS_LIMIT ? (size_t)ldsmax : LDS_LIMIT;
  PHC_REQUIRE(per_wave <= cap, "problem too large for the LDS of one workgroup");
  int wv = (int)(cap / per_wave);
  if (wv > 8) wv = 8;
  *waves = wv;
  return 0;
}

}  // namespace

struct phc_rowsys {
  RowSys S;
  int *dmem;
  int *cmem;   // column systems: the coefficient map
  int nbrcol;  // column systems: the number of columns (0 for row systems)
};

extern "C" {

int phc_rowsys_create_exp(phc_rowsys **out, int dim, const int *monoptr, const int *nvr,
                          const int *idxptr, const int *idx, const int *exp);
int phc_rowsys_destroy(phc_rowsys *s);

int phc_rowsys_create(phc_rowsys **out, int dim, const int *monoptr, const int *nvr,
                      const int *idxptr, const int *idx) {
  return phc_rowsys_create_exp(out, dim, monoptr, nvr, idxptr, idx, nullptr);
}

int phc_rowsys_create_exp(phc_rowsys **out, int dim, const int *monoptr, const int *nvr,
                          const int *idxptr, const int *idx, const int *exp) {
  PHC_REQUIRE(out && monoptr && nvr && idxptr, "phc_rowsys_create: null argument");
  PHC_REQUIRE(dim >= 1 && dim <= 64, "phc_rowsys_create: dim must be in 1..64");
  const int nmono = monoptr[dim];
  PHC_REQUIRE(monoptr[0] == 0 && nmono >= 0, "phc_rowsys_create: bad monomial pointers");
  int nvmax = 0, nbrmax = 0;
  for (int i = 0; i < dim; ++i) {
    PHC_REQUIRE(monoptr[i + 1] >= monoptr[i], "phc_rowsys_create: bad monomial pointers");
    if (monoptr[i + 1] - monoptr[i] > nbrmax) nbrmax = monoptr[i + 1] - monoptr[i];
  }
  PHC_REQUIRE(idxptr[0] == 0, "phc_rowsys_create: bad index pointers");
  for (int m = 0; m < nmono; ++m) {
    PHC_REQUIRE(nvr[m] >= 1 && nvr[m] <= dim && idxptr[m + 1] - idxptr[m] == nvr[m],
                "phc_rowsys_create: monomial variable counts and index pointers disagree");
    for (int j = idxptr[m]; j < idxptr[m + 1]; ++j)
      PHC_REQUIRE(idx && idx[j] >= 0 && idx[j] < dim, "phc_rowsys_create: variable index out of range");
    if (nvr[m] > nvmax) nvmax = nvr[m];
  }
  const int nidx = idxptr[nmono];
  bool any_exp = false;
  for (int j = 0; exp && j < nidx; ++j) {
    PHC_REQUIRE(exp[j] >= 1 && exp[j] <= 64, "phc_rowsys_create: exponents must be in 1..64");
    any_exp = any_exp || exp[j] > 1;
  }
  const size_t ints = (size_t)(dim + 1) + 2 * (size_t)nmono + 1 + nidx + (any_exp ? nidx : 0);
  phc_rowsys *s = new phc_rowsys();
  if (hipMalloc(&s->dmem, sizeof(int) * (ints > 0 ? ints : 1)) != hipSuccess) {
    delete s;
    phc::set_error("phc_rowsys_create: hipMalloc failed");
    return -1;
  }
  int *p = s->dmem;
  std::vector<int> h;
  h.insert(h.end(), monoptr, monoptr + dim + 1);
  h.insert(h.end(), nvr, nvr + nmono);
  h.insert(h.end(), idxptr, idxptr + nmono + 1);
  if (nidx > 0) h.insert(h.end(), idx, idx + nidx);
  if (any_exp) h.insert(h.end(), exp, exp + nidx);
  PHC_HIP(hipMemcpy(p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
  s->S.dim = dim; s->S.nmono = nmono; s->S.nvmax = nvmax; s->S.nbrmax = nbrmax;
  s->S.monoptr = p; s->S.nvr = p + dim + 1; s->S.idxptr = p + dim + 1 + nmono;
  s->S.idx = p + dim + 1 + 2 * nmono + 1;
  s->S.exp = any_exp ? s->S.idx + nidx : nullptr;
  s->S.ncff = nmono;
  *out = s;
  return 0;
}

int phc_colsys_create(phc_rowsys **out, int dim, int nbrcol, const int *nvr, const int *idx,
                      const int *exps) {
  PHC_REQUIRE(out && nvr, "phc_colsys_create: null argument");
  PHC_REQUIRE(dim >= 1 && dim <= 64, "phc_colsys_create: dim must be in 1..64");
  PHC_REQUIRE(nbrcol >= 1 && nbrcol <= 64, "phc_colsys_create: nbrcol must be in 1..64");
  PHC_REQUIRE(nbrcol == 1 || exps == nullptr,
              "phc_colsys_create: exponents are for monomial systems (nbrcol = 1) only");
  const int ncff = nbrcol * dim;
  std::vector<int> iptr(ncff + 1, 0);
  for (int m = 0; m < ncff; ++m) {
    PHC_REQUIRE(nvr[m] >= (nbrcol == 1 ? 1 : 0) && nvr[m] <= dim,
                "phc_colsys_create: monomial variable count out of range");
    iptr[m + 1] = iptr[m] + nvr[m];
  }
  // equation j: monomial j of every column in column order (the order in
  // which CPU_dbl2_evaluate_columns adds them, Newton/dbl2_systems_host.cpp:
  // 291-323), empty monomials left out
  std::vector<int> monoptr(dim + 1, 0), rnvr, ridxptr(1, 0), ridx, rexp, cmap;
  for (int j = 0; j < dim; ++j) {
    for (int c = 0; c < nbrcol; ++c) {
      const int m = c * dim + j;
      if (nvr[m] == 0) continue;
      rnvr.push_back(nvr[m]);
      cmap.push_back(m);
      for (int k = iptr[m]; k < iptr[m + 1]; ++k) {
        PHC_REQUIRE(idx, "phc_colsys_create: null index array");
        ridx.push_back(idx[k]);
        if (exps) rexp.push_back(exps[k]);
      }
      ridxptr.push_back((int)ridx.size());
    }
    monoptr[j + 1] = (int)rnvr.size();
  }
  for (size_t m = 0; m < rnvr.size(); ++m)
    for (int a = ridxptr[m]; a < ridxptr[m + 1]; ++a)
      for (int b = a + 1; b < ridxptr[m + 1]; ++b)
        PHC_REQUIRE(ridx[a] != ridx[b], "phc_colsys_create: repeated variable in a monomial");
  phc_rowsys *s = nullptr;
  const int rc = phc_rowsys_create_exp(&s, dim, monoptr.data(), rnvr.data(), ridxptr.data(),
                                       ridx.data(), exps ? rexp.data() : nullptr);
  if (rc != 0) return rc;
  int *dmap = nullptr;
  if (!cmap.empty() && hipMalloc(&dmap, sizeof(int) * cmap.size()) != hipSuccess) {
    phc_rowsys_destroy(s);
    phc::set_error("phc_colsys_create: hipMalloc failed");
    return -1;
  }
  if (dmap) PHC_HIP(hipMemcpy(dmap, cmap.data(), sizeof(int) * cmap.size(), hipMemcpyHostToDevice));
  s->cmem = dmap;
  s->nbrcol = nbrcol;
  s->S.column = 1;
  s->S.negone = nbrcol > 1;
  s->S.assign = nbrcol == 1;
  s->S.ncff = ncff;
  s->S.cmap = dmap;
  *out = s;
  return 0;
}

int phc_rowsys_destroy(phc_rowsys *s) {
  if (!s) return 0;
  hipError_t e = hipFree(s->dmem);
  if (s->cmem) {
    const hipError_t e2 = hipFree(s->cmem);
    if (e == hipSuccess) e = e2;
  }
  delete s;
  if (e != hipSuccess) {
    phc::set_error("phc_rowsys_destroy: hipFree failed");
    return -1;
  }
  return 0;
}

int phc_rowsys_lds_bytes(const phc_rowsys *s, int K, int deg, size_t *bytes) {
  PHC_REQUIRE(s && bytes, "phc_rowsys_lds_bytes: null argument");
  PHC_REQUIRE(deg >= 0 && deg <= 63, "phc_rowsys_lds_bytes: deg must be in 0..63");
  *bytes = newton_layout(s->S, deg + 1).bytes(K);
  return 0;
}

}  // extern "C"

namespace {

// the staggered row Newton, real (P = K planes) or complex (2K)
template <bool CX>
int row_newton(phc_rowsys *s, int K, int deg, int nbsteps, int nsys, const double *cst,
               int cst_shared, const double *cff, int cff_shared, const double *mb,
               int mb_shared, double *x, int *steps, int *wdeg, int *tail, int *up, int *bs,
               double *history, double *Q, double *R, void *stream, const char *fn) {
  PHC_REQUIRE(s, std::string(fn) + ": null system");
  const bool column = s->S.column != 0;
  PHC_REQUIRE(!column || mb, std::string(fn) + ": a column system goes to phc_column_newton "
                                               "with its right-hand side");
  if (CX) {
    PHC_REQUIRE(K == 2 || K == 4, std::string(fn) + ": K must be 2 or 4");
    PHC_REQUIRE(s->S.exp == nullptr, std::string(fn) + ": exponents above 1 are real only");
  } else {
    PHC_REQUIRE(K == 1 || K == 2 || K == 4, std::string(fn) + ": K must be 1, 2 or 4");
  }
  PHC_REQUIRE(deg >= 0 && deg <= 63, std::string(fn) + ": deg must be in 0..63");
  PHC_REQUIRE(nbsteps >= 0 && nsys >= 0, std::string(fn) + ": negative size");
  if (column) {
    PHC_REQUIRE(!CX, std::string(fn) + ": column systems are real");
    PHC_REQUIRE(x && (cff || s->nbrcol == 1),
                std::string(fn) + ": null coefficient or series array");
  } else {
    PHC_REQUIRE(cst && cff && x, std::string(fn) + ": null coefficient or series array");
  }
  if (nsys == 0 || nbsteps == 0) return 0;
  const size_t per = newton_layout(s->S, deg + 1).bytes(CX ? 2 * K : K);
  int wv = 0;
  if (waves_for(per, &wv) != 0) return -2;
  const int grid = (nsys + wv - 1) / wv;
  const size_t lds = per * wv;
  hipStream_t st = (hipStream_t)stream;
#define LAUNCH(KK)                                                                            \
  do {                                                                                        \
    PHC_HIP(hipFuncSetAttribute((const void *)row_newton_kernel<KK, CX>,                     \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));       \
    hipLaunchKernelGGL((row_newton_kernel<KK, CX>), dim3(grid), dim3(64 * wv), lds, st, s->S, \
                       deg, nbsteps, nsys, cst, cst_shared, cff, cff_shared, mb, mb_shared, \
                       x, steps, wdeg, tail, up, bs, history, Q, R);                          \
  } while (0)
  if (K == 1 && !CX) LAUNCH(1);
  else if (K == 2) LAUNCH(2);
  else LAUNCH(4);
#undef LAUNCH
  PHC_HIP(hipGetLastError());
  return 0;
}

template <bool CX>
int houseqr_batch(int K, int n, int nbatch, const double *A, double *Q, double *R, void *stream,
                  const char *fn) {
  // complex octo-double (K = 8) is refused: its first device run stalled
  // (a launch of houseqr_kernel<8, true> that did not return), cause not
  // established
  PHC_REQUIRE(CX ? (K == 2 || K == 4) : (K == 1 || K == 2 || K == 4),
              std::string(fn) + (CX ? ": K must be 2 or 4" : ": K must be 1, 2 or 4"));
  PHC_REQUIRE(n >= 1 && n <= 64 && nbatch >= 0, std::string(fn) + ": n must be in 1..64");
  PHC_REQUIRE(A && Q && R, std::string(fn) + ": null array");
  if (nbatch == 0) return 0;
  const size_t per = sizeof(double) * (CX ? 2 * K : K) * (2 * (size_t)n * n + 4 * n + 3);
  int wv = 0;
  if (waves_for(per, &wv) != 0) return -2;
  const int grid = (nbatch + wv - 1) / wv;
  hipStream_t st = (hipStream_t)stream;
#define LAUNCH(KK)                                                                            \
  do {                                                                                        \
    PHC_HIP(hipFuncSetAttribute((const void *)houseqr_kernel<KK, CX>,                        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(per * wv))); \
    hipLaunchKernelGGL((houseqr_kernel<KK, CX>), dim3(grid), dim3(64 * wv), per * wv, st, n,  \
                       nbatch, A, Q, R);                                                      \
  } while (0)
  if (K == 1 && !CX) LAUNCH(1);
  else if (K == 2) LAUNCH(2);
  else LAUNCH(4);
#undef LAUNCH
  PHC_HIP(hipGetLastError());
  return 0;
}

template <bool CX>
int qrbs_solve_batch(int K, int n, int degp1, int tailidx, int nbatch, const double *mat,
                     double *rhs, double *sol, double *Q, double *R, int *flags, void *stream,
                     const char *fn) {
  // complex octo-double (K = 8) is refused: its first device run stalled
  // (a launch of houseqr_kernel<8, true> that did not return), cause not
  // established
  PHC_REQUIRE(CX ? (K == 2 || K == 4) : (K == 1 || K == 2 || K == 4),
              std::string(fn) + (CX ? ": K must be 2 or 4" : ": K must be 1, 2 or 4"));
  PHC_REQUIRE(n >= 1 && n <= 64 && degp1 >= 1 && degp1 <= 64 && nbatch >= 0,
              std::string(fn) + ": n and degp1 must be in 1..64");
  PHC_REQUIRE(tailidx >= 1, std::string(fn) + ": tailidx must be at least 1");
  PHC_REQUIRE(mat && rhs && sol && Q && R && flags, std::string(fn) + ": null array");
  if (nbatch == 0) return 0;
  const size_t per =
      sizeof(double) * (CX ? 2 * K : K) * (2 * (size_t)n * n + 4 * n + 3 + 2 * (size_t)degp1 * n);
  int wv = 0;
  if (waves_for(per, &wv) != 0) return -2;
  const int grid = (nbatch + wv - 1) / wv;
  hipStream_t st = (hipStream_t)stream;
#define LAUNCH(KK)                                                                            \
  do {                                                                                        \
    PHC_HIP(hipFuncSetAttribute((const void *)qrbs_solve_kernel<KK, CX>,                     \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)(per * wv))); \
    hipLaunchKernelGGL((qrbs_solve_kernel<KK, CX>), dim3(grid), dim3(64 * wv), per * wv, st,  \
                       n, degp1, tailidx, nbatch, mat, rhs, sol, Q, R, flags);                \
  } while (0)
  if (K == 1 && !CX) LAUNCH(1);
  else if (K == 2) LAUNCH(2);
  else LAUNCH(4);
#undef LAUNCH
  PHC_HIP(hipGetLastError());
  return 0;
}

template <bool CX>
int linear_residue_batch(int K, int n, int degp1, int tailidx, int nbatch, const double *mat,
                         const double *rhs, const double *sol, double *res, double *resmax,
                         void *stream, const char *fn) {
  // complex octo-double (K = 8) is refused: its first device run stalled
  // (a launch of houseqr_kernel<8, true> that did not return), cause not
  // established
  PHC_REQUIRE(CX ? (K == 2 || K == 4) : (K == 1 || K == 2 || K == 4),
              std::string(fn) + (CX ? ": K must be 2 or 4" : ": K must be 1, 2 or 4"));
  PHC_REQUIRE(n >= 1 && degp1 >= 1 && nbatch >= 0 && tailidx >= 0,
              std::string(fn) + ": bad sizes");
  PHC_REQUIRE(mat && rhs && sol && res && resmax, std::string(fn) + ": null array");
  if (nbatch == 0) return 0;
  const int wv = 4, grid = (nbatch + wv - 1) / wv;
  hipStream_t st = (hipStream_t)stream;
#define LAUNCH(KK)                                                                            \
  hipLaunchKernelGGL((linear_residue_kernel<KK, CX>), dim3(grid), dim3(64 * wv), 0, st, n,    \
                     degp1, tailidx, nbatch, mat, rhs, sol, res, resmax)
  if (K == 1 && !CX) LAUNCH(1);
  else if (K == 2) LAUNCH(2);
  else LAUNCH(4);
#undef LAUNCH
  PHC_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int phc_row_newton(phc_rowsys *s, int K, int deg, int nbsteps, int nsys, const double *cst,
                   int cst_shared, const double *cff, int cff_shared, double *x, int *steps,
                   int *wdeg, int *tail, int *up, int *bs, double *history, double *Q, double *R,
                   void *stream) {
  return row_newton<false>(s, K, deg, nbsteps, nsys, cst, cst_shared, cff, cff_shared, nullptr,
                           0, x, steps, wdeg, tail, up, bs, history, Q, R, stream,
                           "phc_row_newton");
}

int phc_column_newton(phc_rowsys *s, int K, int deg, int nbsteps, int nsys, const double *cff,
                      int cff_shared, const double *mb, int mb_shared, double *x, int *steps,
                      int *wdeg, int *tail, int *up, int *bs, double *history, double *Q,
                      double *R, void *stream) {
  PHC_REQUIRE(s && s->S.column, "phc_column_newton: not a column system (phc_colsys_create)");
  PHC_REQUIRE(mb, "phc_column_newton: null right-hand side");
  // a monomial system's coefficients are the unit series
  // (dbl2_unit_series_vector at each step, Newton/dbl2_newton_method.cpp:292)
  return row_newton<false>(s, K, deg, nbsteps, nsys, nullptr, 0, s->nbrcol == 1 ? nullptr : cff,
                           cff_shared, mb, mb_shared, x, steps, wdeg, tail, up, bs, history, Q, R,
                           stream, "phc_column_newton");
}

int phc_cmplx_row_newton(phc_rowsys *s, int K, int deg, int nbsteps, int nsys, const double *cst,
                         int cst_shared, const double *cff, int cff_shared, double *x,
                         int *steps, int *wdeg, int *tail, int *up, int *bs, double *history,
                         double *Q, double *R, void *stream) {
  return row_newton<true>(s, K, deg, nbsteps, nsys, cst, cst_shared, cff, cff_shared, nullptr, 0,
                          x, steps, wdeg, tail, up, bs, history, Q, R, stream,
                          "phc_cmplx_row_newton");
}

int phc_houseqr_batch(int K, int n, int nbatch, const double *A, double *Q, double *R,
                      void *stream) {
  return houseqr_batch<false>(K, n, nbatch, A, Q, R, stream, "phc_houseqr_batch");
}

int phc_cmplx_houseqr_batch(int K, int n, int nbatch, const double *A, double *Q, double *R,
                            void *stream) {
  return houseqr_batch<true>(K, n, nbatch, A, Q, R, stream, "phc_cmplx_houseqr_batch");
}

int phc_qrbs_solve_batch(int K, int n, int degp1, int tailidx, int nbatch, const double *mat,
                         double *rhs, double *sol, double *Q, double *R, int *flags,
                         void *stream) {
  return qrbs_solve_batch<false>(K, n, degp1, tailidx, nbatch, mat, rhs, sol, Q, R, flags,
                                 stream, "phc_qrbs_solve_batch");
}

int phc_cmplx_qrbs_solve_batch(int K, int n, int degp1, int tailidx, int nbatch,
                               const double *mat, double *rhs, double *sol, double *Q, double *R,
                               int *flags, void *stream) {
  return qrbs_solve_batch<true>(K, n, degp1, tailidx, nbatch, mat, rhs, sol, Q, R, flags, stream,
                                "phc_cmplx_qrbs_solve_batch");
}

int phc_linear_residue_batch(int K, int n, int degp1, int tailidx, int nbatch,
                             const double *mat, const double *rhs, const double *sol,
                             double *res, double *resmax, void *stream) {
  return linear_residue_batch<false>(K, n, degp1, tailidx, nbatch, mat, rhs, sol, res, resmax,
                                     stream, "phc_linear_residue_batch");
}

int phc_cmplx_linear_residue_batch(int K, int n, int degp1, int tailidx, int nbatch,
                                   const double *mat, const double *rhs, const double *sol,
                                   double *res, double *resmax, void *stream) {
  return linear_residue_batch<true>(K, n, degp1, tailidx, nbatch, mat, rhs, sol, res, resmax,
                                    stream, "phc_cmplx_linear_residue_batch");
}

}  // extern "C"
