// Multiple-double arithmetic for host and gfx950 device code.
//
// md<2> (double-double) follows the QD library operations used by PHCpack:
//   Norms/double_double_functions.cpp:9-277  (accurate add, sloppy mul,
//   three-quotient div, Karp sqrt).
// md<4> and md<8> (quad- and octo-double) follow the CAMPARY generated code
// used by PHCpack: baileyAdd_fast<K,K,K>, baileyMul_fast<K,K,K>,
// baileyMul_fast<K,1,K>, fast_renorm and renorm with a running pointer:
//   Norms/quad_double_functions.cpp:12-380,  Norms/octo_double_functions.cpp:13-1530.
// The loops below state that generated code once for any K; the order of
// every floating-point operation is the same as in the unrolled reference.
//
// Limb 0 is the most significant limb (hi, hihi, hihihi ...).
//
// two_prod uses a fused multiply-add.  The reference host code uses Dekker's
// split; both return the exact product error for operands whose product
// neither overflows nor underflows, so results are identical there.  Below
// |a b| = 2^-916 the error can fall into the subnormal range, where the two
// round it differently (trailing limbs near 1e-140 and below then differ).
// All code including this header must be compiled with -ffp-contract=off so
// that the plain products of the reference (e.g. "p2 += a_hi*b_lo + a_lo*b_hi")
// keep their two roundings.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define MD_HD __host__ __device__ __forceinline__
#else
#define MD_HD static inline
#endif

namespace md {

MD_HD double quick_two_sum(double a, double b, double &err) {
  double s = a + b;
  err = b - (s - a);
  return s;
}

MD_HD double two_sum(double a, double b, double &err) {
  double s = a + b;
  double bb = s - a;
  err = (a - (s - bb)) + (b - bb);
  return s;
}

MD_HD double two_diff(double a, double b, double &err) {
  double s = a - b;
  double bb = s - a;
  err = (a - (s - bb)) - (b + bb);
  return s;
}

MD_HD double two_prod(double a, double b, double &err) {
  double p = a * b;
  err = __builtin_fma(a, b, -p);
  return p;
}

template <int K> struct mdr {
  double x[K];
};

// ----------------------------------------------------------------- K = 2 ---

using dd = mdr<2>;

MD_HD dd dd_make(double hi, double lo) { dd r; r.x[0] = hi; r.x[1] = lo; return r; }

MD_HD dd add(const dd &a, const dd &b) {
  double s1, s2, t1, t2;
  s1 = two_sum(a.x[0], b.x[0], s2);
  t1 = two_sum(a.x[1], b.x[1], t2);
  s2 += t1;
  s1 = quick_two_sum(s1, s2, s2);
  s2 += t2;
  dd c;
  c.x[0] = quick_two_sum(s1, s2, c.x[1]);
  return c;
}

MD_HD dd sub(const dd &a, const dd &b) {
  double s1, s2, t1, t2;
  s1 = two_diff(a.x[0], b.x[0], s2);
  t1 = two_diff(a.x[1], b.x[1], t2);
  s2 += t1;
  s1 = quick_two_sum(s1, s2, s2);
  s2 += t2;
  dd c;
  c.x[0] = quick_two_sum(s1, s2, c.x[1]);
  return c;
}

MD_HD dd mul(const dd &a, const dd &b) {
  double p1, p2;
  p1 = two_prod(a.x[0], b.x[0], p2);
  p2 += (a.x[0] * b.x[1] + a.x[1] * b.x[0]);
  dd c;
  c.x[0] = quick_two_sum(p1, p2, c.x[1]);
  return c;
}

// ddf_mul_d_dd (double times double-double)
MD_HD dd mul_d(const dd &b, double a) {
  double p1, p2;
  p1 = two_prod(a, b.x[0], p2);
  p2 += (b.x[1] * a);
  dd c;
  c.x[0] = quick_two_sum(p1, p2, c.x[1]);
  return c;
}

// ddf_inc_d
MD_HD dd add_d(const dd &a, double b) {
  double s1, s2;
  s1 = two_sum(a.x[0], b, s2);
  s2 += a.x[1];
  dd c;
  c.x[0] = quick_two_sum(s1, s2, c.x[1]);
  return c;
}

MD_HD dd div(const dd &a, const dd &b) {
  double q1, q2, q3;
  q1 = a.x[0] / b.x[0];
  dd c = sub(a, mul_d(b, q1));
  q2 = c.x[0] / b.x[0];
  c = sub(c, mul_d(b, q2));
  q3 = c.x[0] / b.x[0];
  dd r;
  r.x[0] = quick_two_sum(q1, q2, r.x[1]);
  return add_d(r, q3);
}

// ddf_sqr (Norms/double_double_functions.cpp:206-214): the square adds
// lo*lo, which mul(a, a) omits.  two_prod is exact, as ddf_two_sqr is.
MD_HD dd sqr(const dd &a) {
  double p1, p2;
  p1 = two_prod(a.x[0], a.x[0], p2);
  p2 += 2.0 * a.x[0] * a.x[1];
  p2 += a.x[1] * a.x[1];
  dd c;
  c.x[0] = quick_two_sum(p1, p2, c.x[1]);
  return c;
}

// ddf_sqr as used by the Karp square root: (ax)^2 with ax a double
MD_HD dd sqr_d(double a) {
  double p1, p2;
  p1 = two_prod(a, a, p2);
  p2 += 2.0 * a * 0.0;
  p2 += 0.0 * 0.0;
  dd c;
  c.x[0] = quick_two_sum(p1, p2, c.x[1]);
  return c;
}

MD_HD dd sqrt(const dd &a) {
  dd b;
  if (a.x[0] == 0.0 && a.x[1] == 0.0) { b.x[0] = 0.0; b.x[1] = 0.0; return b; }
  if (a.x[0] < 0.0) { b.x[0] = -1.0; b.x[1] = 0.0; return b; }
  const double x = 1.0 / ::sqrt(a.x[0]);
  double ax = a.x[0] * x;
  dd y = sub(a, sqr_d(ax));
  b.x[0] = two_sum(ax, y.x[0] * x * 0.5, b.x[1]);
  return b;
}

// ------------------------------------------------------------ K = 4, 8 -----

// r[p] = v for a run-time p < K without indexing the array at run time (a
// dynamically indexed local array would live in scratch memory)
template <int K> MD_HD void set_limb(double *r, int p, double v) {
#pragma unroll
  for (int j = 0; j < K; ++j) r[j] = (p == j) ? v : r[j];
}

// renorm with a running pointer over K+1 terms into K limbs
// (qdf_renorm4, odf_renorm8).
//   Common case first: f[1] != 0 and no partial sum pr is zero before the
// last step.  Then the pointer advances at every step and the limbs are
// f[0] and the K-1 quick_two_sum results in order (a zero pr after the last
// step rewrites the last limb with the value it has, or with +0 where that
// value is a zero, which is +0 already since pr != 0 before that step), so
// no selects are needed.  Any other lane takes the general sweep below,
// which the wavefront skips when no lane needs it.
//   General sweep: the limb the pointer rereads is always the one written
// last, so it is kept in `last` instead of being read back.
template <int K> MD_HD void renorm(const double *f, double *r) {
  {
    double p = f[1];
    bool plain = f[1] != 0.0;
    r[0] = f[0];
#pragma unroll
    for (int i = 2; i <= K; ++i) {
      r[i - 1] = quick_two_sum(p, f[i], p);
      if (i < K) plain = plain && (p != 0.0);
    }
    if (__builtin_expect(plain, 1)) return;
  }
  double pr, last;
  int ptr;
#pragma unroll
  for (int i = 0; i < K; ++i) r[i] = 0.0;
  if (f[1] == 0.0) {
    pr = f[0]; ptr = 0;
    r[0] = last = quick_two_sum(pr, f[2], pr);
  } else {
    r[0] = f[0]; pr = f[1]; ptr = 1;
    r[1] = last = quick_two_sum(pr, f[2], pr);
  }
#pragma unroll
  for (int i = 3; i <= K; ++i) {
    if (pr == 0.0) pr = last; else ptr = ptr + 1;
    last = quick_two_sum(pr, f[i], pr);
    set_limb<K>(r, ptr, last);
  }
  if (pr == 0.0) pr = last; else ptr = ptr + 1;
  if (ptr < K && pr != 0.0) { set_limb<K>(r, ptr, pr); ptr = ptr + 1; }
#pragma unroll
  for (int i = 0; i < K; ++i) if (i >= ptr) r[i] = 0.0;
}

// fast_renorm: one sweep of quick_two_sums, then renorm
template <int K> MD_HD mdr<K> fast_renorm(const double *x) {
  double f[K + 1];
  double pr = quick_two_sum(x[K - 1], x[K], f[K]);
#pragma unroll
  for (int i = K - 2; i >= 1; --i) pr = quick_two_sum(x[i], pr, f[i + 1]);
  f[0] = quick_two_sum(x[0], pr, f[1]);
  mdr<K> c;
  renorm<K>(f, c.x);
  return c;
}

// baileyAdd_fast<K,K,K>
template <int K> MD_HD mdr<K> add(const mdr<K> &a, const mdr<K> &b) {
  double f[K + 1], e;
  f[K] = 0.0;
  f[K - 1] = two_sum(a.x[K - 1], b.x[K - 1], e);
  f[K] += e;
#pragma unroll
  for (int i = K - 2; i >= 0; --i) {
    f[i] = two_sum(a.x[i], b.x[i], e);
#pragma unroll
    for (int j = i + 1; j < K; ++j) f[j] = two_sum(f[j], e, e);
    f[K] += e;
  }
  return fast_renorm<K>(f);
}

template <int K> MD_HD mdr<K> neg(const mdr<K> &a) {
  mdr<K> c;
#pragma unroll
  for (int i = 0; i < K; ++i) c.x[i] = -a.x[i];
  return c;
}

// qdf_sub / odf_sub: copy b, negate, increment by a
template <int K> MD_HD mdr<K> sub(const mdr<K> &a, const mdr<K> &b) {
  return add<K>(neg<K>(b), a);
}

// baileyMul_fast<K,K,K>
template <int K> MD_HD mdr<K> mul(const mdr<K> &a, const mdr<K> &b) {
  double f[K + 1], p, e;
  f[K] = a.x[1] * b.x[K - 1];
#pragma unroll
  for (int i = 2; i < K; ++i) f[K] += a.x[i] * b.x[K - i];
#pragma unroll
  for (int L = K - 1; L >= 0; --L) {
    f[L] = two_prod(a.x[0], b.x[L], e);
#pragma unroll
    for (int j = L + 1; j < K; ++j) f[j] = two_sum(f[j], e, e);
    f[K] += e;
#pragma unroll
    for (int i = 1; i <= L; ++i) {
      p = two_prod(a.x[i], b.x[L - i], e);
#pragma unroll
      for (int j = L + 1; j < K; ++j) f[j] = two_sum(f[j], e, e);
      f[K] += e;
      f[L] = two_sum(f[L], p, e);
#pragma unroll
      for (int j = L + 1; j < K; ++j) f[j] = two_sum(f[j], e, e);
      f[K] += e;
    }
  }
  return fast_renorm<K>(f);
}

// baileyMul_fast<K,1,K>
template <int K> MD_HD mdr<K> mul_d(const mdr<K> &a, double b) {
  double f[K + 1], e;
  f[K] = 0.0;
  f[K - 1] = two_prod(a.x[K - 1], b, e);
  f[K] += e;
#pragma unroll
  for (int i = K - 2; i >= 0; --i) {
    f[i] = two_prod(a.x[i], b, e);
#pragma unroll
    for (int j = i + 1; j < K; ++j) f[j] = two_sum(f[j], e, e);
    f[K] += e;
  }
  return fast_renorm<K>(f);
}

// renorm_add1: a + b with b a double (qdf_inc_d, odf_inc_d)
template <int K> MD_HD mdr<K> add_d(const mdr<K> &a, double b) {
  double f[K + 1];
  double pr = two_sum(a.x[K - 1], b, f[K]);
#pragma unroll
  for (int i = K - 2; i >= 1; --i) pr = two_sum(a.x[i], pr, f[i + 1]);
  f[0] = two_sum(a.x[0], pr, f[1]);
  mdr<K> c;
  renorm<K>(f, c.x);
  return c;
}

template <int K> MD_HD mdr<K> div(const mdr<K> &a, const mdr<K> &b) {
  double q[K + 1];
  q[0] = a.x[0] / b.x[0];
  mdr<K> c = sub<K>(a, mul_d<K>(b, q[0]));
#pragma unroll
  for (int i = 1; i < K; ++i) {
    q[i] = c.x[0] / b.x[0];
    c = sub<K>(c, mul_d<K>(b, q[i]));
  }
  q[K] = c.x[0] / b.x[0];
  return fast_renorm<K>(q);
}

template <int K> MD_HD mdr<K> mul_pwr2(const mdr<K> &a, double b) {
  mdr<K> c;
#pragma unroll
  for (int i = 0; i < K; ++i) c.x[i] = a.x[i] * b;
  return c;
}

template <int K> MD_HD mdr<K> widen_half(const mdr<K / 2> &a) {
  mdr<K> c;
#pragma unroll
  for (int i = 0; i < K; ++i) c.x[i] = (i < K / 2) ? a.x[i] : 0.0;
  return c;
}

template <int K> MD_HD mdr<K / 2> lead_half(const mdr<K> &a) {
  mdr<K / 2> c;
#pragma unroll
  for (int i = 0; i < K / 2; ++i) c.x[i] = a.x[i];
  return c;
}

// qdf_sqrt / odf_sqrt: square root at half the precision, then one Newton
// correction b = (b^2 + a) / (2b) at full precision.
MD_HD mdr<4> sqrt(const mdr<4> &a) {
  dd h = sqrt(lead_half<4>(a));
  mdr<4> b = widen_half<4>(h);
  mdr<4> z = mul<4>(b, b);
  z = add<4>(z, a);
  z = div<4>(z, b);
  return mul_pwr2<4>(z, 0.5);
}

MD_HD mdr<8> sqrt(const mdr<8> &a) {
  mdr<4> h = sqrt(lead_half<8>(a));
  mdr<8> b = widen_half<8>(h);
  mdr<8> z = mul<8>(b, b);
  z = add<8>(z, a);
  z = div<8>(z, b);
  return mul_pwr2<8>(z, 0.5);
}

// ------------------------------------------------ uniform generic interface

// qdf_sqr and odf_sqr give the bits of the product a*a
template <int K> MD_HD mdr<K> sqr(const mdr<K> &a) { return mul<K>(a, a); }

template <int K> MD_HD mdr<K> from_d(double a) {
  mdr<K> c;
#pragma unroll
  for (int i = 0; i < K; ++i) c.x[i] = (i == 0) ? a : 0.0;
  return c;
}

template <int K> MD_HD mdr<K> operator+(const mdr<K> &a, const mdr<K> &b) { return add(a, b); }
template <int K> MD_HD mdr<K> operator-(const mdr<K> &a, const mdr<K> &b) { return sub(a, b); }
template <int K> MD_HD mdr<K> operator*(const mdr<K> &a, const mdr<K> &b) { return mul(a, b); }
template <int K> MD_HD mdr<K> operator/(const mdr<K> &a, const mdr<K> &b) { return div(a, b); }
template <int K> MD_HD mdr<K> operator-(const mdr<K> &a) {
  mdr<K> c;
#pragma unroll
  for (int i = 0; i < K; ++i) c.x[i] = -a.x[i];
  return c;
}

}  // namespace md
