// Uniform real and complex number types over K = 1, 2, 4, 8 limbs for the
// device kernels, with limb-planar loads and stores.
//
// Limb-planar layout: an array of m numbers with K limbs is K planes of m
// doubles (plane i holds limb i).  Complex arrays hold the K real planes then
// the K imaginary planes.  A wavefront that reads number j, j+1, ... of one
// plane reads consecutive doubles, so every limb access is coalesced.
#pragma once
#include "md.hpp"

namespace md {

// K = 1: plain doubles with the same function names
MD_HD double add(double a, double b) { return a + b; }
MD_HD double sub(double a, double b) { return a - b; }
MD_HD double mul(double a, double b) { return a * b; }
MD_HD double div(double a, double b) { return a / b; }
MD_HD double mul_d(double a, double b) { return a * b; }
MD_HD double add_d(double a, double b) { return a + b; }
MD_HD double neg(double a) { return -a; }
MD_HD double sqr(double a) { return a * a; }

template <int K> struct RealT { typedef mdr<K> type; };
template <> struct RealT<1> { typedef double type; };

template <int K> struct R {
  typedef typename RealT<K>::type T;
  MD_HD static T zero() { return from(0.0); }
  MD_HD static T from(double a) {
    T c;
#pragma unroll
    for (int i = 0; i < K; ++i) c.x[i] = (i == 0) ? a : 0.0;
    return c;
  }
  MD_HD static double hi(const T &a) { return a.x[0]; }
  MD_HD static T load(const double *p, size_t stride, size_t j) {
    T c;
#pragma unroll
    for (int i = 0; i < K; ++i) c.x[i] = p[i * stride + j];
    return c;
  }
  MD_HD static void store(double *p, size_t stride, size_t j, const T &c) {
#pragma unroll
    for (int i = 0; i < K; ++i) p[i * stride + j] = c.x[i];
  }
  MD_HD static T sqrt(const T &a) { return md::sqrt(a); }
};

template <> struct R<1> {
  typedef double T;
  MD_HD static T zero() { return 0.0; }
  MD_HD static T from(double a) { return a; }
  MD_HD static double hi(const T &a) { return a; }
  MD_HD static T load(const double *p, size_t, size_t j) { return p[j]; }
  MD_HD static void store(double *p, size_t, size_t j, const T &c) { p[j] = c; }
  MD_HD static T sqrt(const T &a) { return ::sqrt(a); }
};

template <int K> struct cplx {
  typedef typename RealT<K>::type T;
  T re, im;
};

template <int K> struct C {
  typedef cplx<K> Z;
  typedef typename RealT<K>::type T;
  MD_HD static Z zero() { Z z; z.re = R<K>::zero(); z.im = R<K>::zero(); return z; }
  MD_HD static Z from(double re, double im) { Z z; z.re = R<K>::from(re); z.im = R<K>::from(im); return z; }
  // complex array: planes re0..re(K-1), im0..im(K-1)
  MD_HD static Z load(const double *p, size_t stride, size_t j) {
    Z z;
    z.re = R<K>::load(p, stride, j);
    z.im = R<K>::load(p + (size_t)K * stride, stride, j);
    return z;
  }
  MD_HD static void store(double *p, size_t stride, size_t j, const Z &z) {
    R<K>::store(p, stride, j, z.re);
    R<K>::store(p + (size_t)K * stride, stride, j, z.im);
  }
  MD_HD static Z add(const Z &a, const Z &b) { Z c; c.re = md::add(a.re, b.re); c.im = md::add(a.im, b.im); return c; }
  MD_HD static Z sub(const Z &a, const Z &b) { Z c; c.re = md::sub(a.re, b.re); c.im = md::sub(a.im, b.im); return c; }
  MD_HD static Z neg(const Z &a) { Z c; c.re = md::neg(a.re); c.im = md::neg(a.im); return c; }
  // product in the order of CPU_cmplx2_product's first term
  // (Convolutions/dbl2_convolutions_host.cpp:42-52)
  MD_HD static Z mul(const Z &x, const Z &y) {
    Z z;
    z.re = md::mul(x.re, y.re);
    z.re = md::sub(z.re, md::mul(x.im, y.im));
    z.im = md::mul(x.im, y.re);
    z.im = md::add(z.im, md::mul(x.re, y.im));
    return z;
  }
  // z += x*y in the order of CPU_cmplx2_product's inner loop (:75-85)
  MD_HD static void mac(Z &z, const Z &x, const Z &y) {
    z.re = md::add(z.re, md::mul(x.re, y.re));
    z.re = md::sub(z.re, md::mul(x.im, y.im));
    z.im = md::add(z.im, md::mul(x.im, y.re));
    z.im = md::add(z.im, md::mul(x.re, y.im));
  }
  // z -= x*y
  MD_HD static void msub(Z &z, const Z &x, const Z &y) {
    z.re = md::sub(z.re, md::mul(x.re, y.re));
    z.re = md::add(z.re, md::mul(x.im, y.im));
    z.im = md::sub(z.im, md::mul(x.im, y.re));
    z.im = md::sub(z.im, md::mul(x.re, y.im));
  }
  MD_HD static Z scale(const Z &x, const T &s) { Z z; z.re = md::mul(x.re, s); z.im = md::mul(x.im, s); return z; }
  MD_HD static Z scale_d(const Z &x, double s) { Z z; z.re = md::mul_d(x.re, s); z.im = md::mul_d(x.im, s); return z; }
  MD_HD static T norm2(const Z &a) { return md::add(md::mul(a.re, a.re), md::mul(a.im, a.im)); }
  // a / b = a * conj(b) / |b|^2
  MD_HD static Z div(const Z &a, const Z &b) {
    T den = norm2(b);
    T re = md::add(md::mul(a.re, b.re), md::mul(a.im, b.im));
    T im = md::sub(md::mul(a.im, b.re), md::mul(a.re, b.im));
    Z c;
    c.re = md::div(re, den);
    c.im = md::div(im, den);
    return c;
  }
  MD_HD static double abs_hi(const Z &a) {
    double r = R<K>::hi(a.re), i = R<K>::hi(a.im);
    return ::sqrt(r * r + i * i);
  }
};

}  // namespace md
