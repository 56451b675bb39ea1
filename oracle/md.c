/* CPU oracle: multiple-double arithmetic (test infrastructure only).
 * See md.h for the reference files restated here. */
#include <math.h>
#include <string.h>
#include "md.h"

double omd_quick_two_sum(double a, double b, double *err) {
  double s = a + b;
  *err = b - (s - a);
  return s;
}

double omd_two_sum(double a, double b, double *err) {
  double s = a + b;
  double bb = s - a;
  *err = (a - (s - bb)) + (b - bb);
  return s;
}

static double two_diff(double a, double b, double *err) {
  double s = a - b;
  double bb = s - a;
  *err = (a - (s - bb)) - (b + bb);
  return s;
}

/* Norms/double_double_functions.cpp:148-168 */
static void split(double a, double *hi, double *lo) {
  const double SPLITTER = 134217729.0;          /* 2^27 + 1 */
  const double THRESH = 6.69692879491417e+299;  /* 2^996 */
  double t;
  if (a > THRESH || a < -THRESH) {
    a *= 3.7252902984619140625e-09;
    t = SPLITTER * a;
    *hi = t - (t - a);
    *lo = a - *hi;
    *hi *= 268435456.0;
    *lo *= 268435456.0;
  } else {
    t = SPLITTER * a;
    *hi = t - (t - a);
    *lo = a - *hi;
  }
}

/* Norms/double_double_functions.cpp:170-182 */
double omd_two_prod(double a, double b, double *err) {
  double ah, al, bh, bl;
  double p = a * b;
  split(a, &ah, &al);
  split(b, &bh, &bl);
  *err = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
  return p;
}

/* ------------------------------------------------------------ K = 2 ---- */

static void dd_add(const double *a, const double *b, double *c) {
  double s1, s2, t1, t2;
  s1 = omd_two_sum(a[0], b[0], &s2);
  t1 = omd_two_sum(a[1], b[1], &t2);
  s2 += t1;
  s1 = omd_quick_two_sum(s1, s2, &s2);
  s2 += t2;
  c[0] = omd_quick_two_sum(s1, s2, &c[1]);
}

static void dd_sub(const double *a, const double *b, double *c) {
  double s1, s2, t1, t2;
  s1 = two_diff(a[0], b[0], &s2);
  t1 = two_diff(a[1], b[1], &t2);
  s2 += t1;
  s1 = omd_quick_two_sum(s1, s2, &s2);
  s2 += t2;
  c[0] = omd_quick_two_sum(s1, s2, &c[1]);
}

static void dd_mul(const double *a, const double *b, double *c) {
  double p1, p2;
  p1 = omd_two_prod(a[0], b[0], &p2);
  p2 += (a[0] * b[1] + a[1] * b[0]);
  c[0] = omd_quick_two_sum(p1, p2, &c[1]);
}

static void dd_mul_d(const double *b, double a, double *c) {
  double p1, p2;
  p1 = omd_two_prod(a, b[0], &p2);
  p2 += (b[1] * a);
  c[0] = omd_quick_two_sum(p1, p2, &c[1]);
}

static void dd_add_d(const double *a, double b, double *c) {
  double s1, s2;
  s1 = omd_two_sum(a[0], b, &s2);
  s2 += a[1];
  c[0] = omd_quick_two_sum(s1, s2, &c[1]);
}

static void dd_div(const double *a, const double *b, double *c) {
  double q1, q2, q3, acc[2], r[2];
  q1 = a[0] / b[0];
  dd_mul_d(b, q1, acc);
  dd_sub(a, acc, r);
  q2 = r[0] / b[0];
  dd_mul_d(b, q2, acc);
  dd_sub(r, acc, r);
  q3 = r[0] / b[0];
  r[0] = omd_quick_two_sum(q1, q2, &r[1]);
  dd_add_d(r, q3, c);
}

static void dd_sqrt(const double *a, double *b) {
  if (a[0] == 0.0 && a[1] == 0.0) { b[0] = 0.0; b[1] = 0.0; return; }
  if (a[0] < 0.0) { b[0] = -1.0; b[1] = 0.0; return; }
  {
    const double x = 1.0 / sqrt(a[0]);
    double ax = a[0] * x;
    double sq[2], y[2], p1, p2, hi, lo;
    /* ddf_sqr(ax, 0.0) with two_sqr by split */
    p1 = ax * ax;
    split(ax, &hi, &lo);
    p2 = ((hi * hi - p1) + 2.0 * hi * lo) + lo * lo;
    p2 += 2.0 * ax * 0.0;
    p2 += 0.0 * 0.0;
    sq[0] = omd_quick_two_sum(p1, p2, &sq[1]);
    dd_sub(a, sq, y);
    b[0] = omd_two_sum(ax, y[0] * x * 0.5, &b[1]);
  }
}

/* ---------------------------------------------------------- K = 4, 8 ---- */

static void renorm(int K, const double *f, double *r) {
  double pr;
  int ptr, i;
  for (i = 0; i < K; i++) r[i] = 0.0;
  if (f[1] == 0.0) {
    pr = f[0]; ptr = 0;
    r[0] = omd_quick_two_sum(pr, f[2], &pr);
  } else {
    r[0] = f[0]; pr = f[1]; ptr = 1;
    r[1] = omd_quick_two_sum(pr, f[2], &pr);
  }
  for (i = 3; i <= K; i++) {
    if (pr == 0.0) pr = r[ptr]; else ptr = ptr + 1;
    r[ptr] = omd_quick_two_sum(pr, f[i], &pr);
  }
  if (pr == 0.0) pr = r[ptr]; else ptr = ptr + 1;
  if (ptr < K && pr != 0.0) { r[ptr] = pr; ptr = ptr + 1; }
  for (i = ptr; i < K; i++) r[i] = 0.0;
}

static void fast_renorm(int K, const double *x, double *r) {
  double f[9], pr;
  int i;
  pr = omd_quick_two_sum(x[K - 1], x[K], &f[K]);
  for (i = K - 2; i >= 1; i--) pr = omd_quick_two_sum(x[i], pr, &f[i + 1]);
  f[0] = omd_quick_two_sum(x[0], pr, &f[1]);
  renorm(K, f, r);
}

static void k_add(int K, const double *a, const double *b, double *c) {
  double f[9], e;
  int i, j;
  f[K] = 0.0;
  f[K - 1] = omd_two_sum(a[K - 1], b[K - 1], &e);
  f[K] += e;
  for (i = K - 2; i >= 0; i--) {
    f[i] = omd_two_sum(a[i], b[i], &e);
    for (j = i + 1; j < K; j++) f[j] = omd_two_sum(f[j], e, &e);
    f[K] += e;
  }
  fast_renorm(K, f, c);
}

static void k_mul(int K, const double *a, const double *b, double *c) {
  double f[9], p, e;
  int i, j, L;
  f[K] = a[1] * b[K - 1];
  for (i = 2; i < K; i++) f[K] += a[i] * b[K - i];
  for (L = K - 1; L >= 0; L--) {
    f[L] = omd_two_prod(a[0], b[L], &e);
    for (j = L + 1; j < K; j++) f[j] = omd_two_sum(f[j], e, &e);
    f[K] += e;
    for (i = 1; i <= L; i++) {
      p = omd_two_prod(a[i], b[L - i], &e);
      for (j = L + 1; j < K; j++) f[j] = omd_two_sum(f[j], e, &e);
      f[K] += e;
      f[L] = omd_two_sum(f[L], p, &e);
      for (j = L + 1; j < K; j++) f[j] = omd_two_sum(f[j], e, &e);
      f[K] += e;
    }
  }
  fast_renorm(K, f, c);
}

static void k_mul_d(int K, const double *a, double b, double *c) {
  double f[9], e;
  int i, j;
  f[K] = 0.0;
  f[K - 1] = omd_two_prod(a[K - 1], b, &e);
  f[K] += e;
  for (i = K - 2; i >= 0; i--) {
    f[i] = omd_two_prod(a[i], b, &e);
    for (j = i + 1; j < K; j++) f[j] = omd_two_sum(f[j], e, &e);
    f[K] += e;
  }
  fast_renorm(K, f, c);
}

static void k_add_d(int K, const double *a, double b, double *c) {
  double f[9], pr;
  int i;
  pr = omd_two_sum(a[K - 1], b, &f[K]);
  for (i = K - 2; i >= 1; i--) pr = omd_two_sum(a[i], pr, &f[i + 1]);
  f[0] = omd_two_sum(a[0], pr, &f[1]);
  renorm(K, f, c);
}

static void k_sub(int K, const double *a, const double *b, double *c) {
  double nb[8] = {0};
  int i;
  for (i = 0; i < K; i++) nb[i] = -b[i];
  k_add(K, nb, a, c);
}

static void k_div(int K, const double *a, const double *b, double *c) {
  double q[9], acc[8], r[8];
  int i;
  q[0] = a[0] / b[0];
  k_mul_d(K, b, q[0], acc);
  k_sub(K, a, acc, r);
  for (i = 1; i < K; i++) {
    q[i] = r[0] / b[0];
    k_mul_d(K, b, q[i], acc);
    k_sub(K, r, acc, r);
  }
  q[K] = r[0] / b[0];
  fast_renorm(K, q, c);
}

static void k_sqrt(int K, const double *a, double *c) {
  double b[8], z[8];
  int i, h = K / 2;
  if (h == 2) dd_sqrt(a, b); else k_sqrt(h, a, b);
  for (i = h; i < K; i++) b[i] = 0.0;
  k_mul(K, b, b, z);
  k_add(K, z, a, z);
  k_div(K, z, b, z);
  for (i = 0; i < K; i++) c[i] = z[i] * 0.5;
}

/* ------------------------------------------------------------ dispatch -- */

void omd_add(int K, const double *a, const double *b, double *c) {
  if (K == 1) c[0] = a[0] + b[0];
  else if (K == 2) dd_add(a, b, c);
  else k_add(K, a, b, c);
}

void omd_sub(int K, const double *a, const double *b, double *c) {
  if (K == 1) c[0] = a[0] - b[0];
  else if (K == 2) dd_sub(a, b, c);
  else k_sub(K, a, b, c);
}

void omd_mul(int K, const double *a, const double *b, double *c) {
  if (K == 1) c[0] = a[0] * b[0];
  else if (K == 2) dd_mul(a, b, c);
  else k_mul(K, a, b, c);
}

void omd_div(int K, const double *a, const double *b, double *c) {
  if (K == 1) c[0] = a[0] / b[0];
  else if (K == 2) dd_div(a, b, c);
  else k_div(K, a, b, c);
}

void omd_sqrt(int K, const double *a, double *c) {
  if (K == 1) c[0] = sqrt(a[0]);
  else if (K == 2) dd_sqrt(a, c);
  else k_sqrt(K, a, c);
}

void omd_mul_d(int K, const double *a, double b, double *c) {
  if (K == 1) c[0] = a[0] * b;
  else if (K == 2) dd_mul_d(a, b, c);
  else k_mul_d(K, a, b, c);
}

void omd_add_d(int K, const double *a, double b, double *c) {
  if (K == 1) c[0] = a[0] + b;
  else if (K == 2) dd_add_d(a, b, c);
  else k_add_d(K, a, b, c);
}

/* Norms/double_double_functions.cpp:184-193 (two_sqr) and :206-214 (sqr):
 * the square adds lo*lo, which the product a*a omits.  The quad and octo
 * double squares (qdf_sqr, odf_sqr) give the same bits as their products. */
static void dd_sqr(const double *a, double *b) {
  double hi, lo, p1, p2;
  p1 = a[0] * a[0];
  split(a[0], &hi, &lo);
  p2 = ((hi * hi - p1) + 2.0 * hi * lo) + lo * lo;
  p2 += 2.0 * a[0] * a[1];
  p2 += a[1] * a[1];
  b[0] = omd_quick_two_sum(p1, p2, &b[1]);
}

void omd_sqr(int K, const double *a, double *c) {
  if (K == 1) c[0] = a[0] * a[0];
  else if (K == 2) dd_sqr(a, c);
  else k_mul(K, a, a, c);
}

void omd_neg(int K, const double *a, double *c) {
  int i;
  for (i = 0; i < K; i++) c[i] = -a[i];
}

void omd_copy(int K, const double *a, double *c) { memmove(c, a, sizeof(double) * K); }

void omd_zero(int K, double *c) {
  int i;
  for (i = 0; i < K; i++) c[i] = 0.0;
}

void omd_from_d(int K, double a, double *c) {
  int i;
  c[0] = a;
  for (i = 1; i < K; i++) c[i] = 0.0;
}

/* complex: Convolutions/dbl2_convolutions_host.cpp:42-52 for the product */
void omd_cadd(int K, const double *a, const double *b, double *c) {
  omd_add(K, a, b, c);
  omd_add(K, a + K, b + K, c + K);
}

void omd_csub(int K, const double *a, const double *b, double *c) {
  omd_sub(K, a, b, c);
  omd_sub(K, a + K, b + K, c + K);
}

void omd_cmul(int K, const double *x, const double *y, double *z) {
  double re[8], im[8], t[8];
  omd_mul(K, x, y, re);
  omd_mul(K, x + K, y + K, t);
  omd_sub(K, re, t, re);
  omd_mul(K, x + K, y, im);
  omd_mul(K, x, y + K, t);
  omd_add(K, im, t, im);
  omd_copy(K, re, z);
  omd_copy(K, im, z + K);
}

void omd_cmac(int K, const double *x, const double *y, double *z) {
  double t[8];
  omd_mul(K, x, y, t);
  omd_add(K, z, t, z);
  omd_mul(K, x + K, y + K, t);
  omd_sub(K, z, t, z);
  omd_mul(K, x + K, y, t);
  omd_add(K, z + K, t, z + K);
  omd_mul(K, x, y + K, t);
  omd_add(K, z + K, t, z + K);
}

void omd_cnorm2(int K, const double *a, double *c) {
  double t[8];
  omd_mul(K, a, a, c);
  omd_mul(K, a + K, a + K, t);
  omd_add(K, c, t, c);
}

/* a/b = a * conj(b) / |b|^2 */
void omd_cdiv(int K, const double *a, const double *b, double *c) {
  double den[8], re[8], im[8], t[8];
  omd_cnorm2(K, b, den);
  omd_mul(K, a, b, re);
  omd_mul(K, a + K, b + K, t);
  omd_add(K, re, t, re);
  omd_mul(K, a + K, b, im);
  omd_mul(K, a, b + K, t);
  omd_sub(K, im, t, im);
  omd_div(K, re, den, c);
  omd_div(K, im, den, c + K);
}

void omd_batch(int op, int K, int n, const double *a, const double *b, double *c) {
  double x[8], y[8], z[8];
  int i, j;
  for (j = 0; j < n; j++) {
    for (i = 0; i < K; i++) { x[i] = a[i * n + j]; y[i] = b ? b[i * n + j] : 0.0; }
    switch (op) {
      case 0: omd_add(K, x, y, z); break;
      case 1: omd_sub(K, x, y, z); break;
      case 2: omd_mul(K, x, y, z); break;
      case 3: omd_div(K, x, y, z); break;
      default: omd_sqrt(K, x, z); break;
    }
    for (i = 0; i < K; i++) c[i * n + j] = z[i];
  }
}
