/* CPU oracle: predictor-corrector tracking of compiled homotopies
 * (test infrastructure only).
 *
 * Restates the series step control of PHCpack's Ada trackers:
 *   Track_One_Path / Predictor_Corrector_Loop
 *     (Deformations/Curves/predictor_corrector_trackers.adb,
 *      predictor_corrector_loops.adb:16-83)
 *   LU_Prediction: pole step pbeta*r with Fabry's ratio (offset 2) and the
 *     predictor feedback halving the step while the predicted residual is
 *     above alpha (standard_predictor_convolutions.adb:1142-1185, Predictor_Feedback)
 *   Convergence_Radius_Estimates.Fabry (Math_Lib/Circuits/convergence_radius_estimates.adb:138-163,354-390)
 *   Bounded_Update (Deformations/Curves/three_way_minima.adb:96-109)
 * and evaluates the homotopy through the program of phcpack_amd/homotopy.py.
 * Every value is computed with the same operations in the same order as the
 * device kernel: sums over the G segments of an equation and over the G
 * parts of a matrix-vector product combine pairwise as a butterfly does
 * ((p0+p1)+(p2+p3) for G = 4), the Jacobian is inverted in place by
 * Gauss-Jordan with partial pivoting (first maximal |re|+|im| of the leading
 * limbs, non-finite values counted as infinite).
 *
 * Parity status: the evaluation (value, Jacobian, Taylor series) is pinned
 * against the reference-order polynomial evaluator (series.c, itself
 * bit-exact with CPU_*_poly_evaldiff) and against closed forms; the step
 * sequence of the Ada tracker cannot be run here (no Ada compiler), so step
 * counts are "parity unpinned" and checked by invariants only.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "md.h"
#include "track.h"

enum { ST_RUNNING = 0, ST_SUCCESS = 1, ST_MINSTEP = 2, ST_MAXSTEPS = 3, ST_SINGULAR = 4,
       ST_DIVERGED = 5 };

#define JT_TERM_BITS 16
#define JT_PAIR_BITS 12

typedef struct {
  int K;
  double *slots, *u, *uj, *A, *rhs, *x0, *col, *W, *dnode, *part;
  int *perm;
} work_t;

#define ZP(a, i) ((a) + (size_t)(i) * 2 * K)

static void czero(int K, double *z) { memset(z, 0, sizeof(double) * 2 * K); }
static void cset(int K, double *z, double re, double im) { czero(K, z); z[0] = re; z[K] = im; }
static void ccopy(int K, const double *a, double *z) { memmove(z, a, sizeof(double) * 2 * K); }
static void cscale_d(int K, const double *a, double s, double *z) {
  omd_mul_d(K, a, s, z);
  omd_mul_d(K, a + K, s, z + K);
}
static void cneg(int K, double *z) {
  int i;
  for (i = 0; i < 2 * K; i++) z[i] = -z[i];
}
/* FP64 tracking (K = 1) forms complex products with fused multiply-adds,
 * as the kernel does; K > 1 keeps the multiple-double operation order. */
static void tc_mul(int K, const double *x, const double *y, double *z) {
  if (K == 1) {
    const double re = fma(x[0], y[0], -(x[1] * y[1]));
    const double im = fma(x[0], y[1], x[1] * y[0]);
    z[0] = re;
    z[1] = im;
  } else {
    omd_cmul(K, x, y, z);
  }
}
static void tc_mac(int K, const double *x, const double *y, double *z) {
  if (K == 1) {
    z[0] = fma(x[0], y[0], z[0]);
    z[0] = fma(-x[1], y[1], z[0]);
    z[1] = fma(x[1], y[0], z[1]);
    z[1] = fma(x[0], y[1], z[1]);
  } else {
    omd_cmac(K, x, y, z);
  }
}
/* z -= x*y in the kernel's order (C<K>::msub) */
static void cmsub(int K, const double *x, const double *y, double *z) {
  double t[8];
  if (K == 1) {
    z[0] = fma(-x[0], y[0], z[0]);
    z[0] = fma(x[1], y[1], z[0]);
    z[1] = fma(-x[1], y[0], z[1]);
    z[1] = fma(-x[0], y[1], z[1]);
    return;
  }
  omd_mul(K, x, y, t);
  omd_sub(K, z, t, z);
  omd_mul(K, x + K, y + K, t);
  omd_add(K, z, t, z);
  omd_mul(K, x + K, y, t);
  omd_sub(K, z + K, t, z + K);
  omd_mul(K, x, y + K, t);
  omd_sub(K, z + K, t, z + K);
}
static double absz(int K, const double *z) { return fabs(z[0]) + fabs(z[K]); }

/* butterfly sum of G complex values in p (result in p[0]) */
static void tree_sum(int K, double *p, int G) {
  int m, g;
  for (m = 1; m < G; m <<= 1)
    for (g = 0; g + m < G; g += 2 * m) omd_cadd(K, ZP(p, g), ZP(p, g + m), ZP(p, g));
}

static double *slot(work_t *w, int off, int mask, int a) {
  const int K = w->K;
  return ZP(w->slots, off + (a & mask));
}

static void coef(const otr_program *P, int K, int t, double *z) {
  cset(K, z, P->term_coef[t], P->term_coef[P->nterm + t]);
}

static void load_weight(const otr_program *P, int K, int j, double *z) {
  const int m = 3 * P->ncomp;
  int i;
  for (i = 0; i < K; i++) {
    z[i] = P->weights[(size_t)i * m + j];
    z[K + i] = P->weights[(size_t)(8 + i) * m + j];
  }
}

static void set_weights(const otr_program *P, work_t *w, double t) {
  const int K = w->K;
  double o0[16], o1[16], o2[16], o2t[16], tmp[16];
  int r;
  for (r = 0; r < P->ncomp; r++) {
    load_weight(P, K, 3 * r, o0);
    load_weight(P, K, 3 * r + 1, o1);
    load_weight(P, K, 3 * r + 2, o2);
    cscale_d(K, o2, t, o2t);
    omd_cadd(K, o1, o2t, tmp);
    cscale_d(K, tmp, t, tmp);
    omd_cadd(K, o0, tmp, ZP(w->W, 3 * r));
    cscale_d(K, o2t, 2.0, tmp);
    omd_cadd(K, o1, tmp, ZP(w->W, 3 * r + 1));
    ccopy(K, o2, ZP(w->W, 3 * r + 2));
  }
}

static void product_pass(const otr_program *P, work_t *w, int k) {
  const int K = w->K;
  double acc[16];
  int i, q;
  for (i = 0; i < P->nprod; i++) {
    const int *nd = P->nodes + 8 * i;
    /* terms q = 0..k alternate between two accumulators (even, odd q) */
    tc_mul(K, slot(w, nd[0], -1, 0), slot(w, nd[2], -1, k), acc);
    if (k >= 1) {
      double a1[16];
      tc_mul(K, slot(w, nd[0], -1, 1), slot(w, nd[2], -1, k - 1), a1);
      for (q = 2; q <= k; q++)
        tc_mac(K, slot(w, nd[0], -1, q), slot(w, nd[2], -1, k - q), (q & 1) ? a1 : acc);
      omd_cadd(K, acc, a1, acc);
    }
    ccopy(K, acc, slot(w, nd[3], nd[4], k));
  }
}

static void correction_pass(const otr_program *P, work_t *w, int k) {
  const int K = w->K;
  double d[16];
  const double *da;
  int i;
  for (i = 0; i < P->nprod; i++) {
    const int *nd = P->nodes + 8 * i;
    da = (nd[5] < 0) ? slot(w, nd[0], -1, k) : ZP(w->dnode, nd[5]);
    tc_mul(K, slot(w, nd[0], -1, 0), slot(w, nd[2], -1, k), d);
    tc_mac(K, da, slot(w, nd[2], -1, 0), d);
    if (nd[6] >= 0) ccopy(K, d, ZP(w->dnode, nd[6]));
    omd_cadd(K, slot(w, nd[3], nd[4], k), d, slot(w, nd[3], nd[4], k));
  }
}

/* rhs_i = -[H_i]_k */
static void residual(const otr_program *P, work_t *w, int k) {
  const int K = w->K, G = P->G;
  double acc[16], c[16];
  int q, l, i, g, t;
  for (q = 0; q < P->npair; q++) {
    const int *pr = P->pairs + 4 * q;
    tc_mul(K, ZP(w->W, 3 * pr[0]), slot(w, pr[1], pr[2], k), acc);
    for (l = 1; l <= P->wdeg && l <= k; l++)
      tc_mac(K, ZP(w->W, 3 * pr[0] + l), slot(w, pr[1], pr[2], k - l), acc);
    ccopy(K, acc, ZP(w->u, q));
  }
  for (i = 0; i < P->n; i++) {
    for (g = 0; g < G; g++) {
      const int s = i * G + g;
      double a4[4][16];
      int r;
      for (r = 0; r < 4; r++) czero(K, a4[r]);
      /* term t goes to accumulator (t - first) mod 4 */
      for (t = P->seg_ptr[s]; t < P->seg_ptr[s + 1]; t++) {
        coef(P, K, t, c);
        tc_mac(K, c, ZP(w->u, P->term_pair[t]), a4[(t - P->seg_ptr[s]) & 3]);
      }
      omd_cadd(K, a4[0], a4[1], a4[0]);
      omd_cadd(K, a4[2], a4[3], a4[2]);
      omd_cadd(K, a4[0], a4[2], acc);
      ccopy(K, acc, ZP(w->part, g));
    }
    tree_sum(K, w->part, G);
    ccopy(K, ZP(w->part, 0), ZP(w->rhs, i));
    cneg(K, ZP(w->rhs, i));
  }
}

static void jacobian(const otr_program *P, work_t *w) {
  const int K = w->K;
  double acc[16], c[16], v[16];
  int q, e, t;
  for (q = 0; q < P->njpair; q++)
    tc_mul(K, ZP(w->W, 3 * P->jpairs[2 * q]), ZP(w->slots, P->jpairs[2 * q + 1]), ZP(w->uj, q));
  for (e = 0; e < P->n * P->n; e++) {
    double a4[4][16];
    int r;
    for (r = 0; r < 4; r++) czero(K, a4[r]);
    for (t = P->jent_ptr[e]; t < P->jent_ptr[e + 1]; t++) {
      const unsigned jt = (unsigned)P->jterms[t];
      const int term = jt & ((1u << JT_TERM_BITS) - 1);
      const int qq = (jt >> JT_TERM_BITS) & ((1u << JT_PAIR_BITS) - 1);
      const int ex = jt >> (JT_TERM_BITS + JT_PAIR_BITS);
      if (ex == 1) ccopy(K, ZP(w->uj, qq), v); else cscale_d(K, ZP(w->uj, qq), (double)ex, v);
      coef(P, K, term, c);
      tc_mac(K, c, v, a4[(t - P->jent_ptr[e]) & 3]);
    }
    omd_cadd(K, a4[0], a4[1], a4[0]);
    omd_cadd(K, a4[2], a4[3], a4[2]);
    omd_cadd(K, a4[0], a4[2], acc);
    ccopy(K, acc, ZP(w->A, e));
  }
}

/* A := (P A)^{-1} in place, perm records P; 0 on a zero/non-finite pivot */
static int gj_invert(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n;
  double one[16], inv[16], tmp[16], *x;
  int j, i, c, p, e;
  cset(K, one, 1.0, 0.0);
  for (i = 0; i < n; i++) w->perm[i] = i;
  for (j = 0; j < n; j++) {
    double amax = -1.0;
    p = j;
    for (i = j; i < n; i++) {
      double v = absz(K, ZP(w->A, i * n + j));
      if (v != v) v = INFINITY;
      if (v > amax) { amax = v; p = i; }
    }
    if (!(amax > 0.0) || amax == INFINITY) return 0;
    if (p != j) {
      for (c = 0; c < n; c++) {
        ccopy(K, ZP(w->A, j * n + c), tmp);
        ccopy(K, ZP(w->A, p * n + c), ZP(w->A, j * n + c));
        ccopy(K, tmp, ZP(w->A, p * n + c));
      }
      i = w->perm[j]; w->perm[j] = w->perm[p]; w->perm[p] = i;
    }
    for (i = 0; i < n; i++) ccopy(K, ZP(w->A, i * n + j), ZP(w->col, i));
    omd_cdiv(K, one, ZP(w->col, j), inv);
    for (c = 0; c < n; c++) {
      x = ZP(w->A, j * n + c);
      if (c == j) ccopy(K, inv, x); else tc_mul(K, x, inv, x);
    }
    for (e = 0; e < n * n; e++) {
      i = e / n;
      c = e - i * n;
      if (i == j) continue;
      if (c == j) {
        tc_mul(K, ZP(w->col, i), inv, ZP(w->A, e));
        cneg(K, ZP(w->A, e));
      } else {
        cmsub(K, ZP(w->col, i), ZP(w->A, j * n + c), ZP(w->A, e));
      }
    }
  }
  return 1;
}

/* y = (P A)^{-1} (P rhs), G parts per row combined by a butterfly */
static void apply_inverse(const otr_program *P, work_t *w, double *y) {
  const int K = w->K, n = P->n, G = P->G;
  int i, g, j;
  for (i = 0; i < n; i++) {
    for (g = 0; g < G; g++) {
      const int jlo = (g * n) / G, jhi = ((g + 1) * n) / G;
      double *acc = ZP(w->part, g);
      czero(K, acc);
      for (j = jlo; j < jhi; j++) tc_mac(K, ZP(w->A, i * n + j), ZP(w->rhs, w->perm[j]), acc);
    }
    tree_sum(K, w->part, G);
    ccopy(K, ZP(w->part, 0), ZP(y, i));
  }
}

static void eval_point(const otr_program *P, work_t *w, int jac) {
  product_pass(P, w, 0);
  residual(P, w, 0);
  if (jac) jacobian(P, w);
}

static double max_abs(int K, const double *v, int n) {
  double m = 0.0;
  int i;
  for (i = 0; i < n; i++) m = fmax(m, absz(K, ZP(v, i)));
  return m;
}

static double max_abs_x(const otr_program *P, work_t *w) {
  const int K = w->K, D = P->deg + 1;
  double m = 0.0;
  int j;
  for (j = 0; j < P->n; j++) m = fmax(m, absz(K, ZP(w->slots, j * D)));
  return m;
}

/* log2 from the binary exponent and 2(z + z^3/3 + z^5/5 + z^7/7),
 * z = (m-1)/(m+1): the same IEEE double operations as the device */
double otr_lg2(double x) {
  int e = 0;
  double m, z, z2, sr;
  if (!(x > 0.0)) return -1100.0;
  if (x == INFINITY) return 1100.0;
  m = 2.0 * frexp(x, &e);
  z = (m - 1.0) / (m + 1.0);
  z2 = z * z;
  sr = z * (1.0 + z2 * (1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (1.0 / 7.0))));
  return (double)(e - 1) + sr * 2.8853900817779268;
}

static unsigned mix_hash(unsigned seed, unsigned path, unsigned step) {
  unsigned h = seed ^ (path * 0x9E3779B1u) ^ (step * 0x85EBCA77u);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

/* the device's MLP inference: hidden unit j per lane, output sums as a
 * 64-lane butterfly */
static int mlp_choose(const otr_controller *ctl, const double *f) {
  const int F = OTR_NFEAT, H = OTR_HIDDEN, C = ctl->ncand;
  const float *W1 = ctl->weights, *b1 = W1 + F * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H;
  const float *b3 = W3 + H * C;
  float h1[OTR_HIDDEN], g[OTR_HIDDEN], v[OTR_HIDDEN];
  int i, j, c, m, best = 0;
  for (j = 0; j < H; j++) {
    float h = b1[j];
    for (i = 0; i < F; i++) h += W1[i * H + j] * (((float)f[i] - ctl->mu[i]) / ctl->sd[i]);
    h1[j] = h < 0.0f ? 0.0f : h;
  }
  for (j = 0; j < H; j++) {
    float a = b2[j];
    for (i = 0; i < H; i++) a += W2[i * H + j] * h1[i];
    g[j] = a < 0.0f ? 0.0f : a;
  }
  for (c = 0; c < C; c++) {
    for (j = 0; j < H; j++) v[j] = W3[j * C + c] * g[j];
    for (m = 1; m < H; m <<= 1)
      for (j = 0; j + m < H; j += 2 * m) v[j] = v[j] + v[j + m];
    if (v[0] + b3[c] >= ctl->logit_threshold) best = c;
  }
  return best;
}

static int newton(const otr_program *P, work_t *w, int iters, double tol, double *dxl, int *count,
                  double *dxh, int evaluated) {
  const int K = w->K, D = P->deg + 1;
  double dx[16 * 64];
  int it, j;
  for (it = 0; it < iters; it++) {
    double m;
    ++*count;
    /* evaluated: products and residual at this point are already there */
    if (it == 0 && evaluated) jacobian(P, w); else eval_point(P, w, 1);
    if (!gj_invert(P, w)) return -1;
    apply_inverse(P, w, dx);
    for (j = 0; j < P->n; j++) omd_cadd(K, ZP(w->slots, j * D), ZP(dx, j), ZP(w->slots, j * D));
    m = max_abs(K, dx, P->n);
    *dxl = m;
    if (dxh) { dxh[3] = dxh[2]; dxh[2] = dxh[1]; dxh[1] = dxh[0]; dxh[0] = m; }
    if (m <= tol * (1.0 + max_abs_x(P, w))) return 1;
  }
  return 0;
}

static int work_alloc(const otr_program *P, int K, work_t *w) {
  const size_t zs = sizeof(double) * 2 * K;
  w->K = K;
  w->slots = calloc((size_t)P->nslots, zs);
  w->u = calloc((size_t)(P->npair > 0 ? P->npair : 1), zs);
  w->uj = calloc((size_t)(P->njpair > 0 ? P->njpair : 1), zs);
  w->A = calloc((size_t)P->n * P->n, zs);
  w->rhs = calloc((size_t)P->n, zs);
  w->x0 = calloc((size_t)P->n, zs);
  w->col = calloc((size_t)P->n, zs);
  w->W = calloc((size_t)3 * P->ncomp, zs);
  w->dnode = calloc((size_t)(P->nfactor > 0 ? P->nfactor : 1), zs);
  w->part = calloc((size_t)(P->G > P->n ? P->G : P->n) + 64, zs);
  w->perm = calloc((size_t)P->n, sizeof(int));
  return w->slots && w->u && w->uj && w->A && w->rhs && w->x0 && w->col && w->W && w->dnode &&
         w->part && w->perm;
}

static void work_free(work_t *w) {
  free(w->slots); free(w->u); free(w->uj); free(w->A); free(w->rhs); free(w->x0);
  free(w->col); free(w->W); free(w->dnode); free(w->part); free(w->perm);
}

static void load_point(const otr_program *P, work_t *w, const double *x, int KS, int npaths,
                       int path) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  int j, i, a;
  for (j = 0; j < n; j++) {
    double *z = ZP(w->slots, j * D);
    czero(K, z);
    for (i = 0; i < K && i < KS; i++) {
      z[i] = x[(size_t)i * npaths * n + (size_t)path * n + j];
      z[K + i] = x[(size_t)(KS + i) * npaths * n + (size_t)path * n + j];
    }
  }
  for (a = 0; a < D; a++) {
    if (a == 0) cset(K, ZP(w->slots, n * D), 1.0, 0.0);
    else czero(K, ZP(w->slots, n * D + a));
  }
}

static void store_point(const otr_program *P, work_t *w, double *x, int npaths, int path) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  int j, i;
  for (j = 0; j < n; j++) {
    const double *z = ZP(w->slots, j * D);
    for (i = 0; i < K; i++) {
      x[(size_t)i * npaths * n + (size_t)path * n + j] = z[i];
      x[(size_t)(K + i) * npaths * n + (size_t)path * n + j] = z[K + i];
    }
  }
}

/* [L/M] Pade approximant of one component's series c[0..L+M] (complex,
 * 2K doubles each), Deformations/Curves/standard_rational_approximations.adb:
 * Denominator_System (:7-32) sets a(i,j) = c_{L-M+i+j-1} (zero where that
 * index is below 1) and b_i = -c_{L+i}; Pade (:80-117) factors a by lufac
 * (Math_Lib/Reduction/standard_complex_linear_solvers.adb:123-178: partial
 * pivoting on |re| + |im|, multipliers -a_ik/a_kk, info /= 0 on a zero
 * pivot) and solves by lusolve (:552-583); the denominator coefficients are
 * the solution reversed, d_i = b_{M+1-i}, and Numerator_Coefficients
 * (:34-56) gives c_i + sum_{j=1..min(i,M)} d_j c_{i-j}.  A singular system
 * leaves the denominator 1 and a zero numerator, or the Taylor numerator
 * when b is zero.  Pivots are compared on the leading limbs.  In place:
 * c_1..c_L become the numerator's, c_{L+1}..c_{L+M} become d_1..d_M; x0
 * (= c_0, kept apart) is zeroed with the numerator. */
/* z += x*y as the reference writes it: the product, then the sum */
static void cadd_mul(int K, const double *x, const double *y, double *z) {
  double u[16];
  tc_mul(K, x, y, u);
  omd_cadd(K, z, u, z);
}

void otr_pade_coefficients(int K, double *c, int L, int M, double *x0) {
  double a[8][8][16], b[8][16], t[16], one[16];
  int piv[8], info = 0, i, j, k;
  if (M < 1 || M > 8) return;
  for (i = 0; i < M; i++) {
    for (j = 0; j < M; j++) {
      const int idx = L - M + i + j + 1;
      if (idx < 1) czero(K, a[i][j]); else ccopy(K, ZP(c, idx), a[i][j]);
    }
    ccopy(K, ZP(c, L + i + 1), b[i]);
    cneg(K, b[i]);
  }
  /* lufac */
  for (k = 0; k < M - 1; k++) {
    int l = k;
    double smax = absz(K, a[k][k]);
    for (i = k + 1; i < M; i++) {
      const double v = absz(K, a[i][k]);
      if (v > smax) { l = i; smax = v; }
    }
    piv[k] = l;
    if (smax == 0.0) { info = k + 1; continue; }
    if (l != k) { ccopy(K, a[l][k], t); ccopy(K, a[k][k], a[l][k]); ccopy(K, t, a[k][k]); }
    cset(K, one, 1.0, 0.0);
    omd_cdiv(K, one, a[k][k], t);
    cneg(K, t);
    for (i = k + 1; i < M; i++) { double u[16]; tc_mul(K, t, a[i][k], u); ccopy(K, u, a[i][k]); }
    for (j = k + 1; j < M; j++) {
      double tj[16];
      ccopy(K, a[l][j], tj);
      if (l != k) { ccopy(K, a[k][j], a[l][j]); ccopy(K, tj, a[k][j]); }
      for (i = k + 1; i < M; i++) cadd_mul(K, tj, a[i][k], a[i][j]);
    }
  }
  piv[M - 1] = M - 1;
  if (a[M - 1][M - 1][0] == 0.0 && a[M - 1][M - 1][K] == 0.0) info = M;
  if (info != 0) {
    int rhs0 = 1;
    for (i = 0; i < M; i++) rhs0 = rhs0 && b[i][0] == 0.0 && b[i][K] == 0.0;
    if (!rhs0) {
      for (i = 1; i <= L; i++) czero(K, ZP(c, i));
      czero(K, x0);
    }
    for (i = 1; i <= M; i++) czero(K, ZP(c, L + i));
    return;
  }
  /* lusolve */
  for (k = 0; k < M - 1; k++) {
    const int l = piv[k];
    ccopy(K, b[l], t);
    if (l != k) { ccopy(K, b[k], b[l]); ccopy(K, t, b[k]); }
    for (i = k + 1; i < M; i++) cadd_mul(K, t, a[i][k], b[i]);
  }
  for (k = M - 1; k >= 0; k--) {
    omd_cdiv(K, b[k], a[k][k], t);
    ccopy(K, t, b[k]);
    cneg(K, t);
    for (j = 0; j < k; j++) cadd_mul(K, t, a[j][k], b[j]);
  }
  /* numerator, from the top down so that c_{i-j} is still the series' */
  for (i = L; i >= 1; i--)
    for (j = 1; j <= (i < M ? i : M); j++) cadd_mul(K, b[M - j], ZP(c, i - j), ZP(c, i));
  for (i = 1; i <= M; i++) ccopy(K, b[M - i], ZP(c, L + i));
}

/* predicted component at step s: the numerator (or Taylor polynomial of
 * degree L when pden = 0) by Horner from c_L down, x0 added last, divided
 * by the denominator 1 + d_1 s + ... + d_M s^M by Horner from d_M down
 * (Evaluate, standard_rational_approximations.adb:151-205) */
void otr_pade_evaluate(int K, const double *c, int L, int pden, const double *x0, double s,
                       double *out) {
  double acc[16], tmp[16], p[16];
  int k;
  ccopy(K, ZP(c, L), acc);
  for (k = L - 1; k >= 1; k--) {
    cscale_d(K, acc, s, tmp);
    omd_cadd(K, tmp, ZP(c, k), acc);
  }
  cscale_d(K, acc, s, tmp);
  omd_cadd(K, x0, tmp, p);
  if (pden >= 1) {
    double one[16];
    cset(K, one, 1.0, 0.0);
    ccopy(K, ZP(c, L + pden), acc);
    for (k = pden - 1; k >= 1; k--) {
      cscale_d(K, acc, s, tmp);
      omd_cadd(K, tmp, ZP(c, L + k), acc);
    }
    cscale_d(K, acc, s, tmp);
    omd_cadd(K, tmp, one, acc);
    omd_cdiv(K, p, acc, out);
  } else {
    ccopy(K, p, out);
  }
}

static int predictor_series(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  double xk[16 * 64];
  int k, j;
  eval_point(P, w, 1);
  if (!gj_invert(P, w)) return 0;
  for (k = 1; k < D; k++) {
    for (j = 0; j < n; j++) czero(K, ZP(w->slots, j * D + k));
    product_pass(P, w, k);
    residual(P, w, k);
    apply_inverse(P, w, xk);
    for (j = 0; j < n; j++) ccopy(K, ZP(xk, j), ZP(w->slots, j * D + k));
    correction_pass(P, w, k);
  }
  return 1;
}

void otr_track(const otr_program *P, int K, int npaths, int p0, int p1, const double *start,
               int KS, double *endp, double *tend, int *status, int *nsteps, int *nrej,
               double *dxlast, int *work, const otr_params *prm, const otr_controller *ctl) {
  const int n = P->n, D = P->deg + 1;
  const int mode = ctl ? ctl->mode : 0;
  const int pden = prm->pade_dendeg;
  const int pnum = (pden == 0 && prm->pade_numdeg == 0) ? D - 1 : prm->pade_numdeg;
  work_t w;
  int path;
  if (n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = p0; path < p1; path++) {
    double t = 0.0, h = prm->maxstep, dxl = 0.0, hprev = prm->maxstep;
    double dxh[4] = {1.0, 1.0, 1.0, 1.0};
    int st = ST_RUNNING, steps = 0, rej = 0, j, k, npred = 0, nnewt = 0, nfeval = 0, ncorr = 0;
    int rejprev = 0, nlog = 0, attempt, pmode = mode;
    /* a path the learned controller (mode 1) does not finish is tracked
       again from its start with the plain pole-step rule; the work counters
       add up both attempts, the other outputs are the last attempt's */
    for (attempt = 0; attempt < 2; attempt++) {
    pmode = (attempt == 0) ? mode : 0;
    t = 0.0; h = prm->maxstep; dxl = 0.0; hprev = prm->maxstep;
    for (j = 0; j < 4; j++) dxh[j] = 1.0;
    st = ST_RUNNING; steps = 0; rej = 0; rejprev = 0;
    load_point(P, &w, start, KS, npaths, path);
    while (st == ST_RUNNING && t < 1.0) {
      const int kd = D - 1 - prm->fabry_offset;
      double r = INFINITY;
      int accepted = 0;
      set_weights(P, &w, t);
      ++npred;
      if (!predictor_series(P, &w)) { st = ST_SINGULAR; break; }
      if (kd >= 1)
        for (j = 0; j < n; j++) {
          const double den = absz(K, ZP(w.slots, j * D + kd));
          if (den > 0.0) r = fmin(r, absz(K, ZP(w.slots, j * D + kd - 1)) / den);
        }
      h = (r < INFINITY && r > 0.0) ? fmin(prm->pbeta * r, prm->maxstep) : prm->maxstep;
      {
        int cand = 0, rej0, nnewt0, i;
        double feat[OTR_NFEAT];
        if (pmode != 0) {
          feat[0] = otr_lg2((r < INFINITY && r > 0.0) ? r : prm->maxstep);
          feat[1] = otr_lg2(hprev);
          feat[2] = t;
          for (k = 1; k <= 8; k++) {
            double c = 0.0;
            for (j = 0; j < n; j++) c = fmax(c, k < D ? absz(K, ZP(w.slots, j * D + k)) : 0.0);
            feat[2 + k] = otr_lg2(c + 1e-300);
          }
          for (i = 0; i < 4; i++) feat[11 + i] = otr_lg2(dxh[i] + 1e-300);
          feat[15] = (double)rejprev;
          if (pmode == 1) cand = mlp_choose(ctl, feat);
          else cand = (int)(mix_hash(ctl->seed, (unsigned)path, (unsigned)(npred - 1)) % (unsigned)ctl->ncand);
          h *= ctl->mult[cand];
          if (ctl->pole_cap > 0.0 && r < INFINITY && r > 0.0) h = fmin(h, ctl->pole_cap * r);
        }
        rej0 = rej;
        nnewt0 = nnewt;
      for (j = 0; j < n; j++) ccopy(K, ZP(w.slots, j * D), ZP(w.x0, j));
      if (pden >= 1)
        for (j = 0; j < n; j++) otr_pade_coefficients(K, ZP(w.slots, j * D), pnum, pden, ZP(w.x0, j));
      while (!accepted) {
        double tn = t + h, hs, res;
        int ok;
        if (tn >= 1.0 - prm->minstep) tn = 1.0;
        hs = tn - t;
        for (j = 0; j < n; j++)
          otr_pade_evaluate(K, ZP(w.slots, j * D), pnum, pden, ZP(w.x0, j), hs, ZP(w.slots, j * D));
        set_weights(P, &w, tn);
        ++nfeval;
        eval_point(P, &w, 0);
        res = max_abs(K, w.rhs, n);
        ok = res <= prm->alpha * (1.0 + max_abs_x(P, &w));
        if (ok) {
          const int nr = newton(P, &w, prm->corsteps, prm->tolcor, &dxl, &nnewt, dxh, 1);
          ++ncorr;
          if (nr < 0) { st = ST_SINGULAR; break; }
          ok = (nr == 1);
        }
        if (ok) { t = tn; accepted = 1; hprev = hs; }
        else {
          ++rej;
          h *= 0.5;
          if (h < prm->minstep) { st = ST_MINSTEP; break; }
        }
      }
        rejprev = rej - rej0;
        if (mode == 2 && nlog < ctl->log_capacity) {
          for (i = 0; i < OTR_NFEAT; i++)
            ctl->log_x[((size_t)path * ctl->log_capacity + nlog) * OTR_NFEAT + i] = (float)feat[i];
          ctl->log_c[(size_t)path * ctl->log_capacity + nlog] = cand;
          /* accepted at the first attempt: the Newton iterations it took */
          ctl->log_y[(size_t)path * ctl->log_capacity + nlog] =
              (rej == rej0 && st == ST_RUNNING) ? nnewt - nnewt0 : 0;
          ++nlog;
        }
      }
      if (st != ST_RUNNING) break;
      ++steps;
      if (max_abs_x(P, &w) > prm->divergence) { st = ST_DIVERGED; break; }
      if (steps >= prm->maxsteps && t < 1.0) { st = ST_MAXSTEPS; break; }
    }
    if (st == ST_RUNNING) {
      int nr;
      set_weights(P, &w, 1.0);
      nr = newton(P, &w, prm->endsteps, prm->tolend, &dxl, &nnewt, NULL, 0);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
    }
    if (!(pmode == 1 && st != ST_SUCCESS)) break;
    }
    store_point(P, &w, endp, npaths, path);
    tend[path] = t;
    status[path] = st;
    nsteps[path] = steps;
    nrej[path] = rej;
    dxlast[path] = dxl;
    if (work) {
      work[path] = npred;
      work[npaths + path] = nnewt;
      work[2 * npaths + path] = nfeval;
      work[3 * npaths + path] = ncorr;
    }
    if (mode == 2) ctl->log_n[path] = nlog;
  }
  work_free(&w);
}

typedef struct {
  const otr_program *P;
  int K, npaths, p0, p1, KS;
  const double *start;
  double *endp, *tend, *dxlast;
  int *status, *nsteps, *nrej, *work;
  const otr_params *prm;
  const otr_controller *ctl;
} job_t;

static void *run_job(void *arg) {
  job_t *j = (job_t *)arg;
  otr_track(j->P, j->K, j->npaths, j->p0, j->p1, j->start, j->KS, j->endp, j->tend, j->status,
            j->nsteps, j->nrej, j->dxlast, j->work, j->prm, j->ctl);
  return NULL;
}

void otr_track_mt(const otr_program *P, int K, int npaths, const double *start, int KS,
                  double *endp, double *tend, int *status, int *nsteps, int *nrej,
                  double *dxlast, int *work, const otr_params *prm, const otr_controller *ctl,
                  int nthreads) {
  pthread_t th[256];
  job_t jobs[256];
  int i;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  for (i = 0; i < nthreads; i++) {
    jobs[i].P = P; jobs[i].K = K; jobs[i].npaths = npaths; jobs[i].KS = KS;
    jobs[i].p0 = (int)((long)npaths * i / nthreads);
    jobs[i].p1 = (int)((long)npaths * (i + 1) / nthreads);
    jobs[i].start = start; jobs[i].endp = endp; jobs[i].tend = tend; jobs[i].dxlast = dxlast;
    jobs[i].status = status; jobs[i].nsteps = nsteps; jobs[i].nrej = nrej; jobs[i].prm = prm;
    jobs[i].work = work;
    jobs[i].ctl = ctl;
    pthread_create(&th[i], NULL, run_job, &jobs[i]);
  }
  for (i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
}

void otr_refine(const otr_program *P, int K, int npaths, double t, const double *xin, int KS,
                double *xout, int iters, double tol, double *dxlast, double *res, int *status) {
  work_t w;
  int path;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = 0; path < npaths; path++) {
    double dxl = 0.0;
    int nr = 1, its = 0;
    load_point(P, &w, xin, KS, npaths, path);
    set_weights(P, &w, t);
    if (iters > 0) nr = newton(P, &w, iters, tol, &dxl, &its, NULL, 0);
    eval_point(P, &w, 0);
    res[path] = max_abs(K, w.rhs, P->n);
    store_point(P, &w, xout, npaths, path);
    dxlast[path] = dxl;
    status[path] = nr;
  }
  work_free(&w);
}

void otr_eval(const otr_program *P, int K, double t, const double *x, double *val, double *jac) {
  const int n = P->n;
  work_t w;
  int i, j, e;
  if (!work_alloc(P, K, &w)) { work_free(&w); return; }
  load_point(P, &w, x, K, 1, 0);
  set_weights(P, &w, t);
  eval_point(P, &w, 1);
  for (i = 0; i < n; i++) {
    cneg(K, ZP(w.rhs, i));  /* the residual routine stores -H */
    for (j = 0; j < K; j++) {
      val[(size_t)j * n + i] = ZP(w.rhs, i)[j];
      val[(size_t)(K + j) * n + i] = ZP(w.rhs, i)[K + j];
    }
  }
  for (e = 0; e < n * n; e++)
    for (j = 0; j < K; j++) {
      jac[(size_t)j * n * n + e] = ZP(w.A, e)[j];
      jac[(size_t)(K + j) * n * n + e] = ZP(w.A, e)[K + j];
    }
  work_free(&w);
}

int otr_series(const otr_program *P, int K, double t, const double *x, double *series) {
  const int n = P->n, D = P->deg + 1;
  work_t w;
  int j, k, i, ok;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return 0; }
  load_point(P, &w, x, K, 1, 0);
  set_weights(P, &w, t);
  ok = predictor_series(P, &w);
  for (j = 0; j < n; j++)
    for (k = 0; k < D; k++)
      for (i = 0; i < K; i++) {
        series[((size_t)i * n + j) * D + k] = ZP(w.slots, j * D + k)[i];
        series[((size_t)(K + i) * n + j) * D + k] = ZP(w.slots, j * D + k)[K + i];
      }
  work_free(&w);
  return ok;
}
