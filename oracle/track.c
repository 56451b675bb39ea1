/* CPU oracle: predictor-corrector tracking of compiled homotopies
 * (test infrastructure only).
 *
 * Restates the series step control of PHCpack's Ada trackers:
 *   Track_One_Path / Track_All_Paths (Deformations/Curves/
 *     predictor_corrector_trackers.adb:29-76,160-245) and
 *     Predictor_Corrector_Loop (predictor_corrector_loops.adb:16-83);
 *   LU_Prediction (standard_predictor_convolutions.adb:1142-1185): the pole
 *     step pbeta*r with Fabry's ratio (Math_Lib/Circuits/
 *     convergence_radius_estimates.adb:28-38,138-163,354-390), the curvature
 *     step of Hesse_Pade (:558-572,648-900; Step_Distance in
 *     series_and_predictors.adb:813-825; Solution_Error in
 *     homotopy_pade_approximants.adb:558-622), Minimum and Bounded_Update
 *     (three_way_minima.adb), and Predictor_Feedback (:958-993) on the
 *     mixed residual (Math_Lib/.../standard_mixed_residuals.adb:472-484);
 *   the corrector LU_Newton_Steps with mixed residuals
 *     (standard_inlined_newton_circuits.adb:157-175,388-422);
 * and evaluates the homotopy through the program of phcpack_amd/homotopy.py.
 * Departures from the reference: t advances by the step at which the point
 * was corrected (the reference advances t before the feedback, keeps it
 * when a step is halved, and lets the next step's series Newton move the
 * point there); the series is solved order by order from the corrected
 * point; a corrector that gives up below minstep ends the path (status
 * minstep); a singular Jacobian ends the path; the step-control quantities
 * (radius, singular values, residual moduli) are computed in FP64 from the
 * leading limbs; the singular values come from powers of the Gram matrix
 * and a Rayleigh quotient (otr_sigma_max_sq), not from an SVD.
 * Every value is computed with the same operations in the same order as
 * the device kernel: sums over the G segments of an equation and over the
 * G parts of a matrix-vector product combine pairwise as a butterfly does
 * ((p0+p1)+(p2+p3) for G = 4), the Jacobian is inverted in place by
 * Gauss-Jordan with partial pivoting (first maximal |re|+|im| of the leading
 * limbs, non-finite values counted as infinite).
 *
 * Parity status: the evaluation (value, Jacobian, Taylor series) is pinned
 * against the reference-order polynomial evaluator (series.c, itself
 * bit-exact with CPU_*_poly_evaldiff) and against closed forms; the pieces
 * of the step rule against closed forms and numpy (tests/test_oracle_step.py);
 * the Ada tracker itself cannot be run here (no Ada compiler), so whole
 * step sequences are "parity unpinned" and checked by invariants.
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
  double *rn, *mr, *H, *SP, *SQ;
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
  w->rn = calloc((size_t)(P->n + 1 + P->nprod), sizeof(double));
  w->mr = calloc((size_t)P->n, sizeof(double));
  w->H = calloc((size_t)2 * P->n * P->n, sizeof(double));
  w->SP = calloc((size_t)2 * P->n * P->n, sizeof(double));
  w->SQ = calloc((size_t)2 * P->n * P->n, sizeof(double));
  return w->slots && w->u && w->uj && w->A && w->rhs && w->x0 && w->col && w->W && w->dnode &&
         w->part && w->perm && w->rn && w->mr && w->H && w->SP && w->SQ;
}

static void work_free(work_t *w) {
  free(w->slots); free(w->u); free(w->uj); free(w->A); free(w->rhs); free(w->x0);
  free(w->col); free(w->W); free(w->dnode); free(w->part); free(w->perm);
  free(w->rn); free(w->mr); free(w->H); free(w->SP); free(w->SQ);
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

/* ---------------------------------------------------- the step rule ---- */

/* x^(1/k) for x >= 0 from IEEE double operations only (host and device
 * agree bit for bit): log2 from the binary exponent and the series of
 * ln((1+z)/(1-z)) to z^21, 2^y from the binary exponent and the Taylor
 * series of e^g to g^16, then one Newton step on r^k = x. */
double otr_root(double x, int k) {
  int e = 0, i;
  double m, z, z2, s, y, fl, g, p, r, rk1;
  if (x != x) return x;
  if (!(x > 0.0)) return 0.0;
  if (x == INFINITY || k == 1) return x;
  m = 2.0 * frexp(x, &e);
  z = (m - 1.0) / (m + 1.0);
  z2 = z * z;
  s = 1.0 / 21.0;
  s = 1.0 / 19.0 + z2 * s;
  s = 1.0 / 17.0 + z2 * s;
  s = 1.0 / 15.0 + z2 * s;
  s = 1.0 / 13.0 + z2 * s;
  s = 1.0 / 11.0 + z2 * s;
  s = 1.0 / 9.0 + z2 * s;
  s = 1.0 / 7.0 + z2 * s;
  s = 1.0 / 5.0 + z2 * s;
  s = 1.0 / 3.0 + z2 * s;
  s = 1.0 + z2 * s;
  y = ((double)(e - 1) + (z * s) * 2.8853900817779268) / (double)k;
  fl = floor(y);
  g = (y - fl) * 0.69314718055994531;
  p = 1.0;
  for (i = 16; i >= 1; i--) p = 1.0 + (g / (double)i) * p;
  r = ldexp(p, (int)fl);
  rk1 = 1.0;
  for (i = 1; i < k; i++) rk1 = rk1 * r;
  return r - (rk1 * r - x) / ((double)k * rk1);
}

/* complex double helpers of the step rule (the device's operation order) */
static void zmac(double *z, const double *x, const double *y) { /* z += x*y */
  z[0] = fma(x[0], y[0], z[0]);
  z[0] = fma(-x[1], y[1], z[0]);
  z[1] = fma(x[1], y[0], z[1]);
  z[1] = fma(x[0], y[1], z[1]);
}
static void zcjmac(double *z, const double *x, const double *y) { /* z += conj(x)*y */
  z[0] = fma(x[0], y[0], z[0]);
  z[0] = fma(x[1], y[1], z[0]);
  z[1] = fma(-x[1], y[0], z[1]);
  z[1] = fma(x[0], y[1], z[1]);
}
static void zmul(const double *x, const double *y, double *z) {
  const double re = fma(x[0], y[0], -(x[1] * y[1]));
  const double im = fma(x[0], y[1], x[1] * y[0]);
  z[0] = re;
  z[1] = im;
}
static double zrad(double re, double im) { return sqrt(re * re + im * im); }

/* largest eigenvalue of M^H M (sigma_max(M)^2) for M n x n complex,
 * interleaved (re, im) row-major: the Gram matrix scaled by its trace and
 * squared six times, then the Rayleigh quotient |M v|^2 / |v|^2 of the
 * column v of largest diagonal entry.  The Gram matrix and its powers are
 * Hermitian: entries i <= j are computed, entries i > j are their
 * conjugates.  P and Q are 2*n*n scratch doubles. */
double otr_sigma_max_sq(int n, const double *M, double *P, double *Q) {
  int i, j, k, s, jm;
  double tr, sc, num, den, best;
  for (i = 0; i < n; i++)
    for (j = i; j < n; j++) {
      double acc[2] = {0.0, 0.0};
      for (k = 0; k < n; k++) zcjmac(acc, M + 2 * (k * n + i), M + 2 * (k * n + j));
      P[2 * (i * n + j)] = acc[0];
      P[2 * (i * n + j) + 1] = acc[1];
      P[2 * (j * n + i)] = acc[0];
      P[2 * (j * n + i) + 1] = (i == j) ? acc[1] : -acc[1];
    }
  tr = 0.0;
  for (i = 0; i < n; i++) tr = tr + P[2 * (i * n + i)];
  if (tr != tr) return tr;
  if (!(tr > 0.0)) return 0.0;
  for (s = 0; s <= 6; s++) {
    sc = 1.0 / tr;
    for (i = 0; i < 2 * n * n; i++) P[i] = P[i] * sc;
    if (s == 6) break;
    for (i = 0; i < n; i++)
      for (j = i; j < n; j++) {
        double acc[2] = {0.0, 0.0};
        for (k = 0; k < n; k++) zmac(acc, P + 2 * (i * n + k), P + 2 * (k * n + j));
        Q[2 * (i * n + j)] = acc[0];
        Q[2 * (i * n + j) + 1] = acc[1];
        Q[2 * (j * n + i)] = acc[0];
        Q[2 * (j * n + i) + 1] = (i == j) ? acc[1] : -acc[1];
      }
    memcpy(P, Q, sizeof(double) * 2 * n * n);
    tr = 0.0;
    for (i = 0; i < n; i++) tr = tr + P[2 * (i * n + i)];
  }
  jm = 0;
  best = P[0];
  for (i = 1; i < n; i++)
    if (P[2 * (i * n + i)] > best) { best = P[2 * (i * n + i)]; jm = i; }
  num = 0.0;
  den = 0.0;
  for (i = 0; i < n; i++) {
    double y[2] = {0.0, 0.0};
    for (k = 0; k < n; k++) zmac(y, M + 2 * (i * n + k), P + 2 * (k * n + jm));
    num = num + y[0] * y[0] + y[1] * y[1];
  }
  for (k = 0; k < n; k++) {
    const double *v = P + 2 * (k * n + jm);
    den = den + v[0] * v[0] + v[1] * v[1];
  }
  return num / den;
}

/* leading limbs of an n x n matrix of K-limb complex numbers, interleaved */
static void lead_matrix(int K, int n, const double *A, double *M) {
  int e;
  for (e = 0; e < n * n; e++) {
    M[2 * e] = ZP(A, e)[0];
    M[2 * e + 1] = ZP(A, e)[K];
  }
}

/* the mixed residual mean_i |f_i| / (|f|_i + 1) at the point in the order-0
 * slots, after eval_point (rhs = -f, weights set): |f|_i is equation i with
 * the moduli of its coefficients (the components of a monomial summed
 * first) evaluated at the moduli of x (Standard_Mixed_Residuals.AbsVal and
 * Mixed_Residual, standard_mixed_residuals.adb:7-18,472-484; EvalCffRad,
 * standard_predictor_convolutions.adb:935-956), all in FP64 on the leading
 * limbs; segment sums meet as a butterfly */
static double mixed_residual(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n, D = P->deg + 1, G = P->G;
  double *rn = w->rn, mix = 0.0;
  int i, j, g, t;
  for (j = 0; j < n; j++) rn[j] = zrad(ZP(w->slots, j * D)[0], ZP(w->slots, j * D)[K]);
  rn[n] = 1.0;
  for (i = 0; i < P->nprod; i++) {
    const int *nd = P->nodes + 8 * i;
    rn[n + 1 + i] = rn[nd[7]] * rn[nd[2] / D];
  }
  for (i = 0; i < n; i++) {
    double part[64];
    for (g = 0; g < G; g++) {
      const int s = i * G + g;
      double acc = 0.0, gz[2] = {0.0, 0.0};
      for (t = P->seg_ptr[s]; t < P->seg_ptr[s + 1]; t++) {
        const int q = P->term_pair[t], m = P->pairs[4 * q + 3], r = P->pairs[4 * q];
        const double c[2] = {P->term_coef[t], P->term_coef[P->nterm + t]};
        const double wz[2] = {ZP(w->W, 3 * r)[0], ZP(w->W, 3 * r)[K]};
        double cw[2];
        zmul(c, wz, cw);
        gz[0] = gz[0] + cw[0];
        gz[1] = gz[1] + cw[1];
        if (t + 1 == P->seg_ptr[s + 1] || P->pairs[4 * P->term_pair[t + 1] + 3] != m) {
          acc = acc + zrad(gz[0], gz[1]) * rn[m];
          gz[0] = gz[1] = 0.0;
        }
      }
      part[g] = acc;
    }
    {
      int mm;
      for (mm = 1; mm < G; mm <<= 1)
        for (g = 0; g + mm < G; g += 2 * mm) part[g] = part[g] + part[g + mm];
    }
    w->mr[i] = part[0];
  }
  for (i = 0; i < n; i++)
    mix = mix + zrad(ZP(w->rhs, i)[0], ZP(w->rhs, i)[K]) / (w->mr[i] + 1.0);
  return mix / (double)n;
}

double otr_mixed_residual(const otr_program *P, int K, double t, const double *x) {
  work_t w = {0};
  double mix;
  if (!work_alloc(P, K, &w)) { work_free(&w); return -1.0; }
  load_point(P, &w, x, K, 1, 0);
  set_weights(P, &w, t);
  eval_point(P, &w, 0);
  mix = mixed_residual(P, &w);
  work_free(&w);
  return mix;
}

/* sum over the equations of sigma_max(Hessian_i)^2 (Distance's accsum,
 * standard_predictor_convolutions.adb:558-572): the program's constant for
 * equations of degree <= 2, else the Hessian at the order-0 point (right
 * after eval_point, every node at order 0) from the Hessian terms */
static double hessian_sum(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n, nn = n * (n + 1) / 2;
  double acc = 0.0;
  int i, a, b, h;
  for (i = 0; i < n; i++) {
    double s = P->hsig[i];
    if (s < 0.0) {
      for (a = 0; a < n; a++)
        for (b = a; b < n; b++) {
          const int e = i * nn + a * n - a * (a - 1) / 2 + (b - a);
          double z[2] = {0.0, 0.0};
          for (h = P->hent_ptr[e]; h < P->hent_ptr[e + 1]; h++) {
            const int term = P->hterms[3 * h], hp = P->hterms[3 * h + 1];
            const double mult = (double)P->hterms[3 * h + 2];
            const int r = P->hpairs[2 * hp], off = P->hpairs[2 * hp + 1];
            const double c[2] = {P->term_coef[term], P->term_coef[P->nterm + term]};
            const double wz[2] = {ZP(w->W, 3 * r)[0], ZP(w->W, 3 * r)[K]};
            const double nd[2] = {ZP(w->slots, off)[0], ZP(w->slots, off)[K]};
            double cw[2], v[2];
            zmul(c, wz, cw);
            zmul(cw, nd, v);
            z[0] = z[0] + mult * v[0];
            z[1] = z[1] + mult * v[1];
          }
          w->H[2 * (a * n + b)] = z[0];
          w->H[2 * (a * n + b) + 1] = z[1];
          w->H[2 * (b * n + a)] = z[0];
          w->H[2 * (b * n + a) + 1] = z[1];
        }
      s = sqrt(otr_sigma_max_sq(n, w->H, w->SP, w->SQ));
    }
    acc = acc + s * s;
  }
  return acc;
}

double otr_hessian_sum(const otr_program *P, int K, double t, const double *x) {
  work_t w = {0};
  double hs2;
  if (!work_alloc(P, K, &w)) { work_free(&w); return -1.0; }
  load_point(P, &w, x, K, 1, 0);
  set_weights(P, &w, t);
  eval_point(P, &w, 1);
  hs2 = hessian_sum(P, &w);
  work_free(&w);
  return hs2;
}

/* Taylor coefficients x_kb..x_{ke-1} (A holds the inverse Jacobian, the
 * coefficients below kb are computed) */
static void series_orders(const otr_program *P, work_t *w, int kb, int ke) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  double xk[16 * 64];
  int k, j;
  for (k = kb; k < ke; k++) {
    for (j = 0; j < n; j++) czero(K, ZP(w->slots, j * D + k));
    product_pass(P, w, k);
    residual(P, w, k);
    apply_inverse(P, w, xk);
    for (j = 0; j < n; j++) ccopy(K, ZP(xk, j), ZP(w->slots, j * D + k));
    correction_pass(P, w, k);
  }
}

/* Taylor coefficients x_1..x_{ke-1} (ke <= deg + 1) at the point in the
 * order-0 slots (weights set at t); *hs2 the Hessian sum at the point; 0 on
 * a singular Jacobian, which then stays in A, else A holds its inverse */
static int predictor_series(const otr_program *P, work_t *w, double *hs2, int ke) {
  eval_point(P, w, 1);
  *hs2 = hessian_sum(P, w);
  if (!gj_invert(P, w)) return 0;
  series_orders(P, w, 1, ke);
  return 1;
}

/* the pole radius of Fabry's ratio with the given offset, minimized over
 * the components (Convergence_Radius_Estimates.Fabry, convergence_radius_
 * estimates.adb:28-38 Is_Zero, :138-163 per series, :354-390 over the
 * vector; the first component whose c_kd is not zero starts the minimum),
 * in FP64 on the leading limbs; returns 1 when every c_kd is zero */
static int fabry_radius(const otr_program *P, work_t *w, int kd, double *r) {
  const int K = w->K, D = P->deg + 1;
  int j, fail = 1;
  *r = 0.0;
  for (j = 0; j < P->n; j++) {
    const double *a = ZP(w->slots, j * D + kd - 1), *b = ZP(w->slots, j * D + kd);
    const double ar = a[0], ai = a[K], br = b[0], bi = b[K];
    double nrm, zr, zi, rad;
    if (br + 1.0 == 1.0 && bi + 1.0 == 1.0) continue;
    nrm = br * br + bi * bi;
    zr = (ar * br + ai * bi) / nrm;
    zi = (ai * br - ar * bi) / nrm;
    rad = zrad(zr, zi);
    if (fail) { *r = rad; fail = 0; }
    else if (rad < *r) *r = rad;
  }
  return fail;
}

/* The [L/M] Pade coefficients of one component in place (as
 * otr_pade_coefficients; M = 0 keeps the Taylor polynomial) and the error
 * estimate of Homotopy_Pade_Approximants.Solution_Error_Estimate
 * (homotopy_pade_approximants.adb:558-568 Threshold_Index, :597-622): top =
 * the largest i <= L+M+2 with |c_i| > 1e-5 (|.| = |re| + |im|, c_0 = x0),
 * est = c_top + sum_{i=1..min(M,top)} d_i c_{top-i}, est = num_top - est when
 * top <= L, est = 1 when top < 0 or |est| < 1e-5; leading limbs in err */
void otr_pade_error(int K, double *c, int L, int M, double *x0, double *err) {
  double orig[9][16], res[16], one[16];
  int top = -1, i;
  for (i = L + M + 2; i >= 0; i--) {
    const double v = absz(K, i == 0 ? x0 : ZP(c, i));
    if (v > 1e-5) { top = i; break; }
  }
  for (i = 0; i <= M && i <= top; i++) ccopy(K, (top - i == 0) ? x0 : ZP(c, top - i), orig[i]);
  if (M >= 1) otr_pade_coefficients(K, c, L, M, x0);
  cset(K, one, 1.0, 0.0);
  if (top < 0) {
    ccopy(K, one, res);
  } else {
    ccopy(K, orig[0], res);
    for (i = 1; i <= M; i++) {
      if (i > top) break;
      cadd_mul(K, ZP(c, L + i), orig[i], res);
    }
    if (top <= L) omd_csub(K, top == 0 ? x0 : ZP(c, top), res, res);
  }
  if (absz(K, res) < 1e-5) ccopy(K, one, res);
  err[0] = res[0];
  err[1] = res[K];
}

/* Newton corrector with the mixed residual (LU_Newton_Steps,
 * standard_inlined_newton_circuits.adb:388-422, each step as LU_Newton_Step
 * :157-175): converged when the mixed residual and max|dx| are both <= tol,
 * then one extra step unless the residual or the update is zero; stops when
 * both grow.  Returns 1 converged, 0 not, -1 singular Jacobian.
 * evaluated: the products and the residual at the point are in place. */
static int lu_newton_steps(const otr_program *P, work_t *w, int maxit, double tol, int extra,
                           int evaluated, double *dxl, int *count, int *nfirst, double *dxh) {
  const int K = w->K, D = P->deg + 1;
  double dx[16 * 64], prev_err = 1.0, prev_mix = 1.0;
  int k, j, cnt = 0;
  for (k = 1; k <= maxit + extra; k++) {
    double err, res, mix;
    ++*count;
    if (k == 1 && !evaluated) { eval_point(P, w, 1); ++*nfirst; } else jacobian(P, w);
    if (!gj_invert(P, w)) return -1;
    apply_inverse(P, w, dx);
    for (j = 0; j < P->n; j++) omd_cadd(K, ZP(w->slots, j * D), ZP(dx, j), ZP(w->slots, j * D));
    err = max_abs(K, dx, P->n);
    *dxl = err;
    if (dxh) { dxh[3] = dxh[2]; dxh[2] = dxh[1]; dxh[1] = dxh[0]; dxh[0] = err; }
    eval_point(P, w, 0);
    res = max_abs(K, w->rhs, P->n);
    mix = mixed_residual(P, w);
    if (mix <= tol && err <= tol) {
      if (cnt == extra || res == 0.0 || err == 0.0) return 1;
      cnt++;
    } else if (k > 1 && mix > prev_mix && err > prev_err) {
      return 0;
    }
    prev_err = err;
    prev_mix = mix;
  }
  return 0;
}

/* a + b + c minimum of Three_Way_Minima.Minimum (three_way_minima.adb:49-62) */
static double minimum3(double a, double b, double c, int *which) {
  if (a < b) {
    if (a < c) { *which = 0; return a; }
    *which = 2;
    return c;
  }
  if (b < c) { *which = 1; return b; }
  *which = 2;
  return c;
}

#define OTR_TRACE 20

/* one path (both attempts of the learned controller); trace (may be NULL)
 * receives OTR_TRACE doubles per step of the last attempt, at most cap:
 * t, Fabry radius (-1 when every c_kd is zero), pole step, curvature
 * step, eta, nrm, Minimum (0 pole, 1 curvature, 2 maxstep), the step
 * of the rule, the step after the controller and Bounded_Update, the
 * step at the end, feedback halvings, corrector halvings, last feedback
 * mixed residual, Newton iterations of the last corrector, accepted (1)
 * or not, sigma_min(J), Hessian sum, t after the step, candidate, series
 * degree of the step */
static void track_path(const otr_program *P, work_t *w, int npaths, int path, const double *start,
                       int KS, const double *tstart, double *endp, double *tend, int *status,
                       int *nsteps, int *nrej, double *dxlast, int *work, const otr_params *prm,
                       const otr_controller *ctl, double *trace, int cap, int *ntrace) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  const int mode = ctl ? ctl->mode : 0;
  const int pden = prm->pade_dendeg;
  const int pnum = (pden == 0 && prm->pade_numdeg == 0) ? D - 3 : prm->pade_numdeg;
  const int d0 = (ctl && ctl->deg0 ? ctl->deg0 : P->deg) + 1;
  double t = 0.0, dxl = 0.0, hprev = prm->maxstep;
  double dxh[4] = {1.0, 1.0, 1.0, 1.0};
  int st = ST_RUNNING, steps = 0, rej = 0, j, k, npred = 0, nnewt = 0, nfeval = 0, ncorr = 0;
  int nbpole = 0, nbhess = 0, nbmaxm = 0, nfirst = 0;
  int rejprev = 0, nlog = 0, attempt, pmode = mode;
  /* a path the learned controller (mode 1) does not finish, stopped at a
     t below retry_below (0: 0.99), is tracked again from its start with
     the plain rule; the work counters add up both attempts, the other
     outputs are the last attempt's */
  const double retry = (ctl && ctl->retry_below > 0.0) ? ctl->retry_below : 0.99;
  for (attempt = 0; attempt < 2; attempt++) {
    pmode = (attempt == 0) ? mode : 0;
    t = tstart ? tstart[path] : 0.0;
    dxl = 0.0; hprev = prm->maxstep;
    for (j = 0; j < 4; j++) dxh[j] = 1.0;
    st = ST_RUNNING; steps = 0; rej = 0; rejprev = 0;
    if (ntrace) *ntrace = 0;
    load_point(P, w, start, KS, npaths, path);
    while (st == ST_RUNNING && t < 1.0) {
      double r, pole, curv, eta, nrm, hs2, sig1, step, tn, mix = 0.0, rule, step1;
      int Dk, pnumk;
      double err[128];
      const double t0 = t;
      int pfail, which, evaluated, cand = 0, rej0, nnewt0, nfb, accepted = 0, nit = 0;
      set_weights(P, w, t);
      ++npred;
      /* the controller chooses after d0 coefficients; its candidate's
         series degree may add orders (the plain rule: all D at once) */
      Dk = pmode != 0 ? d0 : D;
      if (!predictor_series(P, w, &hs2, Dk)) { st = ST_SINGULAR; break; }
      /* pole radius (LU_Prediction, standard_predictor_convolutions.adb:1168-1172) */
      pfail = fabry_radius(P, w, Dk - 1 - prm->fabry_offset, &r);
      /* the controller reads the Taylor series (before the Pade
         coefficients replace it) */
      if (pmode != 0) {
        double feat[OTR_NFEAT];
        feat[0] = otr_lg2(pfail ? prm->maxstep : r);
        feat[1] = otr_lg2(hprev);
        feat[2] = t;
        for (k = 1; k <= 8; k++) {
          double c = 0.0;
          for (j = 0; j < n; j++) c = fmax(c, k < Dk ? absz(K, ZP(w->slots, j * D + k)) : 0.0);
          feat[2 + k] = otr_lg2(c + 1e-300);
        }
        for (j = 0; j < 4; j++) feat[11 + j] = otr_lg2(dxh[j] + 1e-300);
        feat[15] = (double)rejprev;
        if (pmode == 1) cand = mlp_choose(ctl, feat);
        else cand = (int)(mix_hash(ctl->seed, (unsigned)path, (unsigned)(npred - 1)) % (unsigned)ctl->ncand);
        if (mode == 2 && nlog < ctl->log_capacity) {
          for (j = 0; j < OTR_NFEAT; j++)
            ctl->log_x[((size_t)path * ctl->log_capacity + nlog) * OTR_NFEAT + j] = (float)feat[j];
          ctl->log_c[(size_t)path * ctl->log_capacity + nlog] = cand;
        }
        {
          const int dc = (ctl->degree[cand] ? ctl->degree[cand] : P->deg) + 1;
          if (dc > Dk) {
            series_orders(P, w, Dk, dc);
            Dk = dc;
            pfail = fabry_radius(P, w, Dk - 1 - prm->fabry_offset, &r);
          }
        }
      }
      pole = pfail ? prm->maxstep : prm->pbeta * r;
      /* the [L/M] predictor of a series of Dk coefficients */
      pnumk = pnum - (D - Dk);
      /* predictor coefficients and their error estimates (Solution_Error) */
      for (j = 0; j < n; j++) ccopy(K, ZP(w->slots, j * D), ZP(w->x0, j));
      for (j = 0; j < n; j++)
        otr_pade_error(K, ZP(w->slots, j * D), pnumk, pden, ZP(w->x0, j), err + 2 * j);
      nrm = 0.0;
      for (j = 0; j < n; j++) nrm = nrm + err[2 * j] * err[2 * j] + err[2 * j + 1] * err[2 * j + 1];
      nrm = sqrt(nrm);
      /* curvature step (Hesse_Pade, Distance, Step_Distance of
         series_and_predictors.adb:813-825): sigma_min(J) = 1/sigma_max(J^{-1}) */
      lead_matrix(K, n, w->A, w->H);
      sig1 = 1.0 / sqrt(otr_sigma_max_sq(n, w->H, w->SP, w->SQ));
      eta = (2.0 * sig1) / sqrt(hs2);
      curv = otr_root((prm->cbeta * eta) / nrm, Dk - 1);
      rule = minimum3(pole, curv, prm->maxstep, &which);
      if (which == 0) ++nbpole; else if (which == 1) ++nbhess; else ++nbmaxm;
      step = rule;
      if (pmode != 0) {
        step *= ctl->mult[cand];
        if (ctl->pole_cap > 0.0 && !pfail && r > 0.0) step = fmin(step, ctl->pole_cap * r);
      }
      /* Bounded_Update (three_way_minima.adb:96-109) */
      if (t + step >= 1.0 - prm->minstep) { step = 1.0 - t; tn = 1.0; }
      else tn = t + step;
      step1 = step;
      rej0 = rej;
      nnewt0 = nnewt;
      /* Predictor_Feedback (standard_predictor_convolutions.adb:958-993) */
      evaluated = 0;
      nfb = 0;
      for (;;) {
        for (j = 0; j < n; j++)
          otr_pade_evaluate(K, ZP(w->slots, j * D), pnumk, pden, ZP(w->x0, j), step, ZP(w->slots, j * D));
        set_weights(P, w, tn);
        ++nfeval;
        eval_point(P, w, 0);
        mix = mixed_residual(P, w);
        if (mix < prm->alpha) { evaluated = 1; break; }
        step = step / 2.0;
        ++rej;
        ++nfb;
        tn = t + step;
        if (step < prm->minstep) break;
      }
      /* corrector (Predictor_Corrector_Loop, predictor_corrector_loops.adb:59-81) */
      for (;;) {
        int nr, it0 = nnewt;
        if (!evaluated) set_weights(P, w, tn);
        nr = lu_newton_steps(P, w, prm->corsteps, prm->tolcor, 1, evaluated, &dxl, &nnewt,
                             &nfirst, dxh);
        ++ncorr;
        nit = nnewt - it0;
        if (nr < 0) st = ST_SINGULAR;
        else if (nr == 1) { t = tn; hprev = step; accepted = 1; break; }
        else {
          step = step / 2.0;
          ++rej;
          if (step < prm->minstep) st = ST_MINSTEP;
        }
        if (st != ST_RUNNING) {
          /* the path ends at its last accepted point (t, x0) */
          for (j = 0; j < n; j++) ccopy(K, ZP(w->x0, j), ZP(w->slots, j * D));
          break;
        }
        tn = t + step;
        for (j = 0; j < n; j++)
          otr_pade_evaluate(K, ZP(w->slots, j * D), pnumk, pden, ZP(w->x0, j), step, ZP(w->slots, j * D));
        evaluated = 0;
      }
      if (trace && ntrace && *ntrace < cap) {
        double *q = trace + (size_t)OTR_TRACE * *ntrace;
        q[0] = t0;
        q[1] = pfail ? -1.0 : r;
        q[2] = pole; q[3] = curv; q[4] = eta; q[5] = nrm; q[6] = which; q[7] = rule;
        q[8] = step1; q[9] = step; q[10] = nfb; q[11] = rej - rej0 - nfb; q[12] = mix;
        q[13] = nit; q[14] = accepted; q[15] = sig1; q[16] = hs2; q[17] = t; q[18] = cand;
        q[19] = Dk - 1;
        ++*ntrace;
      }
      rejprev = rej - rej0;
      if (mode == 2 && nlog < ctl->log_capacity) {
        /* accepted at the first attempt: the Newton iterations it took */
        ctl->log_y[(size_t)path * ctl->log_capacity + nlog] =
            (rej == rej0 && st == ST_RUNNING) ? nnewt - nnewt0 : 0;
        ++nlog;
      }
      if (st != ST_RUNNING) break;
      ++steps;
      if (max_abs_x(P, w) > prm->divergence) { st = ST_DIVERGED; break; }
      if (steps >= prm->maxsteps && t < 1.0) { st = ST_MAXSTEPS; break; }
    }
    if (st == ST_RUNNING) {
      int nr;
      set_weights(P, w, 1.0);
      nr = lu_newton_steps(P, w, prm->endsteps, prm->tolend, 1, 0, &dxl, &nnewt, &nfirst, NULL);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
    }
    if (!(pmode == 1 && st != ST_SUCCESS && t < retry)) break;
  }
  store_point(P, w, endp, npaths, path);
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
    work[4 * npaths + path] = nbpole;
    work[5 * npaths + path] = nbhess;
    work[6 * npaths + path] = nbmaxm;
    work[7 * npaths + path] = nfirst;
  }
  if (mode == 2) ctl->log_n[path] = nlog;
}

void otr_track(const otr_program *P, int K, int npaths, int p0, int p1, const double *start,
               int KS, const double *tstart, double *endp, double *tend, int *status,
               int *nsteps, int *nrej, double *dxlast, int *work, const otr_params *prm,
               const otr_controller *ctl) {
  work_t w = {0};
  int path;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = p0; path < p1; path++)
    track_path(P, &w, npaths, path, start, KS, tstart, endp, tend, status, nsteps, nrej, dxlast,
               work, prm, ctl, NULL, 0, NULL);
  work_free(&w);
}

int otr_track_trace(const otr_program *P, int K, int npaths, int path, const double *start,
                    int KS, double *endp, double *tend, int *status, int *nsteps, int *nrej,
                    double *dxlast, int *work, const otr_params *prm, const otr_controller *ctl,
                    double *trace, int cap) {
  work_t w = {0};
  int nt = 0;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return -1; }
  track_path(P, &w, npaths, path, start, KS, NULL, endp, tend, status, nsteps, nrej, dxlast,
             work, prm, ctl, trace, cap, &nt);
  work_free(&w);
  return nt;
}

typedef struct {
  const otr_program *P;
  int K, npaths, p0, p1, KS;
  const double *start, *tstart;
  double *endp, *tend, *dxlast;
  int *status, *nsteps, *nrej, *work;
  const otr_params *prm;
  const otr_controller *ctl;
} job_t;

static void *run_job(void *arg) {
  job_t *j = (job_t *)arg;
  otr_track(j->P, j->K, j->npaths, j->p0, j->p1, j->start, j->KS, j->tstart, j->endp, j->tend,
            j->status, j->nsteps, j->nrej, j->dxlast, j->work, j->prm, j->ctl);
  return NULL;
}

void otr_track_mt(const otr_program *P, int K, int npaths, const double *start, int KS,
                  const double *tstart, double *endp, double *tend, int *status, int *nsteps,
                  int *nrej,
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
    jobs[i].start = start; jobs[i].tstart = tstart; jobs[i].endp = endp; jobs[i].tend = tend; jobs[i].dxlast = dxlast;
    jobs[i].status = status; jobs[i].nsteps = nsteps; jobs[i].nrej = nrej; jobs[i].prm = prm;
    jobs[i].work = work;
    jobs[i].ctl = ctl;
    pthread_create(&th[i], NULL, run_job, &jobs[i]);
  }
  for (i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
}

void otr_refine(const otr_program *P, int K, int npaths, double t, const double *xin, int KS,
                double *xout, int iters, double tol, double *dxlast, double *res, int *status) {
  work_t w = {0};
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

/* Where a path goes as t -> 1, from the Taylor series of x(t+s) at its last
 * point (the question the endgames of Deformations/End_Games answer: a
 * regular endpoint, a singular one with its winding number, or divergence
 * to infinity).  Ratio method on the last three coefficients of every
 * component, FP64 on the leading limbs: with q_k = c_{k+1}/c_k at k = D-2,
 * the nearest singularity of the component lies at
 * s* = 1 / ((k+1) q_k - k q_{k-1}) and the component behaves like
 * (s* - s)^alpha with alpha = k - (k+1) s* q_k (both exact for
 * c (s* - s)^alpha).  The component whose last coefficient weighs most at
 * the distance 1 - t (|c_{D-1}| (1-t)^(D-1)) gives alpha = Re and
 * rho = Re s*; components with a zero coefficient (Fabry's Is_Zero) are
 * passed over, and when all are, alpha = rho = +inf (no singularity
 * within reach).  rco = 1 / (sigma_max(J) sigma_max(J^{-1})) of the
 * Jacobian at the point (0 when singular), xmax = max_i |x_i|. */
static int is_zero2(const double *z) { return z[0] + 1.0 == 1.0 && z[1] + 1.0 == 1.0; }

static void zdiv(const double *a, const double *b, double *z) {
  const double nrm = b[0] * b[0] + b[1] * b[1];
  z[0] = (a[0] * b[0] + a[1] * b[1]) / nrm;
  z[1] = (a[1] * b[0] - a[0] * b[1]) / nrm;
}

static void endgame_point(const otr_program *P, work_t *w, double t, double *alpha, double *rho,
                          double *rco, double *xmax) {
  const int K = w->K, n = P->n, D = P->deg + 1, k = D - 2;
  double M[2 * 64 * 64], lamJ, lamI, hs2, best = -1.0, dist = 1.0 - t, xm = 0.0;
  int j, i;
  set_weights(P, w, t);
  eval_point(P, w, 1);
  lead_matrix(K, n, w->A, M);
  lamJ = otr_sigma_max_sq(n, M, w->SP, w->SQ);
  *alpha = INFINITY;
  *rho = INFINITY;
  if (!predictor_series(P, w, &hs2, P->deg + 1)) {
    *rco = 0.0;
  } else {
    lead_matrix(K, n, w->A, M);
    lamI = otr_sigma_max_sq(n, M, w->SP, w->SQ);
    *rco = 1.0 / sqrt(lamJ * lamI);
  }
  for (j = 0; j < n; j++) {
    const double *z = ZP(w->slots, j * D);
    const double r = zrad(z[0], z[K]);
    if (r > xm) xm = r;
  }
  *xmax = xm;
  if (*rco == 0.0 || k < 2) return;
  for (j = 0; j < n; j++) {
    const double *c0 = ZP(w->slots, j * D + k - 1), *c1 = ZP(w->slots, j * D + k),
                 *c2 = ZP(w->slots, j * D + k + 1);
    double a0[2] = {c0[0], c0[K]}, a1[2] = {c1[0], c1[K]}, a2[2] = {c2[0], c2[K]};
    double q0[2], q1[2], den[2], one[2] = {1.0, 0.0}, ss[2], sq[2], pw = 1.0, wt;
    if (is_zero2(a0) || is_zero2(a1) || is_zero2(a2)) continue;
    zdiv(a1, a0, q0);
    zdiv(a2, a1, q1);
    den[0] = (double)(k + 1) * q1[0] - (double)k * q0[0];
    den[1] = (double)(k + 1) * q1[1] - (double)k * q0[1];
    if (den[0] == 0.0 && den[1] == 0.0) continue;
    zdiv(one, den, ss);
    zmul(ss, q1, sq);
    for (i = 0; i <= k; i++) pw = pw * dist;
    wt = zrad(a2[0], a2[1]) * pw;
    if (wt > best) {
      best = wt;
      *alpha = (double)k - (double)(k + 1) * sq[0];
      *rho = ss[0];
    }
  }
}

void otr_endgame(const otr_program *P, int K, int npaths, const double *x, int KS,
                 const double *tend, double *alpha, double *rho, double *rco, double *xmax) {
  work_t w = {0};
  int path;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = 0; path < npaths; path++) {
    load_point(P, &w, x, KS, npaths, path);
    endgame_point(P, &w, tend[path], alpha + path, rho + path, rco + path, xmax + path);
  }
  work_free(&w);
}

void otr_eval(const otr_program *P, int K, double t, const double *x, double *val, double *jac) {
  const int n = P->n;
  work_t w = {0};
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
  work_t w = {0};
  int j, k, i, ok;
  if (P->n > 64 || !work_alloc(P, K, &w)) { work_free(&w); return 0; }
  load_point(P, &w, x, K, 1, 0);
  set_weights(P, &w, t);
  {
    double hs2;
    ok = predictor_series(P, &w, &hs2, D);
  }
  for (j = 0; j < n; j++)
    for (k = 0; k < D; k++)
      for (i = 0; i < K; i++) {
        series[((size_t)i * n + j) * D + k] = ZP(w.slots, j * D + k)[i];
        series[((size_t)(K + i) * n + j) * D + k] = ZP(w.slots, j * D + k)[K + i];
      }
  work_free(&w);
  return ok;
}
