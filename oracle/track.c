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
 * and evaluates the homotopy through the program of phcpack_amd/homotopy.py
 * in the operation order the device kernel uses, value by value.
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

typedef struct {
  int K;
  double *slots, *u, *uj, *A, *part, *rhs, *x0, *W, *dnode, *invd;
  int *piv;
} work_t;

#define ZS(K) (2 * (K))
#define ZP(a, i) ((a) + (size_t)(i) * ZS(K))

static double *nodec(const otr_program *P, work_t *w, int m, int a) {
  const int K = w->K;
  return ZP(w->slots, P->off[m] + a % P->slen[m]);
}

static void load_coef(int K, const double *base, int m, int j, double *z) {
  int i;
  for (i = 0; i < K; i++) {
    z[i] = base[(size_t)i * m + j];
    z[K + i] = base[(size_t)(8 + i) * m + j];
  }
}

static void czero(int K, double *z) { memset(z, 0, sizeof(double) * 2 * K); }
static void cset(int K, double *z, double re) { czero(K, z); z[0] = re; }
static void ccopy(int K, const double *a, double *z) { memmove(z, a, sizeof(double) * 2 * K); }
static void cscale_d(int K, const double *a, double s, double *z) {
  omd_mul_d(K, a, s, z);
  omd_mul_d(K, a + K, s, z + K);
}
static void cneg(int K, double *z) {
  int i;
  for (i = 0; i < 2 * K; i++) z[i] = -z[i];
}
/* z -= x*y in the kernel's order (C<K>::msub) */
static void cmsub(int K, const double *x, const double *y, double *z) {
  double t[8];
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

static void set_weights(const otr_program *P, work_t *w, double t) {
  const int K = w->K;
  double o0[16], o1[16], o2[16], o2t[16], tmp[16];
  int r;
  for (r = 0; r < P->ncomp; r++) {
    load_coef(K, P->weights, 3 * P->ncomp, 3 * r, o0);
    load_coef(K, P->weights, 3 * P->ncomp, 3 * r + 1, o1);
    load_coef(K, P->weights, 3 * P->ncomp, 3 * r + 2, o2);
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
  int lv, i, q;
  for (lv = 1; lv <= P->nlev; lv++)
    for (i = P->lev_ptr[lv - 1]; i < P->lev_ptr[lv]; i++) {
      const int a = P->left[i], b = P->right[i];
      omd_cmul(K, nodec(P, w, a, 0), nodec(P, w, b, k), acc);
      for (q = 1; q <= k; q++) omd_cmac(K, nodec(P, w, a, q), nodec(P, w, b, k - q), acc);
      ccopy(K, acc, nodec(P, w, P->n + 1 + i, k));
    }
}

static void correction_pass(const otr_program *P, work_t *w, int k) {
  const int K = w->K;
  double d[16], zero[16], *da;
  int lv, i;
  czero(K, zero);
  for (lv = 1; lv <= P->nlev; lv++)
    for (i = P->lev_ptr[lv - 1]; i < P->lev_ptr[lv]; i++) {
      const int a = P->left[i], b = P->right[i];
      da = (a < P->n) ? nodec(P, w, a, k) : (a == P->n ? zero : ZP(w->dnode, a - P->n - 1));
      omd_cmul(K, nodec(P, w, a, 0), nodec(P, w, b, k), d);
      omd_cmac(K, da, nodec(P, w, b, 0), d);
      ccopy(K, d, ZP(w->dnode, i));
      omd_cadd(K, nodec(P, w, P->n + 1 + i, k), d, nodec(P, w, P->n + 1 + i, k));
    }
}

static void residual(const otr_program *P, work_t *w, int k) {
  const int K = w->K;
  double acc[16], c[16];
  int q, l, s, t, i;
  for (q = 0; q < P->npair; q++) {
    const int r = P->pair_comp[q], m = P->pair_node[q];
    omd_cmul(K, ZP(w->W, 3 * r), nodec(P, w, m, k), acc);
    for (l = 1; l <= P->wdeg && l <= k; l++)
      omd_cmac(K, ZP(w->W, 3 * r + l), nodec(P, w, m, k - l), acc);
    ccopy(K, acc, ZP(w->u, q));
  }
  for (s = 0; s < P->nseg; s++) {
    czero(K, acc);
    for (t = P->seg_ptr[s]; t < P->seg_ptr[s + 1]; t++) {
      load_coef(K, P->term_coef, P->nterm, t, c);
      omd_cmac(K, c, ZP(w->u, P->term_pair[t]), acc);
    }
    ccopy(K, acc, ZP(w->part, s));
  }
  for (i = 0; i < P->n; i++) {
    ccopy(K, ZP(w->part, P->eq_seg[i]), acc);
    for (s = P->eq_seg[i] + 1; s < P->eq_seg[i + 1]; s++) omd_cadd(K, acc, ZP(w->part, s), acc);
    ccopy(K, acc, ZP(w->rhs, i));
  }
}

static void jacobian(const otr_program *P, work_t *w) {
  const int K = w->K;
  double acc[16], c[16];
  int q, e, t;
  for (q = 0; q < P->njpair; q++)
    omd_cmul(K, ZP(w->W, 3 * P->jpair_comp[q]), nodec(P, w, P->jpair_node[q], 0), ZP(w->uj, q));
  for (e = 0; e < P->n * P->n; e++) {
    czero(K, acc);
    for (t = P->jent_ptr[e]; t < P->jent_ptr[e + 1]; t++) {
      load_coef(K, P->jterm_coef, P->njterm, t, c);
      omd_cmac(K, c, ZP(w->uj, P->jterm_pair[t]), acc);
    }
    ccopy(K, acc, ZP(w->A, e));
  }
}

static int lu_factor(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n;
  double one[16], inv[16], tmp[16];
  int j, i, c, p;
  cset(K, one, 1.0);
  for (j = 0; j < n; j++) {
    double amax = absz(K, ZP(w->A, j * n + j));
    p = j;
    for (i = j + 1; i < n; i++) {
      const double v = absz(K, ZP(w->A, i * n + j));
      if (v > amax) { amax = v; p = i; }
    }
    if (!(amax > 0.0)) return 0;
    if (p != j)
      for (c = 0; c < n; c++) {
        ccopy(K, ZP(w->A, j * n + c), tmp);
        ccopy(K, ZP(w->A, p * n + c), ZP(w->A, j * n + c));
        ccopy(K, tmp, ZP(w->A, p * n + c));
      }
    w->piv[j] = p;
    omd_cdiv(K, one, ZP(w->A, j * n + j), inv);
    for (i = j + 1; i < n; i++) omd_cmul(K, ZP(w->A, i * n + j), inv, ZP(w->A, i * n + j));
    ccopy(K, inv, ZP(w->invd, j));
    for (i = j + 1; i < n; i++)
      for (c = j + 1; c < n; c++)
        cmsub(K, ZP(w->A, i * n + j), ZP(w->A, j * n + c), ZP(w->A, i * n + c));
  }
  return 1;
}

static void lu_solve(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n;
  double tmp[16];
  int j, i;
  for (j = 0; j < n; j++) {
    const int p = w->piv[j];
    if (p != j) {
      ccopy(K, ZP(w->rhs, j), tmp);
      ccopy(K, ZP(w->rhs, p), ZP(w->rhs, j));
      ccopy(K, tmp, ZP(w->rhs, p));
    }
  }
  for (j = 0; j < n; j++)
    for (i = j + 1; i < n; i++) cmsub(K, ZP(w->A, i * n + j), ZP(w->rhs, j), ZP(w->rhs, i));
  for (j = n - 1; j >= 0; j--) {
    omd_cmul(K, ZP(w->rhs, j), ZP(w->invd, j), ZP(w->rhs, j));
    for (i = 0; i < j; i++) cmsub(K, ZP(w->A, i * n + j), ZP(w->rhs, j), ZP(w->rhs, i));
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
  double m = 0.0;
  int j;
  for (j = 0; j < P->n; j++) m = fmax(m, absz(w->K, nodec(P, w, j, 0)));
  return m;
}

static int newton(const otr_program *P, work_t *w, int iters, double tol, double *dxl,
                  int *count) {
  const int K = w->K;
  int it, i, j;
  for (it = 0; it < iters; it++) {
    double dx;
    ++*count;
    eval_point(P, w, 1);
    if (!lu_factor(P, w)) return -1;
    for (i = 0; i < P->n; i++) cneg(K, ZP(w->rhs, i));
    lu_solve(P, w);
    dx = max_abs(K, w->rhs, P->n);
    for (j = 0; j < P->n; j++) omd_cadd(K, nodec(P, w, j, 0), ZP(w->rhs, j), nodec(P, w, j, 0));
    *dxl = dx;
    if (dx <= tol * (1.0 + max_abs_x(P, w))) return 1;
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
  w->part = calloc((size_t)P->nseg, zs);
  w->rhs = calloc((size_t)P->n, zs);
  w->x0 = calloc((size_t)P->n, zs);
  w->W = calloc((size_t)3 * P->ncomp, zs);
  w->dnode = calloc((size_t)(P->nprod > 0 ? P->nprod : 1), zs);
  w->invd = calloc((size_t)P->n, zs);
  w->piv = calloc((size_t)P->n, sizeof(int));
  return w->slots && w->u && w->uj && w->A && w->part && w->rhs && w->x0 && w->W && w->dnode &&
         w->invd && w->piv;
}

static void work_free(work_t *w) {
  free(w->slots); free(w->u); free(w->uj); free(w->A); free(w->part); free(w->rhs);
  free(w->x0); free(w->W); free(w->dnode); free(w->invd); free(w->piv);
}

static void load_point(const otr_program *P, work_t *w, const double *x, int KS, int npaths,
                       int path) {
  const int K = w->K, n = P->n;
  int j, i;
  for (j = 0; j < n; j++) {
    double *z = nodec(P, w, j, 0);
    czero(K, z);
    for (i = 0; i < K && i < KS; i++) {
      z[i] = x[(size_t)i * npaths * n + (size_t)path * n + j];
      z[K + i] = x[(size_t)(KS + i) * npaths * n + (size_t)path * n + j];
    }
  }
}

static void store_point(const otr_program *P, work_t *w, double *x, int npaths, int path) {
  const int K = w->K, n = P->n;
  int j, i;
  for (j = 0; j < n; j++) {
    const double *z = nodec(P, w, j, 0);
    for (i = 0; i < K; i++) {
      x[(size_t)i * npaths * n + (size_t)path * n + j] = z[i];
      x[(size_t)(K + i) * npaths * n + (size_t)path * n + j] = z[K + i];
    }
  }
}

static void set_one(const otr_program *P, work_t *w) {
  const int K = w->K, D = P->deg + 1;
  int a;
  for (a = 0; a < D; a++) {
    if (a == 0) cset(K, nodec(P, w, P->n, a), 1.0);
    else czero(K, nodec(P, w, P->n, a));
  }
}

/* Taylor coefficients x_1..x_deg at the point in the order-0 slots;
 * returns 0 on a singular Jacobian */
static int predictor_series(const otr_program *P, work_t *w) {
  const int K = w->K, n = P->n, D = P->deg + 1;
  int k, j, i;
  eval_point(P, w, 1);
  if (!lu_factor(P, w)) return 0;
  for (k = 1; k < D; k++) {
    for (j = 0; j < n; j++) czero(K, nodec(P, w, j, k));
    product_pass(P, w, k);
    residual(P, w, k);
    for (i = 0; i < n; i++) cneg(K, ZP(w->rhs, i));
    lu_solve(P, w);
    for (j = 0; j < n; j++) ccopy(K, ZP(w->rhs, j), nodec(P, w, j, k));
    correction_pass(P, w, k);
  }
  return 1;
}

void otr_track(const otr_program *P, int K, int npaths, int p0, int p1, const double *start,
               int KS, double *endp, double *tend, int *status, int *nsteps, int *nrej,
               double *dxlast, int *work, const otr_params *prm) {
  const int n = P->n, D = P->deg + 1;
  work_t w;
  int path;
  if (!work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = p0; path < p1; path++) {
    double t = 0.0, h = prm->maxstep, dxl = 0.0;
    int st = ST_RUNNING, steps = 0, rej = 0, j, k, npred = 0, nnewt = 0, nfeval = 0;
    load_point(P, &w, start, KS, npaths, path);
    set_one(P, &w);
    while (st == ST_RUNNING && t < 1.0) {
      const int kd = D - 1 - prm->fabry_offset;
      double rad = -1.0;
      int accepted = 0;
      set_weights(P, &w, t);
      ++npred;
      if (!predictor_series(P, &w)) { st = ST_SINGULAR; break; }
      if (kd >= 1)
        for (j = 0; j < n; j++) {
          const double den = absz(K, nodec(P, &w, j, kd));
          if (den > 0.0) {
            const double r = absz(K, nodec(P, &w, j, kd - 1)) / den;
            if (rad < 0.0 || r < rad) rad = r;
          }
        }
      h = (rad > 0.0) ? fmin(prm->pbeta * rad, prm->maxstep) : prm->maxstep;
      for (j = 0; j < n; j++) ccopy(K, nodec(P, &w, j, 0), ZP(w.x0, j));
      while (!accepted) {
        double tn = t + h, hs, res;
        int ok;
        if (tn >= 1.0 - prm->minstep) tn = 1.0;
        hs = tn - t;
        for (j = 0; j < n; j++) {
          double acc[16], tmp[16];
          ccopy(K, nodec(P, &w, j, D - 1), acc);
          for (k = D - 2; k >= 1; k--) {
            cscale_d(K, acc, hs, tmp);
            omd_cadd(K, tmp, nodec(P, &w, j, k), acc);
          }
          cscale_d(K, acc, hs, tmp);
          omd_cadd(K, ZP(w.x0, j), tmp, nodec(P, &w, j, 0));
        }
        set_weights(P, &w, tn);
        ++nfeval;
        eval_point(P, &w, 0);
        res = max_abs(K, w.rhs, n);
        ok = res <= prm->alpha * (1.0 + max_abs_x(P, &w));
        if (ok) {
          const int nr = newton(P, &w, prm->corsteps, prm->tolcor, &dxl, &nnewt);
          if (nr < 0) { st = ST_SINGULAR; break; }
          ok = (nr == 1);
        }
        if (ok) { t = tn; accepted = 1; }
        else {
          ++rej;
          h *= 0.5;
          if (h < prm->minstep) { st = ST_MINSTEP; break; }
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
      nr = newton(P, &w, prm->endsteps, prm->tolend, &dxl, &nnewt);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
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
    }
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
} job_t;

static void *run_job(void *arg) {
  job_t *j = (job_t *)arg;
  otr_track(j->P, j->K, j->npaths, j->p0, j->p1, j->start, j->KS, j->endp, j->tend, j->status,
            j->nsteps, j->nrej, j->dxlast, j->work, j->prm);
  return NULL;
}

void otr_track_mt(const otr_program *P, int K, int npaths, const double *start, int KS,
                  double *endp, double *tend, int *status, int *nsteps, int *nrej,
                  double *dxlast, int *work, const otr_params *prm, int nthreads) {
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
    pthread_create(&th[i], NULL, run_job, &jobs[i]);
  }
  for (i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
}

void otr_refine(const otr_program *P, int K, int npaths, double t, const double *xin, int KS,
                double *xout, int iters, double tol, double *dxlast, double *res, int *status) {
  work_t w;
  int path;
  if (!work_alloc(P, K, &w)) { work_free(&w); return; }
  for (path = 0; path < npaths; path++) {
    double dxl = 0.0;
    int nr = 1;
    load_point(P, &w, xin, KS, npaths, path);
    cset(K, nodec(P, &w, P->n, 0), 1.0);
    set_weights(P, &w, t);
    int its = 0;
    if (iters > 0) nr = newton(P, &w, iters, tol, &dxl, &its);
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
  set_one(P, &w);
  set_weights(P, &w, t);
  eval_point(P, &w, 1);
  for (i = 0; i < n; i++)
    for (j = 0; j < K; j++) {
      val[(size_t)j * n + i] = ZP(w.rhs, i)[j];
      val[(size_t)(K + j) * n + i] = ZP(w.rhs, i)[K + j];
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
  if (!work_alloc(P, K, &w)) { work_free(&w); return 0; }
  load_point(P, &w, x, K, 1, 0);
  set_one(P, &w);
  set_weights(P, &w, t);
  ok = predictor_series(P, &w);
  for (j = 0; j < n; j++)
    for (k = 0; k < D; k++)
      for (i = 0; i < K; i++) {
        series[((size_t)i * n + j) * D + k] = nodec(P, &w, j, k)[i];
        series[((size_t)(K + i) * n + j) * D + k] = nodec(P, &w, j, k)[K + i];
      }
  work_free(&w);
  return ok;
}
