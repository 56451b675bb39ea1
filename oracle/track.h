/* CPU oracle: path tracking of compiled homotopies (test infrastructure only).
 *
 * The program and parameter structs have the layout of phc_program and
 * phc_track_params in include/phcmd.h so the same host data feed both. */
#ifndef PHC_ORACLE_TRACK_H
#define PHC_ORACLE_TRACK_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n, deg, ncomp, wdeg, G;
  int nprod, nlev, nslots, nfactor;
  const int *nodes, *lev_ptr;
  int npair;
  const int *pairs;
  int nterm;
  const int *seg_ptr, *term_pair;
  const double *term_coef;
  int njpair;
  const int *jpairs;
  int njterm;
  const int *jent_ptr, *jterms;
  const double *weights;
  const double *hsig;
  int nhterm;
  const int *hent_ptr, *hterms;
  int nhpair;
  const int *hpairs;
} otr_program;

typedef struct {
  int maxsteps, corsteps, endsteps, fabry_offset;
  double maxstep, minstep, pbeta, alpha, tolcor, tolend, divergence;
  int pade_numdeg, pade_dendeg;
  double cbeta;
} otr_params;

/* step-size controller, layout of phc_controller (pointers are host memory) */
#define OTR_NFEAT 16
#define OTR_HIDDEN 64
#define OTR_MAXCAND 16
typedef struct {
  int mode, ncand;
  double mult[OTR_MAXCAND];
  const float *weights, *mu, *sd;
  float logit_threshold;
  unsigned seed;
  int log_capacity;
  float *log_x;
  int *log_c, *log_y, *log_n;
  double pole_cap;
  int deg0;                 /* series degree before the choice (0: the program's) */
  int degree[OTR_MAXCAND];  /* series degree of candidate c's step (0: the program's) */
  double retry_below;       /* mode 1: retry only paths stopped at t < this (0: 0.99) */
} otr_controller;

/* paths p0 <= p < p1 of npaths; arrays as in phc_track_from (tstart NULL:
 * every path starts at t = 0); ctl may be NULL */
void otr_track(const otr_program *P, int K, int npaths, int p0, int p1, const double *start,
               int KS, const double *tstart, double *endp, double *tend, int *status,
               int *nsteps, int *nrej, double *dxlast, int *work, const otr_params *prm,
               const otr_controller *ctl);

/* one path with a log of its steps: trace receives 20 doubles per step of
 * the last attempt (at most cap steps; fields listed at track_path in
 * track.c); returns the number of steps logged */
int otr_track_trace(const otr_program *P, int K, int npaths, int path, const double *start,
                    int KS, double *endp, double *tend, int *status, int *nsteps, int *nrej,
                    double *dxlast, int *work, const otr_params *prm, const otr_controller *ctl,
                    double *trace, int cap);

/* pieces of the step rule, exposed for tests */
double otr_root(double x, int k);
double otr_sigma_max_sq(int n, const double *M, double *P, double *Q);
double otr_mixed_residual(const otr_program *P, int K, double t, const double *x);
double otr_hessian_sum(const otr_program *P, int K, double t, const double *x);
void otr_pade_error(int K, double *c, int L, int M, double *x0, double *err);

/* same as otr_track over all paths with nthreads POSIX threads */
void otr_track_mt(const otr_program *P, int K, int npaths, const double *start, int KS,
                  const double *tstart, double *endp, double *tend, int *status, int *nsteps,
                  int *nrej,
                  double *dxlast, int *work, const otr_params *prm, const otr_controller *ctl,
                  int nthreads);

/* log2 approximation shared with the device (exposed for tests) */
double otr_lg2(double x);

void otr_refine(const otr_program *P, int K, int npaths, double t, const double *xin, int KS,
                double *xout, int iters, double tol, double *dxlast, double *res, int *status);

/* value and Jacobian of H at (x, t) for one point: x [2K][n] planar,
 * val [2K][n], jac [2K][n*n] */
void otr_eval(const otr_program *P, int K, double t, const double *x, double *val, double *jac);

/* Taylor coefficients of the solution curve at (x, t): series [2K][n][deg+1] */
int otr_series(const otr_program *P, int K, double t, const double *x, double *series);

/* per path: alpha, rho (the exponent and the distance of the nearest
 * singularity of the dominant component of x(tend + s), ratio method on
 * its last three Taylor coefficients; +inf when none is within reach),
 * rco = 1/cond2 of the Jacobian estimate and xmax = max |x_i|, at the
 * points x [2KS][npaths][n] and parameters tend [npaths] */
void otr_endgame(const otr_program *P, int K, int npaths, const double *x, int KS,
                 const double *tend, double *alpha, double *rho, double *rco, double *xmax);

/* [L/M] Pade coefficients (1 <= M <= 8) of one component's series
 * c[0..L+M] (complex, 2K doubles each) in place: c_1..c_L the numerator's,
 * c_{L+1}..c_{L+M} the denominator's d_1..d_M; x0 (the constant term, kept
 * apart) is zeroed when the numerator is */
void otr_pade_coefficients(int K, double *c, int L, int M, double *x0);
/* the predicted component at step s: numerator (Taylor polynomial of degree
 * L when pden = 0) with constant term x0, over 1 + d_1 s + ... + d_pden s^pden */
void otr_pade_evaluate(int K, const double *c, int L, int pden, const double *x0, double s,
                       double *out);

#ifdef __cplusplus
}
#endif
#endif
