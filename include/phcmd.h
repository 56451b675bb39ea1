/* phcmd.h - C-ABI of the MI355X-native multiple-double power-series path.
 *
 * Plain pointers, int sizes, opaque plan handles, int status codes
 * (0 = success, -1 = HIP runtime error, -2 = invalid argument; the message is
 * available from phc_last_error()).  "stream" is a hipStream_t passed as
 * void* (NULL = the default stream).  Device pointers are marked "dev".
 *
 * Number layout (shared by every entry point): a number with K limbs
 * (K = 1 double, 2 double-double, 4 quad-double, 8 octo-double) is stored
 * limb-planar; an array of m numbers is K planes of m doubles, plane i holding
 * limb i, limb 0 the most significant (hi, hihi, hihihi).  Complex arrays are
 * the K real planes followed by the K imaginary planes (P = 2K planes).  This
 * is the reference's split-limb convention (one array per limb and per re/im
 * part, Convolutions/dbl2_convolutions_kernels.h) with the arrays stacked.
 */
#ifndef PHCMD_H
#define PHCMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- runtime */

const char *phc_last_error(void);
int phc_version(void);
int phc_device_count(int *n);
int phc_malloc(void **ptr, size_t bytes);
int phc_free(void *ptr);
int phc_memcpy_htod(void *dst_dev, const void *src, size_t bytes);
int phc_memcpy_dtoh(void *dst, const void *src_dev, size_t bytes);
int phc_memset(void *dst_dev, int value, size_t bytes);
int phc_synchronize(void);
int phc_stream_synchronize(void *stream);
int phc_event_create(void **ev);
int phc_event_destroy(void *ev);
int phc_event_record(void *ev, void *stream);
int phc_event_elapsed_ms(void *start, void *stop, float *ms);
/* device-to-device copy by a streaming kernel (bytes and pointers 16-byte
 * aligned); used to calibrate HBM byte counters */
int phc_device_copy(void *dst_dev, const void *src_dev, size_t bytes, void *stream);

/* ------------------------------------------------------------- arithmetic */

/* c = a op b elementwise over n numbers of K limbs (K = 2, 4, 8);
 * op 0 add, 1 sub, 2 mul, 3 div, 4 sqrt(a) (b unused).
 * Replaces the device arithmetic ddg_* / qdg_* / odg_*
 * (Norms/double_double_gpufun.cu:9-305, quad_double_gpufun.cu,
 * octo_double_gpufun.cu), exposed batched. */
int phc_md_batch(int op, int K, int n, const double *a_dev, const double *b_dev,
                 double *c_dev, void *stream);

/* ----------------------------------------------------------------- series */

/* z[b] = x[b]*y[b] truncated at degree deg, b < nbatch; arrays [P][nbatch][deg+1].
 * Replaces GPU_dbl2_product / GPU_cmplx2_product and the dbl4/dbl8 twins
 * (Convolutions/dbl2_convolutions_kernels.cu:372), batched. */
int phc_series_product(int K, int cplx, int deg, int nbatch, const double *x_dev,
                       const double *y_dev, double *z_dev, void *stream);

/* Plan for the value and gradient of one polynomial
 *   p(x) = cst + sum_{i<nbr} cff_i * prod_{j<nvr[i]} x[idx_i[j]]
 * at power series of degree deg (coefficients are series too).
 * idx is the concatenation of the nvr[i] variable indices of monomial i;
 * cst [P][deg+1] and cff [P][nbr][deg+1] are host arrays copied to the device.
 * Replaces the setup half of GPU_dbl2_poly_evaldiff
 * (Polynomials/dbl2_polynomials_kernels.cu:1577-1653, which rebuilds the
 * ConvolutionJobs/AdditionJobs schedule and data array on every call). */
typedef struct phc_poly phc_poly;
int phc_poly_create(phc_poly **plan, int K, int cplx, int dim, int nbr, int deg,
                    const int *nvr, const int *idx, const double *cst, const double *cff);
int phc_poly_destroy(phc_poly *plan);

/* Evaluate the plan's polynomial at nbatch input series at once:
 * input [P][nbatch][dim][deg+1], output [P][nbatch][dim+1][deg+1] with
 * output[b][dim] the value and output[b][i] the derivative to x_i.
 * Same result, bit for bit, as CPU_dbl2_poly_evaldiff
 * (Polynomials/dbl2_polynomials_host.cpp:259) per instance. */
int phc_poly_evaldiff(phc_poly *plan, int nbatch, const double *input_dev, double *output_dev,
                      void *stream);

/* --------------------------------------------------------------- tracking */

/* Compiled homotopy H_i(x,t) = sum_r w_r(t) sum_m c[r,i,m] x^m (built by
 * phc_program_compile below, or by compile_homotopy in
 * phcpack_amd/homotopy.py: the same arrays).  Series coefficients live in
 * "slots" (complex units): the variables x_j at j*(deg+1), the constant 1 at
 * n*(deg+1), product nodes where nodes[] says.
 *   nodes [8*nprod], per product node in DAG-level order:
 *     aoff, amask (left factor: a variable or a product node, full series,
 *     amask = -1), boff (right factor: variable j at j*(deg+1)), soff, smask
 *     (own storage: coefficient a at soff + (a & smask); smask = -1 for full
 *     series, 2^m - 1 for a ring of the last 2^m orders), adn, sdn (index of
 *     the left factor's / own order-k correction in the correction buffer,
 *     -1 for none), left (node number of the left factor: j < n the
 *     variable x_j, n the constant, n+1+i product i).
 *   lev_ptr [nlev+1]: nodes of level l are lev_ptr[l-1]..lev_ptr[l]-1.
 *   pairs [4*npair]: component r, storage offset, mask, node number of a
 *     node whose weighted value u = sum_l W_{r,l} node_{k-l} enters the
 *     residual.
 *   seg_ptr [n*G+1]: terms of segment i*G+g of equation i (G a power of 2,
 *     n*G <= 64); term_pair [nterm], term_coef [2][nterm] (real plane,
 *     imaginary plane; coefficients are complex doubles).  Terms of one
 *     monomial (its components) are adjacent in one segment: the mixed
 *     residual takes the modulus of their sum.
 *   jpairs [2*njpair]: component r, storage offset of node m/x_j.
 *   jent_ptr [n*n+1], jterms [njterm]: Jacobian entry i*n+j sums
 *     coef[term] * e * W_{r,0} * node_0 with jterm = term | jpair << 16 | e << 28.
 *   weights [2][8][3*ncomp]: octo-double coefficients of 1, t, t^2 of w_r.
 *   hsig [n]: largest singular value of the Hessian of equation i, kept
 *     from t = 0 for equations of degree <= 2 as Cached_Singular_Values
 *     (Deformations/Curves/standard_predictor_convolutions.adb:575-646)
 *     keeps it; -1 for equations whose Hessian is evaluated every step.
 *   hent_ptr [n*n*(n+1)/2+1], hterms [3*nhterm], hpairs [2*nhpair]: Hessian
 *     entry (i; a <= b) at i*n*(n+1)/2 + a*n - a*(a-1)/2 + b - a sums
 *     mult * coef[term] * W_{r,0} * node_0 over hterms (term, hpair, mult),
 *     hpair = (component r, storage offset of the node x^(m - e_a - e_b)).
 * Replaces the homotopy set-up of gpumanypaths_dd
 * (Ada/PHCtoC/gpumanypaths_dd.cpp:18-57, which reads PHCpack's containers)
 * and the per-call job planning of GPU_dbl2_poly_evaldiff. */
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
} phc_program;

/* Step control of the tracker; defaults of the reference
 * (Deformations/Curves/homotopy_continuation_parameters.adb:11-22):
 * maxstep 0.1, minstep 1e-6, pbeta 0.5, cbeta 0.005, alpha 1e-3,
 * tolres 1e-8, corsteps 4, maxsteps 1000, [5/1] Pade predictor.
 * Every step follows LU_Prediction and Predictor_Corrector_Loop
 * (Deformations/Curves/standard_predictor_convolutions.adb:1142-1185,
 * predictor_corrector_loops.adb:16-83):
 *   pole step  pbeta*r, r = |c_{d-1-o}/c_{d-o}| minimized over the
 *              components whose c_{d-o} is not zero (Is_Zero: both parts
 *              vanish next to 1), maxstep when all are zero
 *              (Math_Lib/Circuits/convergence_radius_estimates.adb:28-38,138-163,354-390);
 *   curvature  (cbeta*eta/nrm)^(1/deg), eta = 2 sigma_min(J) /
 *              sqrt(sum_i sigma_max(Hessian_i)^2) (Hesse_Pade, Distance,
 *              :558-572,648-900), nrm the 2-norm of the Pade error
 *              estimates (Homotopy_Pade_Approximants.Solution_Error);
 *   step       Minimum(pole, curvature, maxstep) (three_way_minima.adb),
 *              set to 1 - t when t + step >= 1 - minstep (Bounded_Update);
 *   feedback   halve the step while the mixed residual of the predicted
 *              point, mean_i |f_i|/(|f|_i + 1) with |f|_i equation i with
 *              coefficient moduli at |x| (standard_mixed_residuals.adb:472-484),
 *              is >= alpha and the step >= minstep (Predictor_Feedback, :958-993);
 *   corrector  Newton (LU_Newton_Steps, standard_inlined_newton_circuits.adb:
 *              388-422): converged when the mixed residual and max|dx| are
 *              both <= tolcor, then one extra step; stops early when both
 *              grow; on failure the step is halved and the predictor
 *              re-evaluated, down to minstep.
 * The build's departures: the point is tracked on with t advanced by the
 * step at which it was corrected (the reference advances t before the
 * feedback and corrects at the halved step); a path whose corrector gives up
 * below minstep ends with status 2 (the reference continues from the
 * unconverged point); divergence is a status of its own; the series is
 * computed order by order from the corrected point (the reference runs
 * series Newton); the singular values are computed by power iteration on
 * the Gram matrix (see phc_track). */
typedef struct {
  int maxsteps;     /* steps per path before giving up */
  int corsteps;     /* Newton iterations per corrector (plus one extra) */
  int endsteps;     /* Newton iterations at t = 1 (plus one extra) */
  int fabry_offset; /* pole radius |c_{d-1-o}/c_{d-o}| (the reference: 2) */
  double maxstep, minstep, pbeta, alpha;
  double tolcor;    /* corrector: mixed residual and max|dx| <= tolcor */
  double tolend;    /* the same at t = 1 */
  double divergence;/* path declared diverged when max|x| exceeds this */
  /* predictor (Predictor_Feedback evaluates the Pade approximants of
   * Pade_Vector, standard_predictor_convolutions.adb:312,958-994;
   * Denominator_System/Pade, Deformations/Curves/standard_rational_approximations.adb):
   * pade_dendeg 0: Taylor polynomial of degree pade_numdeg (0: deg - 2);
   * pade_dendeg M = 1 or 2: [pade_numdeg/M] Pade approximant per component;
   * pade_numdeg + pade_dendeg + 2 <= deg, as the reference requires for
   * its error estimate (standard_predictor_convolutions.ads:34). */
  int pade_numdeg, pade_dendeg;
  double cbeta;     /* curvature step factor (the reference's beta2) */
} phc_track_params;

/* Checks every index of a program against its sizes (no device work);
 * phc_tracker_create calls it first. */
int phc_program_validate(const phc_program *program);

/* Host compiler of the program (host only, no device work): the homotopy
 * H(x,t) = sum_r w_r(t) P_r(x) with ncomp components, w_r(t) = sum_l
 * weights[r][l][0] t^l + i weights[r][l][1] t^l (l = 0, 1, 2), and P_r
 * given term by term: nterms[r] terms after those of component r-1, term k
 * in equation eqs[k] (0..n-1) with exponents exps[k][n] (0..15) and the
 * complex coefficient in the limb planes coef[2*CL][sum nterms] (CL = 1:
 * double, 2: double-double; real limbs, then imaginary limbs).  The
 * program is built for series of degree deg over `lanes` lanes per path
 * (64), the same arrays as compile_homotopy of phcpack_amd/homotopy.py.
 * A double-double coefficient gives two terms of its monomial (high part,
 * low part).  The homotopy of the reference's batched tracker,
 * gamma (1 - t) q + t p (Ada/PHCtoC/gpumanypaths_dd.cpp:88, alpha = gamma),
 * is ncomp = 2 with weights {0, 1, 0} for p and {gamma, -gamma, 0} for q.
 * phc_compiled_program fills a phc_program whose arrays belong to the
 * compiled object (valid until phc_compiled_destroy). */
typedef struct phc_compiled phc_compiled;
int phc_program_compile(phc_compiled **out, int n, int deg, int lanes, int ncomp,
                        const double *weights, const int *nterms, const int *eqs, const int *exps,
                        int CL, const double *coef);
int phc_compiled_program(const phc_compiled *compiled, phc_program *program);
int phc_compiled_destroy(phc_compiled *compiled);

/* Learned step-size control (no counterpart in the reference, whose step
 * rule - the pole step of LU_Prediction,
 * Deformations/Curves/standard_predictor_convolutions.adb:1142-1185 - stays
 * the judge and the fallback: the predicted-residual feedback and the
 * Newton corrector still accept or reject every step, and a rejected step
 * is halved as before).  The controller multiplies the step h0 =
 * Minimum(pole step, curvature step, maxstep) of the reference's rule by
 * one of ncand candidates mult[c], before Bounded_Update.  Each step the
 * tracker forms PHC_CTL_NFEAT features (log2 of the Fabry radius, of the
 * previous step, t, log2 of the largest series coefficient of orders 1..8,
 * log2 of the last four Newton corrections, rejections of the previous
 * step), standardizes them with mu and sd, and evaluates an FP32 MLP
 * nfeat -> 64 (ReLU) -> 64 (ReLU) -> ncand logits (phc_mlp layout).
 * mode 0: plain rule (the controller is ignored);
 * mode 1: learned: the largest c whose logit >= logit_threshold (c = 0 if none);
 *         a path that does not finish (status other than success) at a
 *         t below retry_below (0: 0.99) is tracked again from its start with
 *         the plain rule in the same launch: its work counters add up both
 *         attempts, its other outputs are the second attempt's; a path that
 *         stops closer to t = 1 (a singular or diverging endpoint, which the
 *         plain rule does not reach either) is not tracked again;
 * mode 2: explore: c drawn from a hash of (seed, path, step), and every step
 *         logged: features log_x[path][i][nfeat], candidate log_c[path][i],
 *         log_y[path][i] = the Newton iterations of the step if its first
 *         attempt was accepted (>= 1), else 0, log_n[path]
 *         records (at most log_capacity per path).
 * In modes 1 and 2 the multiplied step is capped at pole_cap * r (r the
 * Fabry radius) when pole_cap > 0, so that the step stays inside the disk
 * where the series converges; candidate 0 (multiplier 1) is the plain rule.
 * Predictor order: in modes 1 and 2 the tracker first computes the Taylor
 * series to degree deg0 (every degree 0: the program's), forms the
 * features from it (orders above deg0 read as zero, the Fabry radius from
 * c_{deg0-1-o} / c_{deg0-o}), and then continues the series to the chosen
 * candidate's degree[c] (deg0 <= degree[c] <= the program's degree), from
 * which the step takes its pole radius, its [L-(deg-degree[c])/M] Pade
 * predictor and its curvature step (root degree[c]) as the reference's rule
 * does for a series of that degree.  A candidate is thus a (multiplier,
 * series degree) pair, and a lower degree saves the orders it leaves out. */
#define PHC_CTL_NFEAT 16
#define PHC_CTL_HIDDEN 64
#define PHC_CTL_MAXCAND 16
typedef struct {
  int mode, ncand;
  double mult[PHC_CTL_MAXCAND];
  const float *weights, *mu, *sd; /* device */
  float logit_threshold;
  unsigned seed;
  int log_capacity;
  float *log_x;                   /* device */
  int *log_c, *log_y, *log_n;     /* device */
  double pole_cap;                /* 0: no cap */
  int deg0;                       /* series degree before the choice (0: the program's) */
  int degree[PHC_CTL_MAXCAND];    /* series degree of candidate c's step (0: the program's) */
  double retry_below;             /* mode 1: retry only paths stopped at t < this (0: 0.99) */
} phc_controller;

typedef struct phc_tracker phc_tracker;
int phc_tracker_create(phc_tracker **tracker, const phc_program *program);
int phc_tracker_destroy(phc_tracker *tracker);
/* LDS bytes of the staged program and of one path at precision K (1, 2, 4, 8) */
int phc_tracker_lds_bytes(const phc_tracker *tracker, int K, size_t *program_bytes,
                          size_t *path_bytes);

/* Track npaths paths from t = 0 to t = 1 in precision K (1 = double,
 * 2 = double-double) with a persistent kernel: one workgroup per compute
 * unit stages the program in LDS, each of its (at most `waves`, as LDS
 * allows) wavefronts tracks one path at a time, paths taken from an atomic
 * queue.
 * start [2*KS][npaths][n] (KS limbs), endp [2*K][npaths][n]; per path:
 * tend, status (1 success, 2 step below minstep, 3 maxsteps, 4 singular
 * Jacobian, 5 diverged), nsteps, nrej (rejected steps), dxlast, and when
 * work is not NULL the work counters work[c*npaths + path] for c = 0
 * predictor series computed, 1 Newton iterations, 2 predicted-residual
 * evaluations, 3 corrector calls (whose first Newton iteration reuses the
 * predicted-residual evaluation and only forms the Jacobian), 4, 5, 6 the
 * steps set by the pole step, the curvature step and maxstep (the
 * reference's nbpole, nbhess, nbmaxm), 7 Newton runs whose first iteration
 * evaluates the point (corrector retries and the Newton at t = 1), so work
 * holds 8*npaths ints.
 * Singular values: sigma_max(M)^2 is the Rayleigh quotient |M v|^2/|v|^2
 * of v = the column of largest diagonal entry of (M^H M)^64 (the Gram
 * matrix scaled by its trace and squared six times), in FP64 from the
 * leading limbs; sigma_min(J) = 1/sigma_max(J^{-1}) with the inverse the
 * tracker already forms.  The reference's SVD (zsvdc) gives the same
 * values to rounding when the two largest singular values are apart, and
 * otherwise a value within a factor (1 + 1e-3 n) of them (tested).
 * Calls on one tracker are ordered: the path queue counter and the block
 * of step parameters the kernel reads belong to the tracker, so a call on
 * another stream than the tracker's previous call first waits (on the
 * device) for that call's kernels to end.
 * Replaces GPU_Path_mult / path_mult (Path/Path_GPU/path_gpu_mult.cu:84-383)
 * behind the gpumanypaths_dd entry (Ada/PHCtoC/gpumanypaths_dd.cpp:18). */
int phc_track(phc_tracker *tracker, int K, int npaths, const double *start_dev, int KS,
              double *endp_dev, double *tend_dev, int *status_dev, int *nsteps_dev,
              int *nrej_dev, double *dxlast_dev, int *work_dev, const phc_track_params *params,
              const phc_controller *controller, int waves, void *stream);

/* phc_track with a start parameter per path, tstart [npaths] (NULL: 0): path
 * p is tracked from the point start[p] at t = tstart[p] to t = 1.  A path
 * that stops (status 2-5) ends at its last accepted point, with tend its
 * parameter, so a run in another precision can continue it from there (the
 * mixed method's escalation of FP64 stops to double-double). */
int phc_track_from(phc_tracker *tracker, int K, int npaths, const double *start_dev, int KS,
                   const double *tstart_dev, double *endp_dev, double *tend_dev, int *status_dev,
                   int *nsteps_dev, int *nrej_dev, double *dxlast_dev, int *work_dev,
                   const phc_track_params *params, const phc_controller *controller, int waves,
                   void *stream);

/* ------------------------------------------------------- controller MLP */

/* number of floats of an MLP F -> H -> H -> C (W1[F][H], b1, W2[H][H], b2,
 * W3[H][C], b3) */
int phc_mlp_size(int F, int H, int C, int *nparams);
/* logits [B][C] of inputs X [B][F] (device arrays) */
int phc_mlp_forward(int F, int H, int C, const float *w_dev, int B, const float *X_dev,
                    float *logits_dev, void *stream);
/* gradient of the mean binary cross-entropy of logit cand[i] against y[i]
 * over B samples (device arrays; loss_dev may be NULL) */
int phc_mlp_grad(int F, int H, int C, const float *w_dev, int B, const float *X_dev,
                 const int *cand_dev, const float *y_dev, float *grad_dev, float *loss_dev,
                 void *stream);
/* minibatch Adam on N samples for `epochs` passes; w updated in place;
 * losses (host, may be NULL) receives the mean loss of each epoch */
int phc_mlp_train(int F, int H, int C, float *w_dev, int N, const float *X_dev,
                  const int *cand_dev, const float *y_dev, int epochs, int batch, float lr,
                  unsigned seed, float *losses, void *stream);

/* Newton refinement at fixed t in precision K (1, 2, 4; octo-double K = 8
 * is refused with -2 in this build) of points with KS limbs; xout [2*K][npaths][n]; res = max_i |H_i(xout, t)| evaluated in
 * precision K; status 1 converged (max|dx| <= tol*(1+max|x|)), 0 not, -1
 * singular.  iters = 0 only evaluates the residual.  only_dev (NULL: every
 * path) holds the tracker's status per path: paths whose status is not 1
 * (success) are copied through unrefined with res = dx = inf and status =
 * their tracker status, so diverged or stalled paths cost no Newton work.
 * Plays the role of the Newton drivers at the endpoints
 * (Newton/dbl2_newton_method.cpp:429, row Newton step) in dd/qd/od. */
int phc_refine(phc_tracker *tracker, int K, int npaths, double t, const double *xin_dev, int KS,
               double *xout_dev, int iters, double tol, double *dxlast_dev, double *res_dev,
               int *status_dev, const int *only_dev, int waves, void *stream);

/* Where the paths go as t -> 1: the question the endgames of
 * Deformations/End_Games answer (the winding number of a singular endpoint,
 * Ada's *_vlprs_algorithm extrapolation, or divergence to infinity),
 * estimated here from the Taylor series of x(t + s) at each path's last
 * point x [2*KS][npaths][n] and parameter tend [npaths], in precision K (1 or
 * 2): ratio method on the last three Taylor coefficients of each component
 * (q_k = c_{k+1}/c_k at k = deg - 1): the nearest singularity lies at
 * s* = 1/((k+1) q_k - k q_{k-1}) and the component behaves like
 * (s* - s)^alpha with alpha = k - (k+1) s* q_k, exact for c (s* - s)^alpha.
 * Per path: alpha and rho = Re s* of the component whose last coefficient
 * weighs most at the distance 1 - tend (+inf when every component has a
 * zero coefficient); rco = 1/(sigma_max(J) sigma_max(J^{-1})), the inverse
 * 2-norm condition number estimate of the Jacobian at the point (0 when J
 * is singular); xmax = max_i |x_i|.  alpha < 0 with rho ~ 1 - tend: a pole
 * at t = 1 (the path diverges); 0 < alpha < 1: a branch point, winding
 * number 1/alpha.  No counterpart in the reference's GPU code; the
 * classification built on it is phcpack_amd.tracker.classify_endpoints. */
int phc_endgame(phc_tracker *tracker, int K, int npaths, const double *x_dev, int KS,
                const double *tend_dev, double *alpha_dev, double *rho_dev, double *rco_dev,
                double *xmax_dev, int waves, void *stream);

/* ---------------- linearized power-series Newton (real series) ----------
 * Layouts are limb-planar with the batch index after the limb:
 *   matrices [K][nbatch][n][n], matrix series [K][nbatch][degp1][n][n],
 *   vector series [K][nbatch][degp1][n], series x [K][nsys][dim][deg+1].
 * K = 1, 2, 4 (double, double-double, quad-double); results are
 * bit-identical to the reference's host functions named below. */

/* batched Householder QR with explicit Q of n x n matrices (n <= 64):
 * CPU_dbl2_factors_houseqr (Matrices/dbl2_factorizations.cpp:723) and its
 * dbl/dbl4 twins, in place of GPU_dbl2_blocked_houseqr for small n. */
int phc_houseqr_batch(int K, int n, int nbatch, const double *A_dev, double *Q_dev,
                      double *R_dev, void *stream);

/* batched lower block-triangular Toeplitz solve sum_{j<=k} A_j x_{k-j} = b_k
 * with the reference's skip rules (threshold 1e-15 / 1e-28 / 1e-56):
 * CPU_dbl2_qrbs_solve (Newton/dbl2_bals_host.cpp:479), replaces
 * GPU_dbl2_bals_solve (Newton/dbl2_bals_kernels.cu:554).  rhs is updated
 * by the tail as in the reference; Q, R are read when the head is skipped
 * and written when it runs; flags [nbatch][5] = zeroQ, noqr, upidx, bsidx,
 * newtail, in and out (newtail in = the old tail index). */
int phc_qrbs_solve_batch(int K, int n, int degp1, int tailidx, int nbatch, const double *mat_dev,
                         double *rhs_dev, double *sol_dev, double *Q_dev, double *R_dev,
                         int *flags_dev, void *stream);

/* r_i = b_i - sum_{j<=i-tailidx} A_j x_{i-j}, i = tailidx..degp1-1, and the
 * limbs of the entry of largest |leading limb| in resmax [K][nbatch]:
 * CPU_dbl2_linear_residue (Newton/dbl2_bals_host.cpp:548), replaces
 * GPU_dbl2_linear_residue (Newton/dbl2_tail_kernels.cu:383). */
int phc_linear_residue_batch(int K, int n, int degp1, int tailidx, int nbatch,
                             const double *mat_dev, const double *rhs_dev, const double *sol_dev,
                             double *res_dev, double *resmax_dev, void *stream);

/* The complex twins (K = 2, 4; arrays complex limb-planar, 2K planes per
 * array: K real, then K imaginary), bit-identical to CPU_cmplx2_factors_houseqr
 * (Matrices/dbl2_factorizations.cpp:766), CPU_cmplx2_qrbs_solve and
 * CPU_cmplx2_linear_residue (Newton/dbl2_bals_host.cpp) and their cmplx4 /
 * cmplx8 twins: the skip tests on max(|re| + |im|) of the leading limbs
 * (CPU_cmplx_onenorm), resmax [K][nbatch] = |re| + |im| of each limb of
 * the entry of largest |re| + |im|.  K = 8 is refused (-2): the first
 * device run of the complex octo-double QR stalled, cause not established;
 * the CPU oracle covers the octo-double complex functions. */
int phc_cmplx_houseqr_batch(int K, int n, int nbatch, const double *A_dev, double *Q_dev,
                            double *R_dev, void *stream);
int phc_cmplx_qrbs_solve_batch(int K, int n, int degp1, int tailidx, int nbatch,
                               const double *mat_dev, double *rhs_dev, double *sol_dev,
                               double *Q_dev, double *R_dev, int *flags_dev, void *stream);
int phc_cmplx_linear_residue_batch(int K, int n, int degp1, int tailidx, int nbatch,
                                   const double *mat_dev, const double *rhs_dev,
                                   const double *sol_dev, double *res_dev, double *resmax_dev,
                                   void *stream);

/* the supports of a polynomial system for row Newton: polynomial i has
 * monomials monoptr[i]..monoptr[i+1]-1, monomial m the nvr[m] distinct
 * variables idx[idxptr[m]..idxptr[m+1]-1] (the nbr/nvr/idx arguments of
 * CPU_dbl2_poly_evaldiff, Polynomials/dbl2_polynomials_host.h, flattened);
 * host arrays, copied to the device. */
typedef struct phc_rowsys phc_rowsys;
int phc_rowsys_create(phc_rowsys **sys, int dim, const int *monoptr, const int *nvr,
                      const int *idxptr, const int *idx);
/* the same with the exponent of every variable occurrence, exp[] parallel
 * to idx[] (1..64; NULL: all ones): the exp/nbrfac/expfac arguments of
 * CPU_dbl2_evaluate_monomials (Newton/dbl2_systems_host.h:7-46, its
 * common-factor trick, Newton/dbl2_systems_host.cpp:12-111) for the
 * monomials of a row system: each coefficient is multiplied by
 * prod_j x_j^(e_j - 1), one factor at a time in the order of idx, the
 * distinct variables are differentiated as CPU_dbl2_poly_evaldiff does,
 * and each derivative contribution is multiplied by its exponent before it
 * is added to the polynomial's. */
int phc_rowsys_create_exp(phc_rowsys **sys, int dim, const int *monoptr, const int *nvr,
                          const int *idxptr, const int *idx, const int *exp);
int phc_rowsys_destroy(phc_rowsys *sys);
/* LDS bytes one system occupies at degree deg with K doubles per number
 * (pass 2K for the complex Newton) */
int phc_rowsys_lds_bytes(const phc_rowsys *sys, int K, int deg, size_t *bytes);

/* staggered row Newton on nsys systems with the same supports, one
 * wavefront per system: working degrees 0, 1, 2, 4, 7, ... capped at deg,
 * at most nbsteps steps, stopping when the tail index reaches deg.  Each
 * step evaluates and differentiates the system at the series x, solves the
 * linearized block-Toeplitz system and updates x, as dbl2_row_newton_qrstep
 * (Newton/dbl2_newton_method.cpp:429) inside the loop of
 * test_dbl2_row_newton (:1447-1494).  cst [K][nc][dim][deg+1] and
 * cff [K][nf][nmono][deg+1] with nc (nf) = 1 when cst_shared (cff_shared),
 * else nsys; x [K][nsys][dim][deg+1] in/out.  Optional outputs (NULL to
 * skip): steps [nsys]; wdeg, tail, up, bs [nsys][nbsteps]; history
 * [nbsteps][K][nsys][dim][deg+1]; Q, R [K][nsys][dim][dim]. */
int phc_row_newton(phc_rowsys *sys, int K, int deg, int nbsteps, int nsys, const double *cst_dev,
                   int cst_shared, const double *cff_dev, int cff_shared, double *x_dev,
                   int *steps_dev, int *wdeg_dev, int *tail_dev, int *up_dev, int *bs_dev,
                   double *history_dev, double *Q_dev, double *R_dev, void *stream);

/* A system in the column representation of dbl2_column_newton_qrstep
 * (Newton/dbl2_newton_method.cpp:259-427): nbrcol columns of dim monomials,
 * monomial m = c*dim + j (equation j of column c) with nvr[m] distinct
 * variables (0: none), their indices idx[] concatenated in the order of m.
 * nbrcol = 1 is a monomial system x^A_i = mb_i: nvr[i] >= 1, unit
 * coefficients times the common factors of the exponents exps[] (parallel
 * to idx, 1..64; NULL: all ones), evaluated by CPU_dbl2_evaluate_monomials
 * (Newton/dbl2_systems_host.cpp:12-111) and linearized by
 * dbl2_linearize_evaldiff_output (:443-523).  nbrcol > 1: coefficient
 * series per monomial, no exponents, the constant -1 in the last equation,
 * evaluated by CPU_dbl2_evaluate_columns (:232-329) with dbl2_define_rhs
 * (:626-659).  Destroy with phc_rowsys_destroy. */
int phc_colsys_create(phc_rowsys **sys, int dim, int nbrcol, const int *nvr, const int *idx,
                      const int *exps);

/* staggered column Newton (K = 1, 2, 4), bit-identical to the reference's
 * host branch: the loop of test_dbl2_column_newton
 * (Newton/dbl2_newton_method.cpp:1036-1287) over dbl2_column_newton_qrstep,
 * one wavefront per system.  cff [K][nf][nbrcol*dim][deg+1] (nf = 1 when
 * cff_shared, else nsys; ignored for monomial systems, whose coefficients
 * are the unit series), mb [K][nm][dim][deg+1] the right-hand side (nm = 1
 * when mb_shared, else nsys): each step solves J dx = -(f(x) - mb).  Other
 * arguments and outputs as phc_row_newton. */
int phc_column_newton(phc_rowsys *sys, int K, int deg, int nbsteps, int nsys,
                      const double *cff_dev, int cff_shared, const double *mb_dev, int mb_shared,
                      double *x_dev, int *steps_dev, int *wdeg_dev, int *tail_dev, int *up_dev,
                      int *bs_dev, double *history_dev, double *Q_dev, double *R_dev,
                      void *stream);

/* the complex staggered row Newton (K = 2, 4; a system of distinct
 * variables): cst, cff, x, history, Q, R complex limb-planar (2K planes),
 * each step as cmplx2_row_newton_qrstep (Newton/cmplx2_newton_method.cpp:565)
 * in the loop of test_cmplx2_row_newton (:1759), bit-identical to it. */
int phc_cmplx_row_newton(phc_rowsys *sys, int K, int deg, int nbsteps, int nsys,
                         const double *cst_dev, int cst_shared, const double *cff_dev,
                         int cff_shared, double *x_dev, int *steps_dev, int *wdeg_dev,
                         int *tail_dev, int *up_dev, int *bs_dev, double *history_dev,
                         double *Q_dev, double *R_dev, void *stream);

/* ------------- complex blocked Householder QR, tiled back substitution ----
 * K = 2, 4 or 8.  Matrices are complex limb-planar [2K][rows][cols],
 * row-major.  Results agree with the reference's host code to a few units in
 * the last limb (long sums are split across lanes). */

/* workspace for A (nrows x ncols, nrows >= ncols, ncols = szt*nbt) */
typedef struct phc_baqr phc_baqr;
int phc_baqr_create(phc_baqr **plan, int K, int nrows, int ncols, int szt);
int phc_baqr_destroy(phc_baqr *plan);

/* Launch configuration of the plan's matrix products (the trailing updates,
 * W = Y S, the Gram matrix, Q^H b), chosen by the autotuner
 * (phcpack_amd/autotune.py, tune_qr); no counterpart in the reference, whose
 * kernels fix their launch shapes.  gemm_block: outputs per lane, 0 the
 * default (2 x 2 for K = 2, 1 x 1 for K = 4, 8), 1: 1 x 1, 2: 2 x 1,
 * 3: 1 x 2, 4: 2 x 2; split_target: the k range of a product is split into
 * slices, doubling while tiles x slices < split_target (0: 1024), slices
 * of >= 64 terms, within the plan's workspace.  The split changes the order
 * of the long sums (results to rounding); the block does not.  get returns
 * the values in use (the defaults resolved). */
int phc_baqr_set_tuning(phc_baqr *plan, int gemm_block, int split_target);
int phc_baqr_get_tuning(const phc_baqr *plan, int *gemm_block, int *split_target);

/* A = Q R with Q (nrows x nrows) explicit and R (nrows x ncols), in tiles of
 * szt columns: CPU_cmplx4_blocked_houseqr (Matrices/dbl4_baqr_host.cpp:1206),
 * replaces GPU_cmplx4_blocked_houseqr (Matrices/dbl4_baqr_kernels.cu:6197)
 * and its dd twin GPU_cmplx2_blocked_houseqr (dbl2_baqr_kernels.cu). */
int phc_cmplx_blocked_houseqr(phc_baqr *plan, const double *A_dev, double *Q_dev, double *R_dev,
                              void *stream);

/* y = Q^H b for the plan's nrows x nrows Q and b, y of length nrows (y must
 * not alias b): GPU_cmplx4_bals_qhb (Newton/dbl4_bals_kernels.cu:602), the
 * right-hand side step of the least squares solve before the tiled back
 * substitution. */
int phc_cmplx_qhb(phc_baqr *plan, const double *Q_dev, const double *b_dev, double *y_dev,
                  void *stream);

/* The real blocked QR with the same plan: A [K][nrows][ncols], Q
 * [K][nrows][nrows], R [K][nrows][ncols] real limb-planar (K planes).
 * CPU_dbl4_blocked_houseqr (Matrices/dbl4_baqr_host.cpp:1033), replaces
 * GPU_dbl4_blocked_houseqr (Matrices/dbl4_baqr_kernels.cu:5746) and its dd
 * and od twins: the real Householder vector of CPU_dbl4_factors_house
 * (v0 = x0 - |x| or -sigma/(x0 + |x|), so R has a positive diagonal); the
 * panel kernels run on the plan's complex workspace with zero imaginary
 * parts, the matrix products on the real planes only.
 * phc_real_qtb: y = Q^T b (b, y real, length nrows). */
int phc_real_blocked_houseqr(phc_baqr *plan, const double *A_dev, double *Q_dev, double *R_dev,
                             void *stream);
int phc_real_qtb(phc_baqr *plan, const double *Q_dev, const double *b_dev, double *y_dev,
                 void *stream);

/* U x = b for upper triangular U (dim = szt*nbt): the diagonal tiles are
 * inverted, then the tiles are applied from the last one up.  As in the
 * reference, U's diagonal tiles are replaced by their inverses and b is
 * updated.  CPU_cmplx4_upper_tiled_solver (Matrices/dbl4_tabs_host.cpp:345),
 * replaces GPU_cmplx4_upper_tiled_solver (Matrices/dbl4_tabs_kernels.cu:1802).
 * Synchronizes the stream. */
int phc_cmplx_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U_dev, double *b_dev,
                                 double *x_dev, void *stream);

/* The same solve without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference.  The tile inverses
 * live in a static device scratch of the library of 1048576 doubles (8 MiB)
 * per GPU: 2*K*dim*szt must not exceed it (dim 1024 in tiles of 128 in
 * quad-double fits), else -2.  Queued on the stream, no synchronization.
 * Every call on one GPU shares that scratch: calls that overlap on one GPU
 * (two streams, two threads) race and must be ordered by the caller. */
int phc_cmplx_upper_tiled_solver_noalloc(int K, int dim, int szt, int nbt, double *U_dev,
                                         double *b_dev, double *x_dev, void *stream);

/* C(i,j) = [C(i,j) +] sum_k op(A(i,k)) op(B(k,j)) for complex
 * multiple-double operands (K = 2, 4, 8; K = 8 serves octo-double checks of
 * quad-double results) given by base pointer, plane stride and the two
 * element strides (element (r,c) of A at A + a0*r + a1*c in every plane);
 * op conjugates when the flag is set.  The product kernel of the QR (the
 * trailing updates of dbl4_baqr_host.cpp:357,743), exposed for checks such as
 * Q^H Q and Q^H b. */
int phc_cmplx_gemm(int K, int conjA, int conjB, int M, int N, int Kd, const double *A_dev,
                   size_t a_plane, long a0, long a1, const double *B_dev, size_t b_plane, long b0,
                   long b1, double *C_dev, size_t c_plane, long c0, long c1, int accumulate,
                   void *stream);

#ifdef __cplusplus
}
#endif
#endif
