/* phcmd_compat.h - the reference's C++ host entry points of the
 * multiple-double path, with the reference's signatures, implemented on top
 * of the C-ABI of phcmd.h (libphcmd.so).  A caller of the reference's GPU
 * functions relinks against libphcmd.so and keeps its code.
 *
 * As in the reference, each call takes host arrays (one array per limb and
 * per real/imaginary part, matrices as row pointers), allocates its device
 * buffers, copies in, runs on the GPU and copies the results back.
 *
 * Timers and operation counters keep their meaning where this build can
 * give it: *walltimesec and *totlapms-style totals receive the measured
 * time of the call; the per-phase timers of the blocked QR are left as they
 * are (the factorization runs as one HIP graph, its phases are not timed
 * apart); the counters receive the multiple-double operations this build
 * executes (a complex multiply-add counts 4 mul + 4 add).  Errors of the
 * HIP runtime are reported on stderr and leave the outputs unchanged (the
 * reference has no error channel).
 */
#ifndef PHCMD_COMPAT_H
#define PHCMD_COMPAT_H

#include <iostream>
#include <vector>

/* ------------------------------------------------------ job planners ----
 * ConvolutionJob(s) and AdditionJob(s) with the reference's public API
 * (Polynomials/convolution_job.h:9-77, convolution_jobs.h, addition_job.h,
 * addition_jobs.h:12-218): the same layers of the same jobs in the same
 * order, so callers that build the schedules (and walk or print them) keep
 * their code.  GPU_*_poly_evaldiff below takes them by value, as the
 * reference does, and evaluates from a plan of its own.
 * Job types: -1 coefficient, 0 input, 1 forward, 2 backward, 3 cross. */
class ConvolutionJob {
 public:
  ConvolutionJob(int monomialix, int input1tp, int input1ix, int input2tp, int input2ix,
                 int outputtp, int outputix);
  int get_monomial_index(void) const;
  int get_first_type(void) const;
  int get_first_input(void) const;
  int get_second_type(void) const;
  int get_second_input(void) const;
  int get_output_type(void) const;
  int get_output_index(void) const;
  friend std::ostream &operator<<(std::ostream &os, const ConvolutionJob &job);

 private:
  int monidx, inp1tp, inp1ix, inp2tp, inp2ix, outptp, outidx;
};

class ConvolutionJobs {
 public:
  explicit ConvolutionJobs(int dim);
  /* the layers of convolution jobs of a polynomial with nbr monomials,
   * monomial k a product of the nvr[k] distinct variables idx[k][...] */
  void make(int nbr, int *nvr, int **idx, bool verbose);
  int get_dimension(void) const;
  int get_count(void) const;
  int get_layer_count(int k) const;
  int get_depth(void) const;
  ConvolutionJob get_job(int k, int i) const;
  ~ConvolutionJobs(void);

 private:
  int dimension, jobcount, laydepth;
  std::vector<int> freqlaycnt;
  std::vector<std::vector<ConvolutionJob> > jobs;
  void add_job(int layer, const ConvolutionJob &job, bool verbose);
  void make_monomial(int nvr, int *idx, int monidx, bool verbose);
};

class AdditionJob {
 public:
  /* atp, itp: update and increment type (1 forward, 2 backward, 3 cross);
   * monidx2 = -1 when the increment is the constant or a coefficient */
  AdditionJob(int atp, int itp, int monidx1, int monidx2, int ix1, int ix2);
  int get_addition_type(void) const;
  int get_increment_type(void) const;
  int get_update_monomial(void) const;
  int get_update_index(void) const;
  int get_increment_monomial(void) const;
  int get_increment_index(void) const;
  friend std::ostream &operator<<(std::ostream &os, const AdditionJob &job);

 private:
  int adtype, intype, updmon, incmon, updidx, incidx;
};

class AdditionJobs {
 public:
  AdditionJobs(int dim, int nbr);
  void make(int nbr, int *nvr, int **idx, bool verbose);
  int get_number_of_variables(void) const;
  int get_number_of_monomials(void) const;
  int get_count(void) const;
  int get_layer_count(int k) const;
  int get_depth(void) const;
  int get_differential_count(int k) const;
  int get_differential_index(int k, int i) const;
  int position(int n, int *idx, int k);
  AdditionJob get_job(int k, int i) const;
  ~AdditionJobs(void);

 private:
  int nbrvar, nbrmon, jobcount, laydepth;
  std::vector<int> freqlaycnt, difcnt;
  std::vector<std::vector<int> > difidx;
  std::vector<std::vector<AdditionJob> > jobs;
  void add_job(int level, const AdditionJob &job, bool verbose);
  void tree(int kind, int level, int stride, int nbr, int *nvr, int **idx, int var,
            bool verbose);
};

/* ---------------------------------------------- complex job planners ----
 * The complex (vectorized) planners with the reference's public API
 * (Polynomials/complexconv_job.h, complexconv_jobs.h, complexinc_job.h,
 * complexinc_jobs.h, complexadd_job.h, complexadd_jobs.h): every real
 * convolution splits into the four real convolutions re*re, im*im, re*im,
 * im*re (output types 1-3 forward/backward/cross "re^a", 4-6 "re^b",
 * 7-9 "im^a", 10-12 "im^b"; input types -1/-2 coefficient re/im, 0/4 input,
 * 1-3/5-7 forward, backward, cross re/im), an increment job per real and
 * imaginary part combines them, and every real addition job becomes one on
 * the real and one on the imaginary parts, in the reference's layers and
 * order.  GPU_cmplx*vectorized_poly_evaldiff below takes them by value. */
class ComplexConvolutionJob {
 public:
  ComplexConvolutionJob(int monomialix, int input1tp, int input1ix, int input2tp, int input2ix,
                        int outputtp, int outputix);
  int get_monomial_index(void) const;
  int get_first_type(void) const;
  int get_first_input(void) const;
  int get_second_type(void) const;
  int get_second_input(void) const;
  int get_output_type(void) const;
  int get_output_index(void) const;
  friend std::ostream &operator<<(std::ostream &os, const ComplexConvolutionJob &job);

 private:
  int monidx, inp1tp, inp1ix, inp2tp, inp2ix, outptp, outidx;
};

class ComplexConvolutionJobs {
 public:
  explicit ComplexConvolutionJobs(int dim);
  void make(int nbr, int *nvr, int **idx, bool verbose);
  int get_dimension(void) const;
  int get_count(void) const;
  int get_layer_count(int k) const;
  int get_depth(void) const;
  ComplexConvolutionJob get_job(int k, int i) const;
  ~ComplexConvolutionJobs(void);

 private:
  int dimension, jobcount, laydepth;
  std::vector<int> freqlaycnt;
  std::vector<std::vector<ComplexConvolutionJob> > jobs;
};

class ComplexIncrementJob {
 public:
  /* kind 1 forward, 2 backward, 3 cross; isreal: the real (true) or the
   * imaginary parts are combined */
  ComplexIncrementJob(int monomix, int kind, int ix, bool isreal);
  int get_monomial_index(void) const;
  int get_increment_kind(void) const;
  int get_increment_index(void) const;
  bool get_isreal(void) const;
  friend std::ostream &operator<<(std::ostream &os, const ComplexIncrementJob &job);

 private:
  int monidx, inkind, incidx;
  bool rornot;
};

class ComplexIncrementJobs {
 public:
  ComplexIncrementJobs(ComplexConvolutionJobs cnvjobs, bool verbose);
  int get_dimension(void) const;
  int get_count(void) const;
  int get_layer_count(int k) const;
  int get_depth(void) const;
  ComplexIncrementJob get_job(int k, int i) const;
  ~ComplexIncrementJobs(void);

 private:
  int dimension, jobcount, laydepth;
  std::vector<int> freqlaycnt;
  std::vector<std::vector<ComplexIncrementJob> > jobs;
  ComplexIncrementJob make_job(ComplexConvolutionJob cnvjob) const;
};

class ComplexAdditionJob {
 public:
  /* atp 1-3 forward/backward/cross real part, 4-6 imaginary part; itp as
   * the output types of the convolution jobs (0 for a coefficient);
   * monidx2 -1 or -2 for the constant (or a coefficient), real or
   * imaginary part; ix2 -1 or -2 for the constant, real or imaginary */
  ComplexAdditionJob(int atp, int itp, int monidx1, int monidx2, int ix1, int ix2);
  int get_addition_type(void) const;
  int get_increment_type(void) const;
  int get_update_monomial(void) const;
  int get_update_index(void) const;
  int get_increment_monomial(void) const;
  int get_increment_index(void) const;
  friend std::ostream &operator<<(std::ostream &os, const ComplexAdditionJob &job);

 private:
  int adtype, intype, updmon, incmon, updidx, incidx;
};

class ComplexAdditionJobs {
 public:
  ComplexAdditionJobs(int dim, int nbr);
  void make(int nbr, int *nvr, int **idx, bool verbose);
  int get_number_of_variables(void) const;
  int get_number_of_monomials(void) const;
  int get_count(void) const;
  int get_layer_count(int k) const;
  int get_depth(void) const;
  int get_differential_count(int k) const;
  int get_differential_index(int k, int i) const;
  int position(int n, int *idx, int k);
  ComplexAdditionJob get_job(int k, int i) const;
  ~ComplexAdditionJobs(void);

 private:
  int nbrvar, nbrmon, jobcount, laydepth;
  std::vector<int> freqlaycnt, difcnt;
  std::vector<std::vector<int> > difidx;
  std::vector<std::vector<ComplexAdditionJob> > jobs;
};

/* ---------------------------------------- polynomial evaluation --------
 * Polynomials/dbl2_polynomials_kernels.h:812,870 and the dbl, dbl4, dbl8
 * twins: the value (output[dim]) and the dim derivatives (output[k]) of
 * cst + sum_k cff[k] * prod_i input[idx[k][i]] at power series truncated at
 * degree deg, with the result of the reference's CPU_*_poly_evaldiff bit for
 * bit.  BS (deg + 1 in the reference) and the job schedules are accepted and
 * not needed; *cnvlapms receives the kernel time, *addlapms 0 (the sums run
 * in the same kernel), *elapsedms their sum, *walltimesec the wall time of
 * the call. */
void GPU_dbl_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx, double *cst,
                           double **cff, double **input, double **output,
                           ConvolutionJobs cnvjobs, AdditionJobs addjobs, double *cnvlapms,
                           double *addlapms, double *elapsedms, double *walltimesec, int vrblvl);
void GPU_dbl2_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx,
                            double *csthi, double *cstlo, double **cffhi, double **cfflo,
                            double **inputhi, double **inputlo, double **outputhi,
                            double **outputlo, ConvolutionJobs cnvjobs, AdditionJobs addjobs,
                            double *cnvlapms, double *addlapms, double *elapsedms,
                            double *walltimesec, int vrblvl);
void GPU_dbl4_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx,
                            double *csthihi, double *cstlohi, double *csthilo, double *cstlolo,
                            double **cffhihi, double **cfflohi, double **cffhilo,
                            double **cfflolo, double **inputhihi, double **inputlohi,
                            double **inputhilo, double **inputlolo, double **outputhihi,
                            double **outputlohi, double **outputhilo, double **outputlolo,
                            ConvolutionJobs cnvjobs, AdditionJobs addjobs, double *cnvlapms,
                            double *addlapms, double *elapsedms, double *walltimesec,
                            int vrblvl);
void GPU_dbl8_poly_evaldiff(
    int BS, int dim, int nbr, int deg, int *nvr, int **idx, double *csthihihi,
    double *cstlohihi, double *csthilohi, double *cstlolohi, double *csthihilo,
    double *cstlohilo, double *csthilolo, double *cstlololo, double **cffhihihi,
    double **cfflohihi, double **cffhilohi, double **cfflolohi, double **cffhihilo,
    double **cfflohilo, double **cffhilolo, double **cfflololo, double **inputhihihi,
    double **inputlohihi, double **inputhilohi, double **inputlolohi, double **inputhihilo,
    double **inputlohilo, double **inputhilolo, double **inputlololo, double **outputhihihi,
    double **outputlohihi, double **outputhilohi, double **outputlolohi,
    double **outputhihilo, double **outputlohilo, double **outputhilolo,
    double **outputlololo, ConvolutionJobs cnvjobs, AdditionJobs addjobs, double *cnvlapms,
    double *addlapms, double *elapsedms, double *walltimesec, int vrblvl);
/* Polynomials/dbl_polynomials_kernels.h:675, dbl2_polynomials_kernels.h:870,
 * dbl4_polynomials_kernels.h:1254: complex coefficients and series. */
void GPU_cmplx_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx,
                             double *cstre, double *cstim, double **cffre, double **cffim,
                             double **inputre, double **inputim, double **outputre,
                             double **outputim, ConvolutionJobs cnvjobs, AdditionJobs addjobs,
                             double *cnvlapms, double *addlapms, double *elapsedms,
                             double *walltimesec, int vrblvl);
void GPU_cmplx2_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx,
                              double *cstrehi, double *cstrelo, double *cstimhi,
                              double *cstimlo, double **cffrehi, double **cffrelo,
                              double **cffimhi, double **cffimlo, double **inputrehi,
                              double **inputrelo, double **inputimhi, double **inputimlo,
                              double **outputrehi, double **outputrelo, double **outputimhi,
                              double **outputimlo, ConvolutionJobs cnvjobs,
                              AdditionJobs addjobs, double *cnvlapms, double *addlapms,
                              double *elapsedms, double *walltimesec, int vrblvl);
void GPU_cmplx4_poly_evaldiff(
    int BS, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehihi,
    double *cstrelohi, double *cstrehilo, double *cstrelolo, double *cstimhihi,
    double *cstimlohi, double *cstimhilo, double *cstimlolo, double **cffrehihi,
    double **cffrelohi, double **cffrehilo, double **cffrelolo, double **cffimhihi,
    double **cffimlohi, double **cffimhilo, double **cffimlolo, double **inputrehihi,
    double **inputrelohi, double **inputrehilo, double **inputrelolo, double **inputimhihi,
    double **inputimlohi, double **inputimhilo, double **inputimlolo, double **outputrehihi,
    double **outputrelohi, double **outputrehilo, double **outputrelolo,
    double **outputimhihi, double **outputimlohi, double **outputimhilo,
    double **outputimlolo, ConvolutionJobs cnvjobs, AdditionJobs addjobs, double *cnvlapms,
    double *addlapms, double *elapsedms, double *walltimesec, int vrblvl);

/* Polynomials/dbl_polynomials_kernels.h:735, dbl2_polynomials_kernels.h:960,
 * dbl4_polynomials_kernels.h:1397: the complex evaluation with the
 * vectorized schedules.  The reference's kernels split every complex
 * product into real convolutions and agree with CPU_cmplx*_poly_evaldiff to
 * rounding; this library returns CPU_cmplx*_poly_evaldiff's result bit for
 * bit (the schedules are accepted and not needed); timings as above. */
void GPU_cmplxvectorized_poly_evaldiff(int BS, int dim, int nbr, int deg, int *nvr, int **idx,
                                       double *cstre, double *cstim, double **cffre,
                                       double **cffim, double **inputre, double **inputim,
                                       double **outputre, double **outputim,
                                       ComplexConvolutionJobs cnvjobs,
                                       ComplexIncrementJobs incjobs,
                                       ComplexAdditionJobs addjobs, double *cnvlapms,
                                       double *addlapms, double *elapsedms,
                                       double *walltimesec, int vrblvl);
void GPU_cmplx2vectorized_poly_evaldiff(
    int BS, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehi, double *cstrelo,
    double *cstimhi, double *cstimlo, double **cffrehi, double **cffrelo, double **cffimhi,
    double **cffimlo, double **inputrehi, double **inputrelo, double **inputimhi,
    double **inputimlo, double **outputrehi, double **outputrelo, double **outputimhi,
    double **outputimlo, ComplexConvolutionJobs cnvjobs, ComplexIncrementJobs incjobs,
    ComplexAdditionJobs addjobs, double *cnvlapms, double *addlapms, double *elapsedms,
    double *walltimesec, int vrblvl);
void GPU_cmplx4vectorized_poly_evaldiff(
    int BS, int dim, int nbr, int deg, int *nvr, int **idx, double *cstrehihi,
    double *cstrelohi, double *cstrehilo, double *cstrelolo, double *cstimhihi,
    double *cstimlohi, double *cstimhilo, double *cstimlolo, double **cffrehihi,
    double **cffrelohi, double **cffrehilo, double **cffrelolo, double **cffimhihi,
    double **cffimlohi, double **cffimhilo, double **cffimlolo, double **inputrehihi,
    double **inputrelohi, double **inputrehilo, double **inputrelolo, double **inputimhihi,
    double **inputimlohi, double **inputimhilo, double **inputimlolo, double **outputrehihi,
    double **outputrelohi, double **outputrehilo, double **outputrelolo,
    double **outputimhihi, double **outputimlohi, double **outputimhilo,
    double **outputimlolo, ComplexConvolutionJobs cnvjobs, ComplexIncrementJobs incjobs,
    ComplexAdditionJobs addjobs, double *cnvlapms, double *addlapms, double *elapsedms,
    double *walltimesec, int vrblvl);

/* ------------------------------------------------ series products ------
 * Convolutions/dbl2_convolutions_kernels.h:121,145 and the dbl, dbl4, dbl8
 * twins: z = x*y truncated at degree deg; freq repeats the launch (the
 * reference times freq launches), BS and padded/mode select the reference's
 * kernel variant and are accepted for compatibility (one kernel here). */
void GPU_dbl_product(double *x_h, double *y_h, double *z_h, int deg, int freq, int BS, int mode);
void GPU_cmplx_product(double *xre, double *xim, double *yre, double *yim, double *zre,
                       double *zim, int deg, int freq, int BS, int mode);
void GPU_dbl2_product(double *xhi_h, double *xlo_h, double *yhi_h, double *ylo_h, double *zhi_h,
                      double *zlo_h, int deg, int freq, int BS, int padded);
void GPU_cmplx2_product(double *xrehi_h, double *xrelo_h, double *ximhi_h, double *ximlo_h,
                        double *yrehi_h, double *yrelo_h, double *yimhi_h, double *yimlo_h,
                        double *zrehi_h, double *zrelo_h, double *zimhi_h, double *zimlo_h,
                        int deg, int freq, int BS, int mode);
void GPU_dbl4_product(double *xhihi_h, double *xlohi_h, double *xhilo_h, double *xlolo_h,
                      double *yhihi_h, double *ylohi_h, double *yhilo_h, double *ylolo_h,
                      double *zhihi_h, double *zlohi_h, double *zhilo_h, double *zlolo_h,
                      int deg, int freq, int BS, int padded);
void GPU_cmplx4_product(double *xrehihi_h, double *xrelohi_h, double *xrehilo_h,
                        double *xrelolo_h, double *ximhihi_h, double *ximlohi_h,
                        double *ximhilo_h, double *ximlolo_h, double *yrehihi_h,
                        double *yrelohi_h, double *yrehilo_h, double *yrelolo_h,
                        double *yimhihi_h, double *yimlohi_h, double *yimhilo_h,
                        double *yimlolo_h, double *zrehihi_h, double *zrelohi_h,
                        double *zrehilo_h, double *zrelolo_h, double *zimhihi_h,
                        double *zimlohi_h, double *zimhilo_h, double *zimlolo_h, int deg,
                        int freq, int BS, int mode);

/* Convolutions/dbl8_convolutions_kernels.h:238,291: octo-double products. */
void GPU_dbl8_product(double *xhihihi_h, double *xlohihi_h, double *xhilohi_h,
                      double *xlolohi_h, double *xhihilo_h, double *xlohilo_h,
                      double *xhilolo_h, double *xlololo_h, double *yhihihi_h,
                      double *ylohihi_h, double *yhilohi_h, double *ylolohi_h,
                      double *yhihilo_h, double *ylohilo_h, double *yhilolo_h,
                      double *ylololo_h, double *zhihihi_h, double *zlohihi_h,
                      double *zhilohi_h, double *zlolohi_h, double *zhihilo_h,
                      double *zlohilo_h, double *zhilolo_h, double *zlololo_h, int deg, int freq,
                      int BS, int padded);
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
    int freq, int BS, int mode);

/* ------------------------------------------------ blocked QR ------------
 * Matrices/dbl2_baqr_kernels.h:2540 and dbl4_baqr_kernels.h:3739:
 * A (nrows x ncols, nrows >= ncols = szt*nbt) = Q R with Q (nrows x nrows)
 * and R (nrows x ncols). */
/* Matrices/dbl2_baqr_kernels.h:2479, dbl4_baqr_kernels.h:3668,
 * dbl8_baqr_kernels.h:4787: the real blocked QR (phc_real_blocked_houseqr),
 * R with a positive diagonal as the reference's real Householder vector
 * gives; the timing arguments other than walltimesec are left unchanged. */
void GPU_dbl2_blocked_houseqr(int nrows, int ncols, int szt, int nbt, double **Ahi, double **Alo,
                              double **Qhi, double **Qlo, double **Rhi, double **Rlo,
                              double *houselapms, double *RTvlapms, double *tileRlapms,
                              double *vb2Wlapms, double *WYTlapms, double *QWYTlapms,
                              double *Qaddlapms, double *YWTlapms, double *YWTClapms,
                              double *Raddlapms, double *walltimesec, long long int *addcnt,
                              long long int *mulcnt, long long int *divcnt,
                              long long int *sqrtcnt, bool verbose = true);
void GPU_dbl4_blocked_houseqr(int nrows, int ncols, int szt, int nbt, double **Ahihi,
                              double **Alohi, double **Ahilo, double **Alolo, double **Qhihi,
                              double **Qlohi, double **Qhilo, double **Qlolo, double **Rhihi,
                              double **Rlohi, double **Rhilo, double **Rlolo,
                              double *houselapms, double *RTvlapms, double *tileRlapms,
                              double *vb2Wlapms, double *WYTlapms, double *QWYTlapms,
                              double *Qaddlapms, double *YWTlapms, double *YWTClapms,
                              double *Raddlapms, double *walltimesec, long long int *addcnt,
                              long long int *mulcnt, long long int *divcnt,
                              long long int *sqrtcnt, bool verbose = true);
void GPU_dbl8_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Ahihihi, double **Alohihi, double **Ahilohi,
    double **Alolohi, double **Ahihilo, double **Alohilo, double **Ahilolo, double **Alololo,
    double **Qhihihi, double **Qlohihi, double **Qhilohi, double **Qlolohi, double **Qhihilo,
    double **Qlohilo, double **Qhilolo, double **Qlololo, double **Rhihihi, double **Rlohihi,
    double **Rhilohi, double **Rlolohi, double **Rhihilo, double **Rlohilo, double **Rhilolo,
    double **Rlololo, double *houselapms, double *RTvlapms, double *tileRlapms,
    double *vb2Wlapms, double *WYTlapms, double *QWYTlapms, double *Qaddlapms, double *YWTlapms,
    double *YWTClapms, double *Raddlapms, double *walltimesec, long long int *addcnt,
    long long int *mulcnt, long long int *divcnt, long long int *sqrtcnt, bool verbose = true);

void GPU_cmplx2_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Arehi, double **Arelo, double **Aimhi,
    double **Aimlo, double **Qrehi, double **Qrelo, double **Qimhi, double **Qimlo,
    double **Rrehi, double **Rrelo, double **Rimhi, double **Rimlo, double *houselapms,
    double *RHvlapms, double *tileRlapms, double *vb2Wlapms, double *WYHlapms,
    double *QWYHlapms, double *Qaddlapms, double *YWHlapms, double *YWHClapms,
    double *Raddlapms, double *walltimesec, long long int *addcnt, long long int *mulcnt,
    long long int *divcnt, long long int *sqrtcnt, bool verbose = true);
void GPU_cmplx4_blocked_houseqr(
    int nrows, int ncols, int szt, int nbt, double **Arehihi, double **Arelohi,
    double **Arehilo, double **Arelolo, double **Aimhihi, double **Aimlohi, double **Aimhilo,
    double **Aimlolo, double **Qrehihi, double **Qrelohi, double **Qrehilo, double **Qrelolo,
    double **Qimhihi, double **Qimlohi, double **Qimhilo, double **Qimlolo, double **Rrehihi,
    double **Rrelohi, double **Rrehilo, double **Rrelolo, double **Rimhihi, double **Rimlohi,
    double **Rimhilo, double **Rimlolo, double *houselapms, double *RHvlapms,
    double *tileRlapms, double *vb2Wlapms, double *WYHlapms, double *QWYHlapms,
    double *Qaddlapms, double *YWHlapms, double *YWHClapms, double *Raddlapms,
    double *walltimesec, long long int *addcnt, long long int *mulcnt, long long int *divcnt,
    long long int *sqrtcnt, bool verbose = true);

/* ------------------------------------------------ tiled back substitution
 * Matrices/dbl2_tabs_kernels.h:392 and dbl4_tabs_kernels.h:604: U x = b
 * for upper triangular U (dim = szt*nbt); U and b are left unchanged, x is
 * returned. */
/* Matrices/dbl8_baqr_kernels.h:4882: the octo-double complex blocked QR. */
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
    double **Rimhilolo, double **Rimlololo, double *houselapms, double *RHvlapms,
    double *tileRlapms, double *vb2Wlapms, double *WYHlapms, double *QWYHlapms,
    double *Qaddlapms, double *YWHlapms, double *YWHClapms, double *Raddlapms,
    double *walltimesec, long long int *addcnt, long long int *mulcnt, long long int *divcnt,
    long long int *sqrtcnt, bool verbose = true);

/* Matrices/dbl8_tabs_kernels.h:1011: the octo-double complex tiled solver. */
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
    double *ximlololo, double *invlapms, double *mullapms, double *sublapms, double *totlapms,
    double *walltimesec, long long int *addcnt, long long int *mulcnt, long long int *divcnt);

void GPU_cmplx2_upper_tiled_solver(int dim, int szt, int nbt, double **Urehi, double **Urelo,
                                   double **Uimhi, double **Uimlo, double *brehi, double *brelo,
                                   double *bimhi, double *bimlo, double *xrehi, double *xrelo,
                                   double *ximhi, double *ximlo, double *invlapms,
                                   double *mullapms, double *sublapms, double *totlapms,
                                   double *walltimesec, long long int *addcnt,
                                   long long int *mulcnt, long long int *divcnt);
void GPU_cmplx4_upper_tiled_solver(
    int dim, int szt, int nbt, double **Urehihi, double **Urelohi, double **Urehilo,
    double **Urelolo, double **Uimhihi, double **Uimlohi, double **Uimhilo, double **Uimlolo,
    double *brehihi, double *brelohi, double *brehilo, double *brelolo, double *bimhihi,
    double *bimlohi, double *bimhilo, double *bimlolo, double *xrehihi, double *xrelohi,
    double *xrehilo, double *xrelolo, double *ximhihi, double *ximlohi, double *ximhilo,
    double *ximlolo, double *invlapms, double *mullapms, double *sublapms, double *totlapms,
    double *walltimesec, long long int *addcnt, double *addover, long long int *mulcnt,
    double *mulover, long long int *divcnt, double *divover);

/* Matrices/dbl2_tabs_kernels.h:358 and dbl4_tabs_kernels.h:555: the real
 * solvers, run as the complex ones with zero imaginary parts (U and b left
 * unchanged, x returned); x agrees with CPU_dbl2_upper_tiled_solver to a
 * few units in the last limb. */
void GPU_dbl2_upper_tiled_solver(int dim, int szt, int nbt, double **Uhi, double **Ulo,
                                 double *bhi, double *blo, double *xhi, double *xlo,
                                 double *invlapms, double *mullapms, double *sublapms,
                                 double *totlapms, double *walltimesec, long long int *addcnt,
                                 long long int *mulcnt, long long int *divcnt);
void GPU_dbl4_upper_tiled_solver(int dim, int szt, int nbt, double **Uhihi, double **Ulohi,
                                 double **Uhilo, double **Ulolo, double *bhihi, double *blohi,
                                 double *bhilo, double *blolo, double *xhihi, double *xlohi,
                                 double *xhilo, double *xlolo, double *invlapms,
                                 double *mullapms, double *sublapms, double *totlapms,
                                 double *walltimesec, long long int *addcnt, double *addover,
                                 long long int *mulcnt, double *mulover, long long int *divcnt,
                                 double *divover);

/* ------------------------------------------------ Q^H b ------------------
 * Newton/dbl2_bals_kernels.h:221 and dbl4_bals_kernels.h:305: b := Q^H b
 * for the ncols x ncols Q of the QR. */
void GPU_cmplx2_bals_qhb(int ncols, int szt, int nbt, double **Qrehi, double **Qrelo,
                         double **Qimhi, double **Qimlo, double *brehi, double *brelo,
                         double *bimhi, double *bimlo, double *totqtblapsedms, int vrblvl);
void GPU_cmplx4_bals_qhb(int ncols, int szt, int nbt, double **Qrehihi, double **Qrelohi,
                         double **Qrehilo, double **Qrelolo, double **Qimhihi, double **Qimlohi,
                         double **Qimhilo, double **Qimlolo, double *brehihi, double *brelohi,
                         double *brehilo, double *brelolo, double *bimhihi, double *bimlohi,
                         double *bimhilo, double *bimlolo, double *totqtblapsedms, int vrblvl);

/* ------------------------------------------------ block Toeplitz solve --
 * Newton/dbl_bals_kernels.h:199, dbl2_bals_kernels.h:255,
 * dbl4_bals_kernels.h:353: sum_{j<=k} A_j x_{k-j} = b_k, k < degp1, for the
 * degp1 matrices mat[j] (dim x dim), with the reference's skip rules; Q, R,
 * rhs, sol and the flags are updated as the reference's host solve
 * (Newton/dbl2_bals_host.cpp:479) updates them, bit for bit. */
void GPU_dbl_bals_solve(int dim, int degp1, int szt, int nbt, int tailidx, double ***mat,
                        double **Q, double **R, double **rhs, double **sol, bool *zeroQ,
                        bool *noqr, int *upidx, int *bsidx, int *newtail, double *totqrlapsedms,
                        double *totqtblapsedms, double *totbslapsedms, double *totupdlapsedms,
                        int vrblvl);
void GPU_dbl2_bals_solve(int dim, int degp1, int szt, int nbt, int tailidx, double ***mathi,
                         double ***matlo, double **Qhi, double **Qlo, double **Rhi, double **Rlo,
                         double **rhshi, double **rhslo, double **solhi, double **sollo,
                         bool *zeroQ, bool *noqr, int *upidx, int *bsidx, int *newtail,
                         double *totqrlapsedms, double *totqtblapsedms, double *totbslapsedms,
                         double *totupdlapsedms, int vrblvl);
void GPU_dbl4_bals_solve(int dim, int degp1, int szt, int nbt, int tailidx, double ***mathihi,
                         double ***matlohi, double ***mathilo, double ***matlolo,
                         double **Qhihi, double **Qlohi, double **Qhilo, double **Qlolo,
                         double **Rhihi, double **Rlohi, double **Rhilo, double **Rlolo,
                         double **rhshihi, double **rhslohi, double **rhshilo, double **rhslolo,
                         double **solhihi, double **sollohi, double **solhilo, double **sollolo,
                         bool *zeroQ, bool *noqr, int *upidx, int *bsidx, int *newtail,
                         double *totqrlapsedms, double *totqtblapsedms, double *totbslapsedms,
                         double *totupdlapsedms, int vrblvl);

/* ------------------------------------------------ linear residue -------
 * Newton/dbl_tail_kernels.h:125, dbl2_tail_kernels.h:164,
 * dbl4_tail_kernels.h:244: resvec_i = rhs_i - sum_{j<=i-tailidx} mat_j
 * sol_{i-j} for i = tailidx..degp1-1 (rows below tailidx unchanged), resmax
 * the limbs (absolute values) of the entry of largest |leading limb|; the
 * same bits as CPU_dbl2_linear_residue (Newton/dbl2_bals_host.cpp:548).
 * *lapms accumulates the kernel time; *add and *mul are set to the
 * reference's count (dim*dim per matrix-vector product). */
void GPU_dbl_linear_residue(int dim, int degp1, int szt, int nbt, int tailidx, double ***mat,
                            double **rhs, double **sol, double **resvec, double *resmax,
                            double *totreslapsedms, long long int *add, long long int *mul,
                            int vrblvl);
void GPU_dbl2_linear_residue(int dim, int degp1, int szt, int nbt, int tailidx, double ***mathi,
                             double ***matlo, double **rhshi, double **rhslo, double **solhi,
                             double **sollo, double **resvechi, double **resveclo,
                             double *resmaxhi, double *resmaxlo, double *lapms,
                             long long int *add, long long int *mul, int vrblvl);
void GPU_dbl4_linear_residue(int dim, int degp1, int szt, int nbt, int tailidx,
                             double ***mathihi, double ***matlohi, double ***mathilo,
                             double ***matlolo, double **rhshihi, double **rhslohi,
                             double **rhshilo, double **rhslolo, double **solhihi,
                             double **sollohi, double **solhilo, double **sollolo,
                             double **resvechihi, double **resveclohi, double **resvechilo,
                             double **resveclolo, double *resmaxhihi, double *resmaxlohi,
                             double *resmaxhilo, double *resmaxlolo, double *lapms,
                             long long int *add, long long int *mul, int vrblvl);

#endif
