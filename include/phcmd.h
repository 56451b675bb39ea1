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

#ifdef __cplusplus
}
#endif
#endif
