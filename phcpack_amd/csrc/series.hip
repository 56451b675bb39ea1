// Batched multiple-double arithmetic, truncated power series products and
// polynomial evaluation/differentiation at power series on gfx950.
//
// Reference counterparts:
//   elementwise dd/qd/od ops   Norms/*_gpufun.cu (ddg_*, qdg_*, odg_*)
//   series product             GPU_dbl2_product, Convolutions/dbl2_convolutions_kernels.cu:372
//   polynomial eval/diff       GPU_dbl2_poly_evaldiff, Polynomials/dbl2_polynomials_kernels.cu:1577
// The reference launches one block of deg+1 threads per product and one
// polynomial per call; here one launch covers a batch of independent
// instances (one per power-series Newton system or path), the plan holding
// the monomial structure and coefficients stays resident on the device, and
// each lane owns one series coefficient of one instance.  The order of the
// multiple-double operations per coefficient is that of the reference host
// code (CPU_dbl2_product, CPU_dbl2_poly_speel), so results are bit-identical
// to it.
#include <hip/hip_runtime.h>
#include <vector>
#include <cstring>
#include "common.hpp"
#include "mdtypes.hpp"
#include "../../include/phcmd.h"

using namespace md;

namespace {

// number traits: real (K limbs) or complex (2K planes)
template <int K, bool CPLX> struct Num;
template <int K> struct Num<K, false> {
  typedef typename RealT<K>::type V;
  static constexpr int P = K;
  __device__ static V load(const double *p, size_t s, size_t j) { return R<K>::load(p, s, j); }
  __device__ static void store(double *p, size_t s, size_t j, const V &v) { R<K>::store(p, s, j, v); }
  __device__ static V mul(const V &a, const V &b) { return md::mul(a, b); }
  __device__ static void mac(V &c, const V &a, const V &b) { c = md::add(c, md::mul(a, b)); }
  __device__ static void inc(V &c, const V &a) { c = md::add(c, a); }
  __device__ static V zero() { return R<K>::zero(); }
};
template <int K> struct Num<K, true> {
  typedef cplx<K> V;
  static constexpr int P = 2 * K;
  __device__ static V load(const double *p, size_t s, size_t j) { return C<K>::load(p, s, j); }
  __device__ static void store(double *p, size_t s, size_t j, const V &v) { C<K>::store(p, s, j, v); }
  __device__ static V mul(const V &a, const V &b) { return C<K>::mul(a, b); }
  __device__ static void mac(V &c, const V &a, const V &b) { C<K>::mac(c, a, b); }
  __device__ static void inc(V &c, const V &a) { c = C<K>::add(c, a); }
  __device__ static V zero() { return C<K>::zero(); }
};

// ----------------------------------------------------------- elementwise --

template <int K>
__global__ void __launch_bounds__(256) md_batch_kernel(int op, int n, const double *a, const double *b, double *c) {
  typedef typename RealT<K>::type T;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    T x = R<K>::load(a, n, j), z;
    T y = (op < 4) ? R<K>::load(b, n, j) : R<K>::zero();
    if (op == 0) z = md::add(x, y);
    else if (op == 1) z = md::sub(x, y);
    else if (op == 2) z = md::mul(x, y);
    else if (op == 3) z = md::div(x, y);
    else z = R<K>::sqrt(x);
    R<K>::store(c, n, j, z);
  }
}

// -------------------------------------------------------- series product --

// coefficient k of x*y: x_0*y_k, then += x_i*y_{k-i} for i = 1..k
template <class N>
__device__ typename N::V coef_product(const double *x, size_t sx, size_t xo, const double *y,
                                      size_t sy, size_t yo, int k) {
  typename N::V acc = N::mul(N::load(x, sx, xo), N::load(y, sy, yo + k));
  for (int i = 1; i <= k; ++i) N::mac(acc, N::load(x, sx, xo + i), N::load(y, sy, yo + k - i));
  return acc;
}

// z[b] = x[b]*y[b] for b < nbatch; arrays [P][nbatch][D]
template <int K, bool CPLX>
__global__ void __launch_bounds__(256) product_kernel(int D, int nbatch, const double *x, const double *y, double *z) {
  typedef Num<K, CPLX> N;
  const size_t s = (size_t)nbatch * D;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < s;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t b = t / D;
    const int k = (int)(t % D);
    N::store(z, s, t, coef_product<N>(x, s, b * D, y, s, b * D, k));
  }
}

// ------------------------------------------------ polynomial eval / diff --

// One block holds ipb instances, each owning dpad lanes (lane k = coefficient
// k).  Monomials are visited in order, as in CPU_dbl2_poly_speel
// (Polynomials/dbl2_polynomials_host.cpp:13-109); chained products are
// separated by block barriers because coefficient k of f[j] needs
// coefficients 0..k of f[j-1], which other lanes produce.
template <int K, bool CPLX>
__global__ void __launch_bounds__(256) poly_evaldiff_kernel(int dim, int nbr, int D, const int *__restrict__ nvr,
                                     const int *__restrict__ idxptr, const int *__restrict__ idx,
                                     const double *__restrict__ cst, const double *__restrict__ cff,
                                     int nbatch, const double *__restrict__ input,
                                     double *__restrict__ output, double *fw, double *bw,
                                     double *cr, int ipb, int dpad) {
  typedef Num<K, CPLX> N;
  typedef typename N::V V;
  const int sub = threadIdx.x / dpad;
  const int k = threadIdx.x % dpad;
  const int b = blockIdx.x * ipb + sub;
  const bool on = (b < nbatch) && (k < D) && (sub < ipb);
  const size_t sX = (size_t)nbatch * dim * D;        // input and work plane stride
  const size_t sO = (size_t)nbatch * (dim + 1) * D;  // output plane stride
  const size_t sC = (size_t)nbr * D;                 // coefficient plane stride
  const size_t xb = (size_t)b * dim * D;             // this instance's input / work base
  const size_t ob = (size_t)b * (dim + 1) * D;       // this instance's output base
#define XS(r) (xb + (size_t)(r) * D)
#define OS(r) (ob + (size_t)(r) * D)

  if (on) {
    N::store(output, sO, OS(dim) + k, N::load(cst, D, k));
    for (int i = 0; i < dim; ++i) N::store(output, sO, OS(i) + k, N::zero());
  }
  for (int i = 0; i < nbr; ++i) {
    const int nv = nvr[i];
    const int *ix = idx + idxptr[i];
    const size_t co = (size_t)i * D;
    if (nv == 1) {
      if (on) {
        V f = coef_product<N>(input, sX, XS(ix[0]), cff, sC, co, k);
        V o = N::load(output, sO, OS(dim) + k);
        N::inc(o, f);
        N::store(output, sO, OS(dim) + k, o);
        o = N::load(output, sO, OS(ix[0]) + k);
        N::inc(o, N::load(cff, sC, co + k));
        N::store(output, sO, OS(ix[0]) + k, o);
      }
    } else if (nv == 2) {
      if (on) {
        V f = coef_product<N>(cff, sC, co, input, sX, XS(ix[0]), k);
        N::store(fw, sX, XS(0) + k, f);
        V o = N::load(output, sO, OS(ix[1]) + k);
        N::inc(o, f);
        N::store(output, sO, OS(ix[1]) + k, o);
        V g = coef_product<N>(cff, sC, co, input, sX, XS(ix[1]), k);
        o = N::load(output, sO, OS(ix[0]) + k);
        N::inc(o, g);
        N::store(output, sO, OS(ix[0]) + k, o);
      }
      __syncthreads();
      if (on) {
        V f = coef_product<N>(fw, sX, XS(0), input, sX, XS(ix[1]), k);
        V o = N::load(output, sO, OS(dim) + k);
        N::inc(o, f);
        N::store(output, sO, OS(dim) + k, o);
      }
      __syncthreads();
    } else if (nv > 2) {
      // forward f[j] and backward b[j] chains advance together
      if (on) {
        N::store(fw, sX, XS(0) + k, coef_product<N>(cff, sC, co, input, sX, XS(ix[0]), k));
        N::store(bw, sX, XS(0) + k,
                 coef_product<N>(input, sX, XS(ix[nv - 1]), input, sX, XS(ix[nv - 2]), k));
      }
      __syncthreads();
      for (int j = 1; j < nv; ++j) {
        if (on) {
          N::store(fw, sX, XS(j) + k,
                   coef_product<N>(fw, sX, XS(j - 1), input, sX, XS(ix[j]), k));
          if (j < nv - 2)
            N::store(bw, sX, XS(j) + k,
                     coef_product<N>(bw, sX, XS(j - 1), input, sX, XS(ix[nv - 2 - j]), k));
        }
        __syncthreads();
      }
      V blast;
      if (on) blast = coef_product<N>(bw, sX, XS(nv - 3), cff, sC, co, k);
      __syncthreads();
      if (on) N::store(bw, sX, XS(nv - 3) + k, blast);
      __syncthreads();
      if (on) {
        if (nv == 3) {
          N::store(cr, sX, XS(0) + k, coef_product<N>(fw, sX, XS(0), input, sX, XS(ix[2]), k));
        } else {
          for (int j = 0; j < nv - 3; ++j)
            N::store(cr, sX, XS(j) + k,
                     coef_product<N>(fw, sX, XS(j), bw, sX, XS(nv - 4 - j), k));
          N::store(cr, sX, XS(nv - 3) + k,
                   coef_product<N>(fw, sX, XS(nv - 3), input, sX, XS(ix[nv - 1]), k));
        }
        V o = N::load(output, sO, OS(dim) + k);
        N::inc(o, N::load(fw, sX, XS(nv - 1) + k));
        N::store(output, sO, OS(dim) + k, o);
        o = N::load(output, sO, OS(ix[nv - 1]) + k);
        N::inc(o, N::load(fw, sX, XS(nv - 2) + k));
        N::store(output, sO, OS(ix[nv - 1]) + k, o);
        o = N::load(output, sO, OS(ix[0]) + k);
        N::inc(o, N::load(bw, sX, XS(nv - 3) + k));
        N::store(output, sO, OS(ix[0]) + k, o);
        for (int j = 1; j < nv - 1; ++j) {
          o = N::load(output, sO, OS(ix[j]) + k);
          N::inc(o, N::load(cr, sX, XS(j - 1) + k));
          N::store(output, sO, OS(ix[j]) + k, o);
        }
      }
      __syncthreads();
    }
  }
#undef XS
#undef OS
}

template <int K, bool CPLX>
void launch_poly(int grid, int block, hipStream_t st, int dim, int nbr, int D, const int *nvr,
                 const int *idxptr, const int *idx, const double *cst, const double *cff,
                 int nbatch, const double *in, double *out, double *fw, double *bw, double *cr,
                 int ipb, int dpad) {
  hipLaunchKernelGGL((poly_evaldiff_kernel<K, CPLX>), dim3(grid), dim3(block), 0, st, dim, nbr,
                     D, nvr, idxptr, idx, cst, cff, nbatch, in, out, fw, bw, cr, ipb, dpad);
}

bool valid_k(int K) { return K == 1 || K == 2 || K == 4 || K == 8; }

}  // namespace

// ================================================================ C-ABI ====

struct phc_poly {
  int K, cplx, dim, nbr, deg;
  int *d_nvr = nullptr, *d_idxptr = nullptr, *d_idx = nullptr;
  double *d_cst = nullptr, *d_cff = nullptr;
  std::vector<int> nvr;
  double *work = nullptr;
  size_t work_cap = 0;
};

extern "C" {

int phc_md_batch(int op, int K, int n, const double *a, const double *b, double *c,
                 void *stream) {
  PHC_REQUIRE(valid_k(K) && K > 1, "phc_md_batch: K must be 2, 4 or 8");
  PHC_REQUIRE(op >= 0 && op <= 4, "phc_md_batch: op must be 0..4");
  PHC_REQUIRE(n >= 0, "phc_md_batch: n < 0");
  if (n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int block = 256, grid = phc::grid_for(n, block);
  if (K == 2) hipLaunchKernelGGL(md_batch_kernel<2>, dim3(grid), dim3(block), 0, st, op, n, a, b, c);
  else if (K == 4) hipLaunchKernelGGL(md_batch_kernel<4>, dim3(grid), dim3(block), 0, st, op, n, a, b, c);
  else hipLaunchKernelGGL(md_batch_kernel<8>, dim3(grid), dim3(block), 0, st, op, n, a, b, c);
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_series_product(int K, int cplx, int deg, int nbatch, const double *x, const double *y,
                       double *z, void *stream) {
  PHC_REQUIRE(valid_k(K), "phc_series_product: K must be 1, 2, 4 or 8");
  PHC_REQUIRE(deg >= 0 && nbatch >= 0, "phc_series_product: negative size");
  if (nbatch == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int D = deg + 1, block = 256;
  const int grid = phc::grid_for((size_t)nbatch * D, block);
#define PL(KK, CC) hipLaunchKernelGGL((product_kernel<KK, CC>), dim3(grid), dim3(block), 0, st, D, nbatch, x, y, z)
  if (!cplx) {
    if (K == 1) PL(1, false); else if (K == 2) PL(2, false); else if (K == 4) PL(4, false); else PL(8, false);
  } else {
    if (K == 1) PL(1, true); else if (K == 2) PL(2, true); else if (K == 4) PL(4, true); else PL(8, true);
  }
#undef PL
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_poly_create(phc_poly **plan, int K, int cplx, int dim, int nbr, int deg, const int *nvr,
                    const int *idx, const double *cst, const double *cff) {
  PHC_REQUIRE(plan != nullptr, "phc_poly_create: null plan");
  *plan = nullptr;
  PHC_REQUIRE(valid_k(K), "phc_poly_create: K must be 1, 2, 4 or 8");
  PHC_REQUIRE(dim >= 1 && nbr >= 0 && deg >= 0, "phc_poly_create: bad sizes");
  PHC_REQUIRE(deg + 1 <= 256, "phc_poly_create: degree above 255");
  std::vector<int> ptr(nbr + 1, 0);
  for (int i = 0; i < nbr; ++i) {
    PHC_REQUIRE(nvr[i] >= 1 && nvr[i] <= dim, "phc_poly_create: nvr out of range");
    ptr[i + 1] = ptr[i] + nvr[i];
  }
  for (int j = 0; j < ptr[nbr]; ++j)
    PHC_REQUIRE(idx[j] >= 0 && idx[j] < dim, "phc_poly_create: variable index out of range");
  const int P = cplx ? 2 * K : K, D = deg + 1;
  phc_poly *p = new phc_poly();
  p->K = K; p->cplx = cplx; p->dim = dim; p->nbr = nbr; p->deg = deg;
  p->nvr.assign(nvr, nvr + nbr);
  const size_t ni = (size_t)std::max(ptr[nbr], 1);
  if (hipMalloc(&p->d_nvr, sizeof(int) * std::max(nbr, 1)) != hipSuccess ||
      hipMalloc(&p->d_idxptr, sizeof(int) * (nbr + 1)) != hipSuccess ||
      hipMalloc(&p->d_idx, sizeof(int) * ni) != hipSuccess ||
      hipMalloc(&p->d_cst, sizeof(double) * P * D) != hipSuccess ||
      hipMalloc(&p->d_cff, sizeof(double) * P * std::max(nbr, 1) * D) != hipSuccess) {
    phc_poly_destroy(p);
    phc::set_error("phc_poly_create: device allocation failed");
    return -1;
  }
  if (nbr > 0) {
    PHC_HIP(hipMemcpy(p->d_nvr, nvr, sizeof(int) * nbr, hipMemcpyHostToDevice));
    PHC_HIP(hipMemcpy(p->d_idx, idx, sizeof(int) * ptr[nbr], hipMemcpyHostToDevice));
    PHC_HIP(hipMemcpy(p->d_cff, cff, sizeof(double) * P * nbr * D, hipMemcpyHostToDevice));
  }
  PHC_HIP(hipMemcpy(p->d_idxptr, ptr.data(), sizeof(int) * (nbr + 1), hipMemcpyHostToDevice));
  PHC_HIP(hipMemcpy(p->d_cst, cst, sizeof(double) * P * D, hipMemcpyHostToDevice));
  *plan = p;
  return 0;
}

int phc_poly_destroy(phc_poly *p) {
  if (!p) return 0;
  hipFree(p->d_nvr); hipFree(p->d_idxptr); hipFree(p->d_idx);
  hipFree(p->d_cst); hipFree(p->d_cff); hipFree(p->work);
  delete p;
  return 0;
}

int phc_poly_evaldiff(phc_poly *p, int nbatch, const double *input, double *output, void *stream) {
  PHC_REQUIRE(p != nullptr, "phc_poly_evaldiff: null plan");
  PHC_REQUIRE(nbatch >= 0, "phc_poly_evaldiff: nbatch < 0");
  if (nbatch == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const int P = p->cplx ? 2 * p->K : p->K, D = p->deg + 1;
  const size_t plane = (size_t)nbatch * p->dim * D;
  const size_t need = 3 * plane * P;
  if (need > p->work_cap) {
    PHC_HIP(hipStreamSynchronize(st));
    hipFree(p->work);
    p->work = nullptr;
    p->work_cap = 0;
    PHC_HIP(hipMalloc(&p->work, sizeof(double) * need));
    p->work_cap = need;
  }
  double *fw = p->work, *bw = fw + plane * P, *cr = bw + plane * P;
  int dpad = 1;
  while (dpad < D) dpad *= 2;
  const int block = dpad >= 64 ? dpad : 64;
  const int ipb = block / dpad;
  const int grid = (nbatch + ipb - 1) / ipb;
#define LP(KK, CC) launch_poly<KK, CC>(grid, block, st, p->dim, p->nbr, D, p->d_nvr, p->d_idxptr, p->d_idx, p->d_cst, p->d_cff, nbatch, input, output, fw, bw, cr, ipb, dpad)
  if (!p->cplx) {
    if (p->K == 1) LP(1, false); else if (p->K == 2) LP(2, false); else if (p->K == 4) LP(4, false); else LP(8, false);
  } else {
    if (p->K == 1) LP(1, true); else if (p->K == 2) LP(2, true); else if (p->K == 4) LP(4, true); else LP(8, true);
  }
#undef LP
  PHC_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
