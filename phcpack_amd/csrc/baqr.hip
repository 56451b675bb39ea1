// Blocked Householder QR with explicit Q, and tiled back substitution, of
// complex multiple-double matrices (K = 2 double-double, K = 4 quad-double)
// on gfx950.
//
//   reference (Matrices/)                       here
//   CPU_cmplx4_blocked_houseqr (dbl4_baqr_host.cpp:1206),
//   GPU_cmplx4_blocked_houseqr (dbl4_baqr_kernels.cu:6197)
//     per column: house, panel leftRupdate      house_kernel, panel_dot_kernel,
//                                               panel_update_kernel (many workgroups
//                                               per column, partial sums reduced in
//                                               a fixed order)
//     VB_to_W (dbl4_baqr_host.cpp:91)           W = Y S with S upper triangular from
//                                               the Gram matrix G = Y^H Y: one
//                                               wavefront per row of S
//     blocked leftRupdate (:357):               C += Y (W^H C): two thin products
//       C += (Y W^H) C                          instead of the rowdim x rowdim Y W^H
//     blocked rightQupdate (:743):              Q += (Q W) Y^H, likewise
//       Q += Q (W Y^H)
//   CPU_cmplx4_upper_tiled_solver               tile_inverse_kernel (one workgroup per
//     (dbl4_tabs_host.cpp:345),                 diagonal tile, one lane per column),
//   GPU_cmplx4_upper_tiled_solver               tiled_solve_kernel (tile products in
//     (dbl4_tabs_kernels.cu:1802)               order, one workgroup)
//
// Every product runs through one complex multiple-double GEMM kernel (32 x 32
// output tiles, 8-deep k slices staged in LDS, 2 x 2 outputs per lane, each
// output summed in ascending k).  Results agree with the reference's host
// code to a few units in the last limb (the order of the long sums differs);
// the tests bound the difference.
#include <vector>

#include "common.hpp"
#include "mdtypes.hpp"

using namespace md;

namespace {

template <int K> using Z = cplx<K>;
template <int K> using T = typename RealT<K>::type;

// element (r, c) of a strided planar operand: p[s0*r + s1*c], planes ps apart
struct Op {
  const double *p;
  size_t ps;
  long s0, s1;
};

// acc += op(a) op(b), op = conjugate when the flag is set
template <int K, bool CA, bool CB>
__device__ __forceinline__ void cmac(Z<K> &acc, const Z<K> &a, const Z<K> &b) {
  const T<K> t1 = mul(a.re, b.re), t2 = mul(a.im, b.im);
  const T<K> t3 = mul(a.im, b.re), t4 = mul(a.re, b.im);
  acc.re = add(acc.re, t1);
  acc.re = (CA != CB) ? add(acc.re, t2) : sub(acc.re, t2);
  acc.im = CA ? sub(acc.im, t3) : add(acc.im, t3);
  acc.im = CB ? sub(acc.im, t4) : add(acc.im, t4);
}

constexpr int TM = 32, TN = 32, TKD = 8;

// C(i, j) = [C(i, j) +] sum_k op(A(i, k)) op(B(k, j)), i < M, j < N, k < KD;
// the sum starts from zero and is added to C at the end.
template <int K, bool CA, bool CB>
__global__ void __launch_bounds__(256)
cgemm_kernel(int M, int N, int KD, Op A, Op B, double *Cp, size_t cps, long c0, long c1,
             int accumulate) {
  __shared__ double As[2 * K][TKD][TM];
  __shared__ double Bs[2 * K][TKD][TN];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  Z<K> acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = C<K>::zero();
  for (int k0 = 0; k0 < KD; k0 += TKD) {
    {
      int ii, kk;
      if (A.s0 == 1) { ii = t % TM; kk = t / TM; } else { kk = t % TKD; ii = t / TKD; }
      const int gi = i0 + ii, gk = k0 + kk;
      const bool in = gi < M && gk < KD;
      const size_t e = in ? (size_t)(A.s0 * gi + A.s1 * gk) : 0;
#pragma unroll
      for (int q = 0; q < 2 * K; ++q) As[q][kk][ii] = in ? A.p[q * A.ps + e] : 0.0;
    }
    {
      int jj, kk;
      if (B.s1 == 1) { jj = t % TN; kk = t / TN; } else { kk = t % TKD; jj = t / TKD; }
      const int gj = j0 + jj, gk = k0 + kk;
      const bool in = gj < N && gk < KD;
      const size_t e = in ? (size_t)(B.s0 * gk + B.s1 * gj) : 0;
#pragma unroll
      for (int q = 0; q < 2 * K; ++q) Bs[q][kk][jj] = in ? B.p[q * B.ps + e] : 0.0;
    }
    __syncthreads();
    const int kmax = (KD - k0) < TKD ? (KD - k0) : TKD;
    for (int kk = 0; kk < kmax; ++kk) {
      Z<K> av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = C<K>::load(&As[0][0][0], TKD * TM, kk * TM + ty + 16 * a);
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = C<K>::load(&Bs[0][0][0], TKD * TN, kk * TN + tx + 16 * b);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) cmac<K, CA, CB>(acc[a][b], av[a], bv[b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
      if (i < M && j < N) {
        const size_t e = (size_t)(c0 * i + c1 * j);
        Z<K> v = acc[a][b];
        if (accumulate) v = C<K>::add(C<K>::load(Cp, cps, e), v);
        C<K>::store(Cp, cps, e, v);
      }
    }
}

template <int K>
int cgemm(bool ca, bool cb, int M, int N, int KD, Op A, Op B, double *Cp, size_t cps, long c0,
          long c1, bool accumulate, hipStream_t st) {
  if (M <= 0 || N <= 0) return 0;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  const int acc = accumulate ? 1 : 0;
  if (!ca && !cb)
    hipLaunchKernelGGL((cgemm_kernel<K, false, false>), grid, dim3(256), 0, st, M, N, KD, A, B, Cp, cps, c0, c1, acc);
  else if (ca && !cb)
    hipLaunchKernelGGL((cgemm_kernel<K, true, false>), grid, dim3(256), 0, st, M, N, KD, A, B, Cp, cps, c0, c1, acc);
  else if (!ca && cb)
    hipLaunchKernelGGL((cgemm_kernel<K, false, true>), grid, dim3(256), 0, st, M, N, KD, A, B, Cp, cps, c0, c1, acc);
  else
    hipLaunchKernelGGL((cgemm_kernel<K, true, true>), grid, dim3(256), 0, st, M, N, KD, A, B, Cp, cps, c0, c1, acc);
  PHC_HIP(hipGetLastError());
  return 0;
}

// --------------------------------------------------------- panel kernels

// Householder vector of column colidx of R (rows colidx..nrows-1), as
// CPU_cmplx2_factors_house (Matrices/dbl2_factorizations.cpp:441-533):
// v is written to Y row L from position L on (zeros before), beta to B[L].
template <int K>
__global__ void __launch_bounds__(256)
house_kernel(const double *Rm, size_t rps, int ld, int colidx, int nrows, int L, double *Y,
             size_t yps, int rowdim, double *Bv, int szt) {
  __shared__ T<K> part[256];
  __shared__ T<K> sbeta;
  __shared__ Z<K> sinv;
  __shared__ int szero;
  const int t = threadIdx.x, n = nrows - colidx;
  auto x = [&](int i) { return C<K>::load(Rm, rps, (size_t)(colidx + i) * ld + colidx); };
  T<K> s = R<K>::zero();
  for (int i = 1 + t; i < n; i += 256) {
    const Z<K> xi = x(i);
    s = add(s, sqr(xi.re));
    s = add(s, sqr(xi.im));
  }
  part[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) part[t] = add(part[t], part[t + h]);
    __syncthreads();
  }
  if (t == 0) {
    T<K> sigma = part[0];
    bool zero = true;
    for (int q = 0; q < K; ++q) zero = zero && (sigma.x[q] == 0.0);
    szero = zero ? 1 : 0;
    if (zero) {
      sbeta = R<K>::zero();
      sinv = C<K>::from(1.0, 0.0);
    } else {
      const Z<K> x0 = x(0);
      T<K> sqrx0 = sqr(x0.re);
      sqrx0 = add(sqrx0, sqr(x0.im));
      const T<K> x0rad = R<K>::sqrt(sqrx0);
      sqrx0 = add(sqrx0, sigma);
      T<K> mu = R<K>::sqrt(sqrx0);
      Z<K> v0;
      bool rzero = true;
      for (int q = 0; q < K; ++q) rzero = rzero && (x0rad.x[q] == 0.0);
      if (rzero) {
        v0.re = neg(mu);
        v0.im = R<K>::zero();
      } else {
        mu = div(mu, x0rad);
        v0.re = sub(x0.re, mul(mu, x0.re));
        v0.im = sub(x0.im, mul(mu, x0.im));
      }
      T<K> sqrv0 = sqr(v0.re);
      sqrv0 = add(sqrv0, sqr(v0.im));
      sigma = add(sigma, sqrv0);
      T<K> beta = div(sqrv0, sigma);
      sbeta = mul_d(beta, 2.0);
      sinv.re = div(v0.re, sqrv0);
      sinv.im = neg(div(v0.im, sqrv0));
    }
    R<K>::store(Bv, szt, L, sbeta);
  }
  __syncthreads();
  const Z<K> inv = sinv;
  const bool zero = szero != 0;
  for (int i = t; i < rowdim; i += 256) {
    Z<K> v;
    if (i < L) v = C<K>::zero();
    else if (i == L) v = C<K>::from(1.0, 0.0);
    else {
      const Z<K> xi = x(i - L);
      if (zero) v = xi;
      else {
        v.re = sub(mul(xi.re, inv.re), mul(xi.im, inv.im));
        v.im = add(mul(xi.im, inv.re), mul(xi.re, inv.im));
      }
    }
    C<K>::store(Y, yps, (size_t)L * rowdim + i, v);
  }
}

constexpr int RCH = 64;  // rows per chunk of the panel dot products

// partial sums wp[chunk][j] = sum_{i in chunk} conj(R_ij) v_i over columns
// j = colidx..endcol-1 (one lane per column)
template <int K>
__global__ void __launch_bounds__(128)
panel_dot_kernel(const double *Rm, size_t rps, int ld, int colidx, int endcol, int nrows,
                 const double *Y, size_t yps, int rowdim, int L, double *wp, size_t wps,
                 int szt) {
  __shared__ double vs[2 * K][RCH];
  const int chunk = blockIdx.x, t = threadIdx.x;
  const int r0 = colidx + chunk * RCH, n = nrows - colidx;
  const int rows = (n - chunk * RCH) < RCH ? (n - chunk * RCH) : RCH;
  for (int i = t; i < rows; i += blockDim.x) {
    const Z<K> v = C<K>::load(Y, yps, (size_t)L * rowdim + L + chunk * RCH + i);
    for (int q = 0; q < K; ++q) { vs[q][i] = v.re.x[q]; vs[K + q][i] = v.im.x[q]; }
  }
  __syncthreads();
  const int j = colidx + t;
  if (j >= endcol) return;
  Z<K> s = C<K>::zero();
  for (int i = 0; i < rows; ++i) {
    const Z<K> r = C<K>::load(Rm, rps, (size_t)(r0 + i) * ld + j);
    const Z<K> v = C<K>::load(&vs[0][0], RCH, i);
    cmac<K, true, false>(s, r, v);
  }
  C<K>::store(wp, wps, (size_t)chunk * szt + t, s);
}

// w_j = beta * sum_chunks wp[.][j]; R_ij -= v_i conj(w_j) for the rows of
// this workgroup
template <int K>
__global__ void __launch_bounds__(128)
panel_update_kernel(double *Rm, size_t rps, int ld, int colidx, int endcol, int nrows,
                    const double *Y, size_t yps, int rowdim, int L, const double *wp,
                    size_t wps, int szt, const double *Bv, int nchunk, int rows_per_block) {
  const int t = threadIdx.x, j = colidx + t;
  if (j >= endcol) return;
  const T<K> beta = R<K>::load(Bv, szt, L);
  Z<K> w = C<K>::zero();
  for (int c = 0; c < nchunk; ++c) w = C<K>::add(w, C<K>::load(wp, wps, (size_t)c * szt + t));
  w.re = mul(beta, w.re);
  w.im = mul(beta, w.im);
  const int ib = colidx + blockIdx.x * rows_per_block;
  const int ie = (ib + rows_per_block) < nrows ? ib + rows_per_block : nrows;
  for (int i = ib; i < ie; ++i) {
    const Z<K> v = C<K>::load(Y, yps, (size_t)L * rowdim + L + (i - colidx));
    const size_t e = (size_t)i * ld + j;
    Z<K> r = C<K>::load(Rm, rps, e);
    // r -= v * conj(w)
    r.re = sub(r.re, add(mul(v.re, w.re), mul(v.im, w.im)));
    r.im = sub(r.im, sub(mul(v.im, w.re), mul(v.re, w.im)));
    C<K>::store(Rm, rps, e, r);
  }
}

// S (szt x szt, upper triangular, row-major) with W = Y S:
// S[j][j] = -B_j, S[m][j] = -B_j sum_{k=m}^{j-1} S[m][k] G[k][j]; one
// wavefront per row m (kept in LDS), lanes split the sum over k
template <int K>
__global__ void __launch_bounds__(64)
s_factor_kernel(const double *G, size_t gps, const double *Bv, double *S, size_t sps, int szt) {
  __shared__ double row[2 * K][256];
  const int m = blockIdx.x, l = threadIdx.x;
  for (int j = 0; j < szt; ++j) {
    Z<K> s = C<K>::zero();
    if (j > m) {
      for (int k = m + l; k < j; k += 64)
        s = C<K>::add(s, C<K>::mul(C<K>::load(&row[0][0], 256, k),
                                    C<K>::load(G, gps, (size_t)k * szt + j)));
      for (int o = 32; o > 0; o >>= 1) {
        Z<K> u;
        for (int q = 0; q < K; ++q) {
          u.re.x[q] = __shfl_xor(s.re.x[q], o, 64);
          u.im.x[q] = __shfl_xor(s.im.x[q], o, 64);
        }
        s = C<K>::add(s, u);
      }
    }
    if (l == 0) {
      const T<K> b = neg(R<K>::load(Bv, szt, j));
      Z<K> v;
      if (j < m) v = C<K>::zero();
      else if (j == m) { v.re = b; v.im = R<K>::zero(); }
      else { v.re = mul(b, s.re); v.im = mul(b, s.im); }
      C<K>::store(&row[0][0], 256, j, v);
    }
    __syncthreads();
  }
  for (int j = l; j < szt; j += 64) C<K>::store(S, sps, (size_t)m * szt + j, C<K>::load(&row[0][0], 256, j));
}

// ------------------------------------------------------- tiled solver

// inverse of the diagonal tile `blockIdx.x` of U (dim x dim), one lane per
// column: x = T^{-1} e_j by back substitution (CPU_cmplx2_factors_backward
// order), written to inv[tile][szt][szt]
template <int K>
__global__ void __launch_bounds__(256)
tile_inverse_kernel(const double *U, size_t ups, int dim, int szt, double *inv, size_t ips) {
  const int tile = blockIdx.x, off = tile * szt;
  const size_t ib = (size_t)tile * szt * szt;
  for (int j = threadIdx.x; j < szt; j += blockDim.x) {
    for (int i = szt - 1; i >= 0; --i) {
      Z<K> x = (i == j) ? C<K>::from(1.0, 0.0) : C<K>::zero();
      if (i <= j) {
        for (int k = j; k > i; --k) {
          const Z<K> u = C<K>::load(U, ups, (size_t)(off + i) * dim + off + k);
          const Z<K> xk = C<K>::load(inv, ips, ib + (size_t)k * szt + j);
          x = C<K>::sub(x, C<K>::mul(u, xk));
        }
        const Z<K> d = C<K>::load(U, ups, (size_t)(off + i) * dim + off + i);
        x = C<K>::div(x, d);
      }
      C<K>::store(inv, ips, ib + (size_t)i * szt + j, x);
    }
  }
}

// the tile stages of CPU_cmplx2_upper_tiled_solver (Matrices/dbl2_tabs_host.cpp:289-513)
// with the inverses precomputed: x_last = invT b_last; for k = nbt-1..1:
// b_L -= U_{L,k} x_k (L < k), x_{k-1} = invT_{k-1} b_{k-1}.  One workgroup.
template <int K>
__global__ void __launch_bounds__(256)
tiled_solve_kernel(const double *U, size_t ups, int dim, int szt, int nbt, const double *inv,
                   size_t ips, double *b, size_t bps, double *x, size_t xps) {
  const int t = threadIdx.x;
  auto matvec = [&](int row0, const double *M, size_t mps, size_t mb, int ldm, int col0,
                    int rowx, bool sub_from_b) {
    for (int i = t; i < szt; i += blockDim.x) {
      Z<K> s = C<K>::zero();
      for (int j = 0; j < szt; ++j)
        s = C<K>::add(s, C<K>::mul(C<K>::load(M, mps, mb + (size_t)(row0 + i) * ldm + col0 + j),
                                    C<K>::load(sub_from_b ? x : b, sub_from_b ? xps : bps,
                                               (size_t)rowx + j)));
      if (sub_from_b) {
        const size_t e = (size_t)row0 + i;
        C<K>::store(b, bps, e, C<K>::sub(C<K>::load(b, bps, e), s));
      } else {
        C<K>::store(x, xps, (size_t)rowx + i, s);
      }
    }
  };
  // x_last
  const int last = nbt - 1;
  matvec(0, inv, ips, (size_t)last * szt * szt, szt, 0, last * szt, false);
  __syncthreads();
  for (int k = nbt - 1; k > 0; --k) {
    for (int L = 0; L < k; ++L) matvec(L * szt, U, ups, 0, dim, k * szt, k * szt, true);
    __syncthreads();
    matvec(0, inv, ips, (size_t)(k - 1) * szt * szt, szt, 0, (k - 1) * szt, false);
    __syncthreads();
  }
}

// copy the inverse tiles into the diagonal of U (the reference overwrites them)
template <int K>
__global__ void copy_inverse_kernel(double *U, size_t ups, int dim, int szt, int nbt,
                                    const double *inv, size_t ips) {
  const size_t n = (size_t)nbt * szt * szt;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int tile = (int)(e / ((size_t)szt * szt)), r = (int)(e / szt % szt), c = (int)(e % szt);
    C<K>::store(U, ups, (size_t)(tile * szt + r) * dim + tile * szt + c, C<K>::load(inv, ips, e));
  }
}

template <int K>
__global__ void init_qr_kernel(int nrows, int ncols, const double *A, double *Q, double *Rm) {
  const size_t nq = (size_t)nrows * nrows, nr = (size_t)nrows * ncols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nq + nr;
       e += (size_t)gridDim.x * blockDim.x) {
    if (e < nq) {
      const size_t i = e / nrows, j = e % nrows;
      C<K>::store(Q, nq, e, C<K>::from(i == j ? 1.0 : 0.0, 0.0));
    } else {
      C<K>::store(Rm, nr, e - nq, C<K>::load(A, nr, e - nq));
    }
  }
}

}  // namespace

struct phc_baqr {
  int K, nrows, ncols, szt, nbt, nchunk;
  double *Y, *W, *G, *S, *T1, *T2, *wp, *B;
  size_t yps, gps, t1ps, t2ps, wps;
};

namespace {

template <int K> int run_qr(phc_baqr *p, const double *A, double *Q, double *R, hipStream_t st) {
  const int m = p->nrows, n = p->ncols, szt = p->szt;
  const size_t qps = (size_t)m * m, rps = (size_t)m * n;
  hipLaunchKernelGGL(init_qr_kernel<K>, dim3(phc::grid_for(qps + rps, 256)), dim3(256), 0, st, m, n, A, Q, R);
  PHC_HIP(hipGetLastError());
  for (int k = 0; k < p->nbt; ++k) {
    const int rowoff = k * szt, rowdim = m - rowoff, endcol = (k + 1) * szt;
    for (int L = 0; L < szt; ++L) {
      const int colidx = rowoff + L, nr = m - colidx;
      const int nchunk = (nr + RCH - 1) / RCH;
      hipLaunchKernelGGL(house_kernel<K>, dim3(1), dim3(256), 0, st, R, rps, n, colidx, m, L,
                         p->Y, p->yps, rowdim, p->B, szt);
      hipLaunchKernelGGL(panel_dot_kernel<K>, dim3(nchunk), dim3(128), 0, st, R, rps, n, colidx,
                         endcol, m, p->Y, p->yps, rowdim, L, p->wp, p->wps, szt);
      const int rpb = 16, nblk = (nr + rpb - 1) / rpb;
      hipLaunchKernelGGL(panel_update_kernel<K>, dim3(nblk), dim3(128), 0, st, R, rps, n, colidx,
                         endcol, m, p->Y, p->yps, rowdim, L, p->wp, p->wps, szt, p->B, nchunk, rpb);
    }
    PHC_HIP(hipGetLastError());
    // Y, W stored as szt rows of length rowdim (row stride rowdim)
    const Op Yt{p->Y, p->yps, 1, rowdim};          // Yhat(i, k) = Y[k][i]
    const Op Yc{p->Y, p->yps, rowdim, 1};          // (k, i) -> Y[k][i]
    // G = Yhat^H Yhat
    if (cgemm<K>(true, false, szt, szt, rowdim, Op{p->Y, p->yps, rowdim, 1}, Yt, p->G, p->gps,
                 szt, 1, false, st)) return -1;
    hipLaunchKernelGGL(s_factor_kernel<K>, dim3(szt), dim3(64), 0, st, p->G, p->gps, p->B, p->S,
                       p->gps, szt);
    PHC_HIP(hipGetLastError());
    // W = Yhat S, stored like Y
    if (cgemm<K>(false, false, rowdim, szt, szt, Yt, Op{p->S, p->gps, szt, 1}, p->W, p->yps, 1,
                 rowdim, false, st)) return -1;
    const Op Wt{p->W, p->yps, 1, rowdim};          // What(i, k) = W[k][i]
    if (k < p->nbt - 1) {
      const int coloff = endcol, coldim = n - coloff;
      // T1 = What^H C, C = R[rowoff.., coloff..]
      const Op Cm{R + (size_t)rowoff * n + coloff, rps, n, 1};
      if (cgemm<K>(true, false, szt, coldim, rowdim, Op{p->W, p->yps, rowdim, 1}, Cm, p->T1,
                   p->t1ps, coldim, 1, false, st)) return -1;
      // C += Yhat T1
      if (cgemm<K>(false, false, rowdim, coldim, szt, Yt, Op{p->T1, p->t1ps, coldim, 1},
                   R + (size_t)rowoff * n + coloff, rps, n, 1, true, st)) return -1;
    }
    // T2 = Q[:, rowoff..] What;  Q[:, rowoff..] += T2 Yhat^H
    const Op Qm{Q + rowoff, qps, m, 1};
    if (cgemm<K>(false, false, m, szt, rowdim, Qm, Wt, p->T2, p->t2ps, szt, 1, false, st)) return -1;
    if (cgemm<K>(false, true, m, rowdim, szt, Op{p->T2, p->t2ps, szt, 1}, Yc, Q + rowoff, qps, m, 1,
                 true, st)) return -1;
  }
  return 0;
}

}  // namespace

extern "C" {

int phc_baqr_create(phc_baqr **out, int K, int nrows, int ncols, int szt) {
  PHC_REQUIRE(out, "phc_baqr_create: null plan pointer");
  PHC_REQUIRE(K == 2 || K == 4, "phc_baqr_create: K must be 2 or 4");
  PHC_REQUIRE(szt >= 1 && szt <= 256 && ncols >= szt && ncols % szt == 0,
              "phc_baqr_create: ncols must be a multiple of szt (1..256)");
  PHC_REQUIRE(nrows >= ncols, "phc_baqr_create: nrows must be at least ncols");
  phc_baqr *p = new phc_baqr();
  p->K = K; p->nrows = nrows; p->ncols = ncols; p->szt = szt; p->nbt = ncols / szt;
  p->nchunk = (nrows + RCH - 1) / RCH;
  p->yps = (size_t)szt * nrows;
  p->gps = (size_t)szt * szt;
  p->t1ps = (size_t)szt * ncols;
  p->t2ps = (size_t)nrows * szt;
  p->wps = (size_t)p->nchunk * szt;
  const size_t z = sizeof(double) * 2 * K;
  double **bufs[] = {&p->Y, &p->W, &p->G, &p->S, &p->T1, &p->T2, &p->wp, &p->B};
  const size_t sizes[] = {z * p->yps, z * p->yps, z * p->gps, z * p->gps, z * p->t1ps,
                          z * p->t2ps, z * p->wps, sizeof(double) * K * szt};
  for (int i = 0; i < 8; ++i) *bufs[i] = nullptr;
  for (int i = 0; i < 8; ++i) {
    if (hipMalloc(bufs[i], sizes[i]) != hipSuccess) {
      for (int j = 0; j < i; ++j) hipFree(*bufs[j]);
      delete p;
      phc::set_error("phc_baqr_create: hipMalloc failed");
      return -1;
    }
  }
  *out = p;
  return 0;
}

int phc_baqr_destroy(phc_baqr *p) {
  if (!p) return 0;
  double *bufs[] = {p->Y, p->W, p->G, p->S, p->T1, p->T2, p->wp, p->B};
  for (double *b : bufs) hipFree(b);
  delete p;
  return 0;
}

int phc_cmplx_blocked_houseqr(phc_baqr *p, const double *A, double *Q, double *R, void *stream) {
  PHC_REQUIRE(p && A && Q && R, "phc_cmplx_blocked_houseqr: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int rc = (p->K == 2) ? run_qr<2>(p, A, Q, R, st) : run_qr<4>(p, A, Q, R, st);
  if (rc) return rc;
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_cmplx_gemm(int K, int conjA, int conjB, int M, int N, int KD, const double *A,
                   size_t aps, long a0, long a1, const double *B, size_t bps, long b0, long b1,
                   double *Cm, size_t cps, long c0, long c1, int accumulate, void *stream) {
  PHC_REQUIRE(K == 2 || K == 4, "phc_cmplx_gemm: K must be 2 or 4");
  PHC_REQUIRE(M >= 0 && N >= 0 && KD >= 0, "phc_cmplx_gemm: negative size");
  PHC_REQUIRE(A && B && Cm, "phc_cmplx_gemm: null array");
  hipStream_t st = (hipStream_t)stream;
  const Op opa{A, aps, a0, a1}, opb{B, bps, b0, b1};
  return K == 2 ? cgemm<2>(conjA, conjB, M, N, KD, opa, opb, Cm, cps, c0, c1, accumulate, st)
                : cgemm<4>(conjA, conjB, M, N, KD, opa, opb, Cm, cps, c0, c1, accumulate, st);
}

int phc_cmplx_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b,
                                 double *x, void *stream) {
  PHC_REQUIRE(K == 2 || K == 4, "phc_cmplx_upper_tiled_solver: K must be 2 or 4");
  PHC_REQUIRE(szt >= 1 && nbt >= 1 && dim == szt * nbt,
              "phc_cmplx_upper_tiled_solver: dim must equal szt*nbt");
  PHC_REQUIRE(U && b && x, "phc_cmplx_upper_tiled_solver: null array");
  hipStream_t st = (hipStream_t)stream;
  double *inv = nullptr;
  const size_t ips = (size_t)nbt * szt * szt;
  PHC_HIP(hipMalloc(&inv, sizeof(double) * 2 * K * ips));
  const size_t ups = (size_t)dim * dim;
#define SOLVE(KK)                                                                              \
  do {                                                                                         \
    hipLaunchKernelGGL(tile_inverse_kernel<KK>, dim3(nbt), dim3(256), 0, st, U, ups, dim, szt, \
                       inv, ips);                                                              \
    hipLaunchKernelGGL(tiled_solve_kernel<KK>, dim3(1), dim3(256), 0, st, U, ups, dim, szt,    \
                       nbt, inv, ips, b, (size_t)dim, x, (size_t)dim);                         \
    hipLaunchKernelGGL(copy_inverse_kernel<KK>, dim3(phc::grid_for(ips, 256)), dim3(256), 0, st, U, ups, \
                       dim, szt, nbt, inv, ips);                                               \
  } while (0)
  if (K == 2) SOLVE(2); else SOLVE(4);
#undef SOLVE
  hipError_t e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(st);
  hipFree(inv);
  PHC_HIP(e);
  PHC_HIP(e2);
  return 0;
}

}  // extern "C"
