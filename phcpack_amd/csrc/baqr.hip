// Blocked Householder QR with explicit Q, and tiled back substitution, of
// complex multiple-double matrices (K = 2 double-double, K = 4 quad-double)
// on gfx950.
//
//   reference (Matrices/)                       here
//   CPU_cmplx4_blocked_houseqr (dbl4_baqr_host.cpp:1206),
//   GPU_cmplx4_blocked_houseqr (dbl4_baqr_kernels.cu:6197)
//     per column: house, panel leftRupdate      house_kernel (forms v as well),
//                                               column_dot_kernel and column_upd_kernel
//                                               (workgroups per panel column and split
//                                               of its rows), on a column-major copy
//                                               of the panel; the launches of one
//                                               factorization replay as a HIP graph
//     VB_to_W (dbl4_baqr_host.cpp:91)           W = Y S with S upper triangular from
//                                               the Gram matrix G = Y^H Y: one
//                                               wavefront per row of S
//     blocked leftRupdate (:357):               C += Y (W^H C): two thin products
//       C += (Y W^H) C                          instead of the rowdim x rowdim Y W^H
//     blocked rightQupdate (:743):              Q += (Q W) Y^H, likewise
//       Q += Q (W Y^H)
//   CPU_cmplx4_upper_tiled_solver               tile_inverse_kernel (one wavefront per
//     (dbl4_tabs_host.cpp:345),                 column of every diagonal tile), then
//   GPU_cmplx4_upper_tiled_solver               rowdot_kernel stages (one wavefront
//     (dbl4_tabs_kernels.cu:1802)               per row: x_k = inv_k b_k, then
//                                               b_L -= U_Lk x_k for all L < k)
//   GPU_cmplx4_bals_qhb                         the GEMM with a split k range
//     (Newton/dbl4_bals_kernels.cu:602)
//
// Every product runs through one complex multiple-double GEMM kernel (32 x 32
// output tiles, 8-deep k slices staged in LDS, 2 x 2 outputs per lane); long
// k ranges are split over workgroups whose partial sums are added in slice
// order by a second kernel.  Results agree with the reference's host
// code to a few units in the last limb (the order of the long sums differs);
// the tests bound the difference.
#include <mutex>
#include <vector>

#include "common.hpp"
#include "mdtypes.hpp"
#include "wave.hpp"

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

// A complex limb-planar array, or (RE, the real QR) an array of K real
// planes read with zero imaginary parts and written as its real parts.
template <int K, bool RE> struct Acc {
  static __device__ __forceinline__ Z<K> load(const double *p, size_t ps, size_t i) {
    if constexpr (RE) {
      Z<K> z;
      z.re = R<K>::load(p, ps, i);
      z.im = R<K>::zero();
      return z;
    } else {
      return C<K>::load(p, ps, i);
    }
  }
  static __device__ __forceinline__ void store(double *p, size_t ps, size_t i, const Z<K> &z) {
    if constexpr (RE) R<K>::store(p, ps, i, z.re);
    else C<K>::store(p, ps, i, z);
  }
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

constexpr int TKD = 8;  // k depth of the LDS slices

// Launch configuration of the GEMM, chosen per QR plan (phc_baqr_set_tuning,
// by the autotuner of phcpack_amd/autotune.py):
//   block: register block RA x RB (outputs per lane along i and j; the
//     16 x 16 lanes of a workgroup cover a (16 RA) x (16 RB) tile):
//     0 the default of the limb count (2 x 2 for dd, 1 x 1 for qd and od),
//     1: 1 x 1, 2: 2 x 1, 3: 1 x 2, 4: 2 x 2;
//   target: the k range is split into slices (doubling) while the tiles
//     times the slices stay below target (0: 1024).
struct GemmCfg {
  int block = 0, target = 0;
};
template <int K> constexpr int default_block() { return (K >= 4) ? 1 : 4; }
inline int block_ra(int blk) { return (blk == 2 || blk == 4) ? 2 : 1; }
inline int block_rb(int blk) { return (blk >= 3) ? 2 : 1; }

// C(i, j) = [C(i, j) +] sum_k op(A(i, k)) op(B(k, j)), i < M, j < N, k < KD;
// the sum starts from zero and is added to C at the end.  With a split of
// the k range over gridDim.z slices, slice s writes its partial sums to
// part[s][i][j] (plane stride pps) and csum_kernel adds the slices in order.
// RE: real operands (K planes, no conjugation), real accumulators
template <int K, bool CA, bool CB, int RA, int RB, bool RE>
__global__ void __launch_bounds__(256)
cgemm_kernel(int M, int N, int KD, int kchunk, Op A, Op B, double *Cp, size_t cps, long c0,
             long c1, int accumulate, double *part, size_t pps) {
  constexpr int TM = 16 * RA, TN = 16 * RB, PL = RE ? K : 2 * K;
  __shared__ double As[2 * K][TKD][TM];
  __shared__ double Bs[2 * K][TKD][TN];
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  const int kb = blockIdx.z * kchunk;
  const int ke = (kb + kchunk) < KD ? (kb + kchunk) : KD;
  Z<K> acc[RA][RB];
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) acc[a][b] = C<K>::zero();
  for (int k0 = kb; k0 < ke; k0 += TKD) {
#pragma unroll
    for (int e = t; e < TM * TKD; e += 256) {
      int ii, kk;
      if (A.s0 == 1) { ii = e % TM; kk = e / TM; } else { kk = e % TKD; ii = e / TKD; }
      const int gi = i0 + ii, gk = k0 + kk;
      const bool in = gi < M && gk < ke;
      const size_t o = in ? (size_t)(A.s0 * gi + A.s1 * gk) : 0;
#pragma unroll
      for (int q = 0; q < PL; ++q) As[q][kk][ii] = in ? A.p[q * A.ps + o] : 0.0;
    }
#pragma unroll
    for (int e = t; e < TN * TKD; e += 256) {
      int jj, kk;
      if (B.s1 == 1) { jj = e % TN; kk = e / TN; } else { kk = e % TKD; jj = e / TKD; }
      const int gj = j0 + jj, gk = k0 + kk;
      const bool in = gj < N && gk < ke;
      const size_t o = in ? (size_t)(B.s0 * gk + B.s1 * gj) : 0;
#pragma unroll
      for (int q = 0; q < PL; ++q) Bs[q][kk][jj] = in ? B.p[q * B.ps + o] : 0.0;
    }
    __syncthreads();
    const int kmax = (ke - k0) < TKD ? (ke - k0) : TKD;
    for (int kk = 0; kk < kmax; ++kk) {
      if constexpr (RE) {
        T<K> av[RA], bv[RB];
#pragma unroll
        for (int a = 0; a < RA; ++a) av[a] = R<K>::load(&As[0][0][0], TKD * TM, kk * TM + ty + 16 * a);
#pragma unroll
        for (int b = 0; b < RB; ++b) bv[b] = R<K>::load(&Bs[0][0][0], TKD * TN, kk * TN + tx + 16 * b);
#pragma unroll
        for (int a = 0; a < RA; ++a)
#pragma unroll
          for (int b = 0; b < RB; ++b) acc[a][b].re = add(acc[a][b].re, mul(av[a], bv[b]));
      } else {
        Z<K> av[RA], bv[RB];
#pragma unroll
        for (int a = 0; a < RA; ++a) av[a] = C<K>::load(&As[0][0][0], TKD * TM, kk * TM + ty + 16 * a);
#pragma unroll
        for (int b = 0; b < RB; ++b) bv[b] = C<K>::load(&Bs[0][0][0], TKD * TN, kk * TN + tx + 16 * b);
#pragma unroll
        for (int a = 0; a < RA; ++a)
#pragma unroll
          for (int b = 0; b < RB; ++b) cmac<K, CA, CB>(acc[a][b], av[a], bv[b]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < RA; ++a)
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
      if (i < M && j < N) {
        if (gridDim.z > 1) {
          Acc<K, RE>::store(part, pps, (size_t)blockIdx.z * M * N + (size_t)i * N + j, acc[a][b]);
        } else {
          const size_t e = (size_t)(c0 * i + c1 * j);
          Z<K> v = acc[a][b];
          if (accumulate) {
            if constexpr (RE) v.re = add(R<K>::load(Cp, cps, e), v.re);
            else v = C<K>::add(C<K>::load(Cp, cps, e), v);
          }
          Acc<K, RE>::store(Cp, cps, e, v);
        }
      }
    }
}

// C(i, j) = [C(i, j) +] sum_s part[s][i][j], slices in order
template <int K, bool RE>
__global__ void csum_kernel(int M, int N, int S, const double *part, size_t pps, double *Cp,
                            size_t cps, long c0, long c1, int accumulate) {
  typedef Acc<K, RE> A_;
  const size_t n = (size_t)M * N;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    Z<K> v = A_::load(part, pps, e);
    for (int s = 1; s < S; ++s) v = C<K>::add(v, A_::load(part, pps, (size_t)s * n + e));
    const size_t i = e / N, j = e % N;
    const size_t o = (size_t)(c0 * (long)i + c1 * (long)j);
    if (accumulate) v = C<K>::add(A_::load(Cp, cps, o), v);
    A_::store(Cp, cps, o, v);
  }
}

// split of the k range: enough workgroups to fill the chip (tiles times
// slices up to the target), slices of at least 64 terms, and M*N*S partial
// sums within the workspace
inline int ksplit(int M, int N, int KD, size_t cap, int TM, int TN, int target) {
  const long tiles = (long)((M + TM - 1) / TM) * ((N + TN - 1) / TN);
  int S = 1;
  while (tiles * S < target && KD / (2 * S) >= 64 && (size_t)M * N * 2 * S <= cap) S *= 2;
  return S;
}

template <int K, bool CA, bool CB, bool RE>
void launch_gemm(int blk, dim3 grid, hipStream_t st, int M, int N, int KD, int kchunk, Op A,
                 Op B, double *Cp, size_t cps, long c0, long c1, int acc, double *part,
                 size_t pps) {
#define GEMM(RA_, RB_)                                                                           \
  hipLaunchKernelGGL((cgemm_kernel<K, CA, CB, RA_, RB_, RE>), grid, dim3(256), 0, st, M, N, KD,   \
                     kchunk, A, B, Cp, cps, c0, c1, acc, part, pps)
  switch (blk) {
    case 1: GEMM(1, 1); break;
    case 2: GEMM(2, 1); break;
    case 3: GEMM(1, 2); break;
    default: GEMM(2, 2); break;
  }
#undef GEMM
}

// RE: real operands and result (K planes each); the conjugation flags are
// then ignored
template <int K, bool RE = false>
int cgemm(bool ca, bool cb, int M, int N, int KD, Op A, Op B, double *Cp, size_t cps, long c0,
          long c1, bool accumulate, hipStream_t st, double *part = nullptr, size_t cap = 0,
          GemmCfg cfg = GemmCfg()) {
  if (M <= 0 || N <= 0) return 0;
  const int blk = cfg.block ? cfg.block : default_block<K>();
  const int TM = 16 * block_ra(blk), TN = 16 * block_rb(blk);
  const int S = part ? ksplit(M, N, KD, cap, TM, TN, cfg.target ? cfg.target : 1024) : 1;
  const int kchunk = ((KD + S - 1) / S + TKD - 1) / TKD * TKD;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, S);
  const int acc = accumulate ? 1 : 0;
  const size_t pps = cap;
  if (RE) launch_gemm<K, false, false, true>(blk, grid, st, M, N, KD, kchunk, A, B, Cp, cps, c0, c1, acc, part, pps);
  else if (!ca && !cb) launch_gemm<K, false, false, false>(blk, grid, st, M, N, KD, kchunk, A, B, Cp, cps, c0, c1, acc, part, pps);
  else if (ca && !cb) launch_gemm<K, true, false, false>(blk, grid, st, M, N, KD, kchunk, A, B, Cp, cps, c0, c1, acc, part, pps);
  else if (!ca && cb) launch_gemm<K, false, true, false>(blk, grid, st, M, N, KD, kchunk, A, B, Cp, cps, c0, c1, acc, part, pps);
  else launch_gemm<K, true, true, false>(blk, grid, st, M, N, KD, kchunk, A, B, Cp, cps, c0, c1, acc, part, pps);
  if (S > 1)
    hipLaunchKernelGGL((csum_kernel<K, RE>), dim3(phc::grid_for((size_t)M * N, 256)), dim3(256), 0, st,
                       M, N, S, part, pps, Cp, cps, c0, c1, acc);
  PHC_HIP(hipGetLastError());
  return 0;
}

// --------------------------------------------------------- panel kernels

// the butterfly partner's value of a complex number (wave::bfly per limb)
template <int M, int K> __device__ __forceinline__ Z<K> bfly_cz(const Z<K> &s) {
  Z<K> u;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    u.re.x[q] = wave::bfly<M>(s.re.x[q]);
    u.im.x[q] = wave::bfly<M>(s.im.x[q]);
  }
  return u;
}

// butterfly sum of a complex number over the 64 lanes of a wavefront (every
// lane active; the sum is the same in every lane)
template <int K> __device__ Z<K> wave_csum(Z<K> s) {
  s = C<K>::add(s, bfly_cz<1, K>(s));
  s = C<K>::add(s, bfly_cz<2, K>(s));
  s = C<K>::add(s, bfly_cz<4, K>(s));
  s = C<K>::add(s, bfly_cz<8, K>(s));
  s = C<K>::add(s, bfly_cz<16, K>(s));
  return C<K>::add(s, bfly_cz<32, K>(s));
}

// the butterfly partner's value of a real multiple-double (wave::bfly per limb)
template <int M, int K> __device__ __forceinline__ T<K> bfly_r(const T<K> &s) {
  T<K> u;
#pragma unroll
  for (int q = 0; q < K; ++q) u.x[q] = wave::bfly<M>(s.x[q]);
  return u;
}

// Out-of-line multiple-double operations for the Householder kernel, whose
// serial chain (a sum over the column, two square roots, three divisions)
// runs on a few lanes of one workgroup: one copy of each operation's code
// instead of inlined copies at every use (88 KB of straight-line code).
template <int K> __device__ __noinline__ T<K> nl_add(T<K> a, T<K> b) { return add(a, b); }
template <int K> __device__ __noinline__ T<K> nl_sub(T<K> a, T<K> b) { return sub(a, b); }
template <int K> __device__ __noinline__ T<K> nl_mul(T<K> a, T<K> b) { return mul(a, b); }
template <int K> __device__ __noinline__ T<K> nl_div(T<K> a, T<K> b) { return div(a, b); }
// md::sqrt<4> (a double-double root, then one Newton correction) with its
// product, sum and quotient as the calls above
template <int K> __device__ __noinline__ T<K> nl_sqrt(T<K> a) {
  if constexpr (K == 4) {
    const T<K> b = widen_half<4>(md::sqrt(lead_half<4>(a)));
    return mul_pwr2<4>(nl_div<K>(nl_add<K>(nl_mul<K>(b, b), a), b), 0.5);
  } else {
    return R<K>::sqrt(a);
  }
}
// (a + re^2) + im^2, and re^2 + im^2
template <int K> __device__ __noinline__ T<K> nl_addsq(T<K> a, Z<K> z) {
  return add(add(a, sqr(z.re)), sqr(z.im));
}
template <int K> __device__ __noinline__ T<K> nl_norm2(Z<K> z) {
  return add(sqr(z.re), sqr(z.im));
}
// lane src's value on every lane of the wavefront
template <int K> __device__ __forceinline__ T<K> lane_r(const T<K> &a, int src) {
  T<K> r;
#pragma unroll
  for (int q = 0; q < K; ++q) r.x[q] = __shfl(a.x[q], src);
  return r;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int HW = 1024;  // lanes of the Householder kernel

// Householder vector of column colidx (rows colidx..nrows-1), as
// CPU_cmplx2_factors_house (Matrices/dbl2_factorizations.cpp:441-533), on
// the column-major copy Pc of the panel (column c at Pc + c*m): beta goes to
// B[L], v (v_0 = 1, v_i = x_i / v0, or x_i when sigma = 0) to Y row L from
// position L, zeros to Y row L before position L.
//   every wavefront (nw = blockDim.x / 64 <= 16, as many as the rows need):
//     sigma = sum_{i>0} |x_i|^2, one row per lane and pass, a butterfly per
//     wavefront;
//   wavefront 0: the nw sums (butterfly); nrm = sqrt(|x0|^2 + sigma) on
//     lane 0 and r0 = |x0| on lane 1 in one lane-uniform root; mu = nrm /
//     r0; v0 = x0 - mu x0 (re on lane 0, im on lane 1); |v0|^2; the three
//     quotients beta = 2 |v0|^2 / (sigma + |v0|^2), re and im of
//     conj(v0) / |v0|^2 on lanes 0, 1, 2 in one lane-uniform division;
//   every lane: v_i = x_i * (1/v0) for its rows.
// Sums are in tree order (the reference's is sequential): results agree to
// a few units in the last limb.
// RE (real data, zero imaginary parts; CPU_dbl4_factors_house,
// Matrices/dbl4_factorizations.cpp:541-626): mu = sqrt(x0^2 + sigma),
// v0 = x0 - mu when the leading limb of x0 is <= 0, else -sigma/(x0 + mu),
// beta = 2 v0^2 / (sigma + v0^2) on lane 0; every lane: v_i = x_i / v0.
template <int K, bool RE>
__global__ void __launch_bounds__(HW)
house_kernel(const double *Pc, size_t pps, int m, int rowoff, int colidx, int nrows, int L,
             double *Y, size_t yps, int rowdim, double *Bv, int szt) {
  __shared__ T<K> part[HW / 64];
  __shared__ Z<K> sinv;
  __shared__ T<K> sv0;
  __shared__ int sflag;
  const int t = threadIdx.x, w = t >> 6, n = nrows - colidx;
  const int nw = blockDim.x >> 6, stride = blockDim.x;
  const double *col = Pc + (size_t)(colidx - rowoff) * m + colidx;  // x_i at col[i]
  auto x = [&](int i) { return C<K>::load(col, pps, i); };
  for (int i = t; i < L; i += stride) C<K>::store(Y, yps, (size_t)L * rowdim + i, C<K>::zero());
  const int ir = 1 + t;  // the row of lane t in the first pass
  Z<K> xr;
  T<K> s = R<K>::zero();
  if (ir < n) {
    xr = x(ir);
    s = nl_norm2<K>(xr);
  }
  for (int i = ir + stride; i < n; i += stride) s = nl_addsq<K>(s, x(i));
  s = nl_add<K>(s, bfly_r<1, K>(s));
  s = nl_add<K>(s, bfly_r<2, K>(s));
  s = nl_add<K>(s, bfly_r<4, K>(s));
  s = nl_add<K>(s, bfly_r<8, K>(s));
  s = nl_add<K>(s, bfly_r<16, K>(s));
  s = nl_add<K>(s, bfly_r<32, K>(s));
  if ((t & 63) == 0) part[w] = s;
  __syncthreads();
  if (w == 0) {
    const Z<K> x0 = x(0);
    T<K> u = (t < nw) ? part[t] : R<K>::zero();
    const T<K> sq0 = nl_norm2<K>(x0);
    u = nl_add<K>(u, bfly_r<1, K>(u));
    u = nl_add<K>(u, bfly_r<2, K>(u));
    u = nl_add<K>(u, bfly_r<4, K>(u));
    u = nl_add<K>(u, bfly_r<8, K>(u));
    const T<K> sigma = lane_r<K>(u, 0);
    bool szero = true;
#pragma unroll
    for (int q = 0; q < K; ++q) szero = szero && (sigma.x[q] == 0.0);
    if constexpr (RE) {
      if (t == 0) {
        T<K> beta = R<K>::zero(), v0 = R<K>::from(1.0);
        if (!szero) {
          const T<K> mu = nl_sqrt<K>(nl_add<K>(sqr(x0.re), sigma));
          if (R<K>::hi(x0.re) <= 0.0) v0 = nl_sub<K>(x0.re, mu);
          else v0 = neg(nl_div<K>(sigma, nl_add<K>(x0.re, mu)));
          const T<K> v0p2 = sqr(v0);
          beta = mul_d(nl_div<K>(v0p2, nl_add<K>(sigma, v0p2)), 2.0);
        }
        R<K>::store(Bv, szt, L, beta);
        sflag = szero ? 1 : 0;
        sv0 = v0;
      }
    } else {
    // lane 0: nrm = sqrt(|x0|^2 + sigma), lane 1: r0 = sqrt(|x0|^2)
    T<K> rt = sq0;
    if (t == 0) rt = nl_add<K>(sq0, sigma);
    if (t < 2) rt = nl_sqrt<K>(rt);
    const T<K> nrm = lane_r<K>(rt, 0), r0 = lane_r<K>(rt, 1);
    bool rz = true;
#pragma unroll
    for (int q = 0; q < K; ++q) rz = rz && (r0.x[q] == 0.0);
    const T<K> mu = rz ? nrm : nl_div<K>(nrm, r0);
    // lane 0: v0.re, lane 1: v0.im
    T<K> vc, vsq;
    if (t < 2) {
      const T<K> xc = (t == 0) ? x0.re : x0.im;
      if (rz)
        vc = (t == 0) ? neg(mu) : R<K>::zero();
      else
        vc = nl_sub<K>(xc, nl_mul<K>(mu, xc));
      vsq = nl_mul<K>(vc, vc);
    }
    const T<K> v0re = lane_r<K>(vc, 0), v0im = lane_r<K>(vc, 1);
    const T<K> sqrv0 = nl_add<K>(lane_r<K>(vsq, 0), lane_r<K>(vsq, 1));
    T<K> num, den;
    if (t < 3) {
      if (t == 0) {
        num = sqrv0;
        den = nl_add<K>(sigma, sqrv0);
      } else {
        num = (t == 1) ? v0re : v0im;
        den = sqrv0;
      }
      T<K> q = nl_div<K>(num, den);
      if (t == 0) q = mul_d(q, 2.0);
      if (t == 2) q = neg(q);
      if (szero) q = R<K>::zero();
      if (t == 0) {
        R<K>::store(Bv, szt, L, q);
        sflag = szero ? 1 : 0;
      } else if (t == 1) {
        sinv.re = q;
      } else {
        sinv.im = q;
      }
    }
    }
  }
  __syncthreads();
  const bool flag = sflag != 0;
  auto vof = [&](const Z<K> &xi) {
    if (flag) return xi;
    Z<K> v;
    if constexpr (RE) {
      v.re = div(xi.re, sv0);
      v.im = R<K>::zero();
    } else {
      const Z<K> inv = sinv;
      v.re = sub(mul(xi.re, inv.re), mul(xi.im, inv.im));
      v.im = add(mul(xi.im, inv.re), mul(xi.re, inv.im));
    }
    return v;
  };
  double *yl = Y + (size_t)L * rowdim + L;  // v_i at yl[i]
  if (t == 0) C<K>::store(yl, yps, 0, C<K>::from(1.0, 0.0));
  if (ir < n) C<K>::store(yl, yps, ir, vof(xr));
  for (int i = ir + stride; i < n; i += stride) C<K>::store(yl, yps, i, vof(x(i)));
}

constexpr int CW = 256;  // lanes (rows) per workgroup of the column kernels

// The panel part of the reference's leftRupdate (Matrices/dbl4_baqr_host.cpp:
// 357) for the columns j = colidx + c of the column-major copy Pc, with the
// rows split over workgroups of CW lanes (one row per lane):
//   column_dot_kernel, workgroup (c, s): wpart[s][c] = sum over the rows of
//   split s of conj(R_ij) v_i (a butterfly per wavefront, then the four
//   wavefront sums);
//   column_upd_kernel, workgroup (c, s): w_j = beta_L (wpart[0][c] + ... +
//   wpart[ns-1][c]) and R_ij -= v_i conj(w_j) for the rows of split s.
template <int K>
__global__ void __launch_bounds__(CW)
column_dot_kernel(const double *Pc, size_t pps, int m, int rowoff, int colidx, int nrows,
                  const double *Y, size_t yps, int rowdim, int L, double *wpart, size_t wpps,
                  int szt) {
  __shared__ Z<K> part[CW / 64];
  const int t = threadIdx.x, c = blockIdx.x, i = blockIdx.y * CW + t, n = nrows - colidx;
  const double *col = Pc + (size_t)(colidx + c - rowoff) * m + colidx;  // R_ij at col[i]
  const double *yl = Y + (size_t)L * rowdim + L;                       // v_i at yl[i]
  Z<K> s = C<K>::zero();
  if (i < n) cmac<K, true, false>(s, C<K>::load(col, pps, i), C<K>::load(yl, yps, i));
  s = wave_csum<K>(s);
  if ((t & 63) == 0) part[t >> 6] = s;
  __syncthreads();
  if (t == 0)
    C<K>::store(wpart, wpps, (size_t)blockIdx.y * szt + c,
                C<K>::add(C<K>::add(part[0], part[1]), C<K>::add(part[2], part[3])));
}

template <int K>
__global__ void __launch_bounds__(CW)
column_upd_kernel(double *Pc, size_t pps, int m, int rowoff, int colidx, int nrows,
                  const double *Y, size_t yps, int rowdim, int L, const double *wpart,
                  size_t wpps, int ns, const double *Bv, int szt) {
  const int t = threadIdx.x, c = blockIdx.x, i = blockIdx.y * CW + t, n = nrows - colidx;
  if (i >= n) return;
  double *col = Pc + (size_t)(colidx + c - rowoff) * m + colidx;
  const double *yl = Y + (size_t)L * rowdim + L;
  Z<K> wj = C<K>::load(wpart, wpps, c);
  for (int q = 1; q < ns; ++q) wj = C<K>::add(wj, C<K>::load(wpart, wpps, (size_t)q * szt + c));
  const T<K> beta = R<K>::load(Bv, szt, L);
  wj.re = mul(beta, wj.re);
  wj.im = mul(beta, wj.im);
  const Z<K> v = C<K>::load(yl, yps, i);
  Z<K> r = C<K>::load(col, pps, i);
  // r -= v * conj(w)
  r.re = sub(r.re, add(mul(v.re, wj.re), mul(v.im, wj.im)));
  r.im = sub(r.im, sub(mul(v.im, wj.re), mul(v.re, wj.im)));
  C<K>::store(col, pps, i, r);
}

// the panel columns rowoff..rowoff+szt-1 of R (rows rowoff..m-1) to and from
// the column-major copy Pc
template <int K, bool RE>
__global__ void panel_copy_kernel(double *Rm, size_t rps, int ld, int m, int rowoff, int szt,
                                  double *Pc, size_t pps, int to_panel) {
  const size_t rows = (size_t)(m - rowoff), tot = rows * szt;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / rows), i = rowoff + (int)(e % rows);
    const size_t re = (size_t)i * ld + rowoff + c, pe = (size_t)c * m + i;
    if (to_panel)
      C<K>::store(Pc, pps, pe, Acc<K, RE>::load(Rm, rps, re));
    else
      Acc<K, RE>::store(Rm, rps, re, C<K>::load(Pc, pps, pe));
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
      s = wave_csum<K>(s);
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

// inverse of the diagonal tiles of U (dim x dim): one wavefront per column j
// of tile t, x = T^{-1} e_j by back substitution with the sums over k split
// across lanes; x kept in LDS, written to inv[t][szt][szt]
template <int K>
__global__ void __launch_bounds__(256)
tile_inverse_kernel(const double *U, size_t ups, int dim, int szt, int ntiles, double *inv,
                    size_t ips) {
  __shared__ double xs[4][2 * K][256];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int id = blockIdx.x * 4 + wv;
  if (id >= ntiles * szt) return;
  const int tile = id / szt, j = id % szt, off = tile * szt;
  double *x = &xs[wv][0][0];
  const size_t ib = (size_t)tile * szt * szt;
  for (int i = szt - 1; i >= 0; --i) {
    Z<K> v = C<K>::zero();
    if (i <= j) {
      Z<K> s = C<K>::zero();
      for (int k = i + 1 + l; k <= j; k += 64)
        s = C<K>::add(s, C<K>::mul(C<K>::load(U, ups, (size_t)(off + i) * dim + off + k),
                                    C<K>::load(x, 256, k)));
      s = wave_csum<K>(s);
      v = C<K>::sub(i == j ? C<K>::from(1.0, 0.0) : C<K>::zero(), s);
      v = C<K>::div(v, C<K>::load(U, ups, (size_t)(off + i) * dim + off + i));
    }
    if (l == 0) C<K>::store(x, 256, i, v);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  for (int i = l; i < szt; i += 64) C<K>::store(inv, ips, ib + (size_t)i * szt + j, C<K>::load(x, 256, i));
}

// rows r of y: y[r] = [y[r] -] sum_j M[r][j] v[j], one wavefront per row
// (M row r at m + r*ldm, v at v + j); sub = 1 subtracts from y
template <int K>
__global__ void __launch_bounds__(256)
rowdot_kernel(int rows, int len, const double *m, size_t mps, int ldm, const double *v,
              size_t vps, double *y, size_t yps, int sub) {
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wv;
  if (r >= rows) return;
  Z<K> s = C<K>::zero();
  for (int j = l; j < len; j += 64)
    s = C<K>::add(s, C<K>::mul(C<K>::load(m, mps, (size_t)r * ldm + j), C<K>::load(v, vps, j)));
  s = wave_csum<K>(s);
  if (l == 0) C<K>::store(y, yps, r, sub ? C<K>::sub(C<K>::load(y, yps, r), s) : s);
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

template <int K, bool RE>
__global__ void init_qr_kernel(int nrows, int ncols, const double *A, double *Q, double *Rm) {
  typedef Acc<K, RE> A_;
  const size_t nq = (size_t)nrows * nrows, nr = (size_t)nrows * ncols;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nq + nr;
       e += (size_t)gridDim.x * blockDim.x) {
    if (e < nq) {
      const size_t i = e / nrows, j = e % nrows;
      A_::store(Q, nq, e, C<K>::from(i == j ? 1.0 : 0.0, 0.0));
    } else {
      A_::store(Rm, nr, e - nq, A_::load(A, nr, e - nq));
    }
  }
}

// zeros into n doubles (the imaginary planes of a workspace the real QR
// reads as complex)
__global__ void zero_kernel(double *p, size_t n) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

}  // namespace

struct phc_baqr {
  int K, nrows, ncols, szt, nbt;
  GemmCfg cfg;  // GEMM launch configuration of the plan's products
  double *Y, *W, *G, *S, *T1, *T2, *Pc, *wpart, *B, *part;
  size_t yps, gps, t1ps, t2ps, pps, wpps, pcap;
  // the launch sequence of one factorization, captured for the last (A, Q, R)
  hipStream_t cap;
  hipGraphExec_t exec;
  const double *gA;
  double *gQ, *gR;
  int greal;  // the captured sequence is the real QR's
};

namespace {

// RE: A, Q, R are real (K planes); the panel kernels run on the complex
// workspace with zero imaginary parts, the products on the real planes
template <int K, bool RE>
int run_qr(phc_baqr *p, const double *A, double *Q, double *R, hipStream_t st) {
  const int m = p->nrows, n = p->ncols, szt = p->szt;
  const size_t qps = (size_t)m * m, rps = (size_t)m * n;
  hipLaunchKernelGGL((init_qr_kernel<K, RE>), dim3(phc::grid_for(qps + rps, 256)), dim3(256), 0, st, m,
                     n, A, Q, R);
  if (RE)  // the Gram matrix is read as complex by the S factor
    hipLaunchKernelGGL(zero_kernel, dim3(phc::grid_for((size_t)K * p->gps, 256)), dim3(256), 0, st,
                       p->G + (size_t)K * p->gps, (size_t)K * p->gps);
  PHC_HIP(hipGetLastError());
  for (int k = 0; k < p->nbt; ++k) {
    const int rowoff = k * szt, rowdim = m - rowoff, endcol = (k + 1) * szt;
    hipLaunchKernelGGL((panel_copy_kernel<K, RE>), dim3(phc::grid_for((size_t)rowdim * szt, 256)),
                       dim3(256), 0, st, R, rps, n, m, rowoff, szt, p->Pc, p->pps, 1);
    for (int L = 0; L < szt; ++L) {
      const int colidx = rowoff + L, ncol = endcol - colidx;
      const int nr = m - colidx, ns = (nr + CW - 1) / CW;
      const int hwv = (nr - 1 + 63) / 64, hw = 64 * (hwv < 1 ? 1 : hwv < HW / 64 ? hwv : HW / 64);
      hipLaunchKernelGGL((house_kernel<K, RE>), dim3(1), dim3(hw), 0, st, p->Pc, p->pps, m, rowoff,
                         colidx, m, L, p->Y, p->yps, rowdim, p->B, szt);
      hipLaunchKernelGGL(column_dot_kernel<K>, dim3(ncol, ns), dim3(CW), 0, st, p->Pc, p->pps, m,
                         rowoff, colidx, m, p->Y, p->yps, rowdim, L, p->wpart, p->wpps, szt);
      hipLaunchKernelGGL(column_upd_kernel<K>, dim3(ncol, ns), dim3(CW), 0, st, p->Pc, p->pps, m,
                         rowoff, colidx, m, p->Y, p->yps, rowdim, L, p->wpart, p->wpps, ns, p->B,
                         szt);
    }
    hipLaunchKernelGGL((panel_copy_kernel<K, RE>), dim3(phc::grid_for((size_t)rowdim * szt, 256)),
                       dim3(256), 0, st, R, rps, n, m, rowoff, szt, p->Pc, p->pps, 0);
    PHC_HIP(hipGetLastError());
    // Y, W stored as szt rows of length rowdim (row stride rowdim)
    const Op Yt{p->Y, p->yps, 1, rowdim};  // Yhat(i, k) = Y[k][i]
    const Op Yc{p->Y, p->yps, rowdim, 1};  // (k, i) -> Y[k][i]
    // G = Yhat^H Yhat
    if (cgemm<K, RE>(true, false, szt, szt, rowdim, Yc, Yt, p->G, p->gps, szt, 1, false, st, p->part,
                 p->pcap, p->cfg)) return -1;
    hipLaunchKernelGGL(s_factor_kernel<K>, dim3(szt), dim3(64), 0, st, p->G, p->gps, p->B, p->S,
                       p->gps, szt);
    PHC_HIP(hipGetLastError());
    // W = Yhat S, stored like Y
    if (cgemm<K, RE>(false, false, rowdim, szt, szt, Yt, Op{p->S, p->gps, szt, 1}, p->W, p->yps, 1,
                 rowdim, false, st, p->part, p->pcap, p->cfg)) return -1;
    const Op Wt{p->W, p->yps, 1, rowdim};  // What(i, k) = W[k][i]
    if (k < p->nbt - 1) {
      const int coloff = endcol, coldim = n - coloff;
      // T1 = What^H C with C = R[rowoff.., coloff..];  C += Yhat T1
      const Op Cm{R + (size_t)rowoff * n + coloff, rps, n, 1};
      if (cgemm<K, RE>(true, false, szt, coldim, rowdim, Op{p->W, p->yps, rowdim, 1}, Cm, p->T1,
                   p->t1ps, coldim, 1, false, st, p->part, p->pcap, p->cfg)) return -1;
      if (cgemm<K, RE>(false, false, rowdim, coldim, szt, Yt, Op{p->T1, p->t1ps, coldim, 1},
                   R + (size_t)rowoff * n + coloff, rps, n, 1, true, st, p->part, p->pcap, p->cfg))
        return -1;
    }
    // T2 = Q[:, rowoff..] What;  Q[:, rowoff..] += T2 Yhat^H
    const Op Qm{Q + rowoff, qps, m, 1};
    if (cgemm<K, RE>(false, false, m, szt, rowdim, Qm, Wt, p->T2, p->t2ps, szt, 1, false, st, p->part,
                 p->pcap, p->cfg)) return -1;
    if (cgemm<K, RE>(false, true, m, rowdim, szt, Op{p->T2, p->t2ps, szt, 1}, Yc, Q + rowoff, qps, m,
                 1, true, st, p->part, p->pcap, p->cfg)) return -1;
  }
  return 0;
}

template <int K>
int run_tiled(int dim, int szt, int nbt, double *U, double *b, double *x, double *inv,
              hipStream_t st) {
  const size_t ups = (size_t)dim * dim, ips = (size_t)nbt * szt * szt;
  hipLaunchKernelGGL(tile_inverse_kernel<K>, dim3((nbt * szt + 3) / 4), dim3(256), 0, st, U, ups,
                     dim, szt, nbt, inv, ips);
  // x_last = invT_last b_last; for k = nbt-1..1: b_L -= U_{L,k} x_k (L < k),
  // x_{k-1} = invT_{k-1} b_{k-1}
  const int last = nbt - 1;
  hipLaunchKernelGGL(rowdot_kernel<K>, dim3((szt + 3) / 4), dim3(256), 0, st, szt, szt,
                     inv + (size_t)last * szt * szt, ips, szt, b + (size_t)last * szt, (size_t)dim,
                     x + (size_t)last * szt, (size_t)dim, 0);
  for (int k = nbt - 1; k > 0; --k) {
    hipLaunchKernelGGL(rowdot_kernel<K>, dim3((k * szt + 3) / 4), dim3(256), 0, st, k * szt, szt,
                       U + (size_t)k * szt, ups, dim, x + (size_t)k * szt, (size_t)dim, b,
                       (size_t)dim, 1);
    hipLaunchKernelGGL(rowdot_kernel<K>, dim3((szt + 3) / 4), dim3(256), 0, st, szt, szt,
                       inv + (size_t)(k - 1) * szt * szt, ips, szt, b + (size_t)(k - 1) * szt,
                       (size_t)dim, x + (size_t)(k - 1) * szt, (size_t)dim, 0);
  }
  hipLaunchKernelGGL(copy_inverse_kernel<K>, dim3(phc::grid_for(ips, 256)), dim3(256), 0, st, U,
                     ups, dim, szt, nbt, inv, ips);
  PHC_HIP(hipGetLastError());
  return 0;
}

// Device scratch of the allocation-free tiled solver: the inverses of the
// diagonal tiles, 2*K*nbt*szt*szt doubles, at most kTiledScratch (8 MiB:
// dim 1024 in tiles of 128 in complex quad-double).  Every GPU holds its own
// copy for the life of the process; calls on one GPU share it.
constexpr size_t kTiledScratch = size_t(1) << 20;
static __device__ __align__(256) double scratch[kTiledScratch];

// The device address of `scratch` on the current device, looked up once
// per device (the host-side name is not a device address).
int tiled_scratch(double **out) {
  static std::mutex mu;
  static std::vector<double *> ptrs;
  int dev = 0;
  PHC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
  if (!ptrs[dev]) {
    void *p = nullptr;
    PHC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PHC_REQUIRE(p != nullptr, "phc_cmplx_upper_tiled_solver_noalloc: no scratch address");
    ptrs[dev] = (double *)p;
  }
  *out = ptrs[dev];
  return 0;
}

}  // namespace

extern "C" {

int phc_baqr_create(phc_baqr **out, int K, int nrows, int ncols, int szt) {
  PHC_REQUIRE(out, "phc_baqr_create: null plan pointer");
  PHC_REQUIRE(K == 2 || K == 4 || K == 8, "phc_baqr_create: K must be 2, 4 or 8");
  PHC_REQUIRE(szt >= 1 && szt <= 128 && ncols >= szt && ncols % szt == 0,
              "phc_baqr_create: ncols must be a multiple of szt (1..128)");
  PHC_REQUIRE(nrows >= ncols, "phc_baqr_create: nrows must be at least ncols");
  phc_baqr *p = new phc_baqr();
  p->K = K; p->nrows = nrows; p->ncols = ncols; p->szt = szt; p->nbt = ncols / szt;
  p->yps = (size_t)szt * nrows;
  p->gps = (size_t)szt * szt;
  p->t1ps = (size_t)szt * ncols;
  p->t2ps = (size_t)nrows * szt;
  p->pps = (size_t)szt * nrows;
  p->wpps = (size_t)szt * ((nrows + CW - 1) / CW);
  // split-k partial sums: up to 8 slices of the largest thin product, and
  // the Q^H b products (nrows x 1 in up to 32 slices)
  p->pcap = (size_t)nrows * szt * 8;
  if ((size_t)nrows * 32 > p->pcap) p->pcap = (size_t)nrows * 32;
  const size_t z = sizeof(double) * 2 * K;
  double **bufs[] = {&p->Y,  &p->W,     &p->G, &p->S,   &p->T1,
                     &p->T2, &p->Pc, &p->wpart, &p->B, &p->part};
  const size_t sizes[] = {z * p->yps, z * p->yps, z * p->gps, z * p->gps, z * p->t1ps,
                          z * p->t2ps, z * p->pps, z * p->wpps, sizeof(double) * K * szt,
                          z * p->pcap};
  const int nb = 10;
  for (int i = 0; i < nb; ++i) *bufs[i] = nullptr;
  p->exec = nullptr;
  p->cap = nullptr;
  p->gA = nullptr;
  p->gQ = p->gR = nullptr;
  for (int i = 0; i < nb; ++i) {
    if (hipMalloc(bufs[i], sizes[i]) != hipSuccess) {
      for (int j = 0; j < i; ++j) hipFree(*bufs[j]);
      delete p;
      phc::set_error("phc_baqr_create: hipMalloc failed");
      return -1;
    }
  }
  if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess) {
    for (int j = 0; j < nb; ++j) hipFree(*bufs[j]);
    delete p;
    phc::set_error("phc_baqr_create: hipStreamCreate failed");
    return -1;
  }
  *out = p;
  return 0;
}

int phc_baqr_destroy(phc_baqr *p) {
  if (!p) return 0;
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->cap) hipStreamDestroy(p->cap);
  double *bufs[] = {p->Y, p->W, p->G, p->S, p->T1, p->T2, p->Pc, p->wpart, p->B, p->part};
  for (double *b : bufs) hipFree(b);
  delete p;
  return 0;
}

int phc_baqr_set_tuning(phc_baqr *p, int gemm_block, int split_target) {
  PHC_REQUIRE(p, "phc_baqr_set_tuning: null plan");
  PHC_REQUIRE(gemm_block >= 0 && gemm_block <= 4, "phc_baqr_set_tuning: gemm_block must be 0..4");
  PHC_REQUIRE(split_target >= 0 && split_target <= (1 << 20),
              "phc_baqr_set_tuning: split_target must be 0..2^20");
  p->cfg.block = gemm_block;
  p->cfg.target = split_target;
  // the captured launch sequence holds the old configuration
  if (p->exec) {
    PHC_HIP(hipGraphExecDestroy(p->exec));
    p->exec = nullptr;
  }
  return 0;
}

int phc_baqr_get_tuning(const phc_baqr *p, int *gemm_block, int *split_target) {
  PHC_REQUIRE(p && gemm_block && split_target, "phc_baqr_get_tuning: null argument");
  *gemm_block = p->cfg.block ? p->cfg.block
                : (p->K >= 4 ? default_block<4>() : default_block<2>());
  *split_target = p->cfg.target ? p->cfg.target : 1024;
  return 0;
}

// The ~4 launches per column are captured once into a HIP graph (per
// (A, Q, R) pointers) and replayed: the panel kernels are short, so launch
// gaps would otherwise dominate.
namespace {

int blocked_houseqr(phc_baqr *p, const double *A, double *Q, double *R, hipStream_t st, int real) {
  if (!p->exec || p->gA != A || p->gQ != Q || p->gR != R || p->greal != real) {
    if (p->exec) {
      PHC_HIP(hipGraphExecDestroy(p->exec));
      p->exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    PHC_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
    int rc;
    if (real)
      rc = (p->K == 2) ? run_qr<2, true>(p, A, Q, R, p->cap)
           : (p->K == 4) ? run_qr<4, true>(p, A, Q, R, p->cap)
                         : run_qr<8, true>(p, A, Q, R, p->cap);
    else
      rc = (p->K == 2) ? run_qr<2, false>(p, A, Q, R, p->cap)
           : (p->K == 4) ? run_qr<4, false>(p, A, Q, R, p->cap)
                         : run_qr<8, false>(p, A, Q, R, p->cap);
    const hipError_t e = hipStreamEndCapture(p->cap, &graph);
    if (rc) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    PHC_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    PHC_HIP(ei);
    p->gA = A;
    p->gQ = Q;
    p->gR = R;
    p->greal = real;
  }
  PHC_HIP(hipGraphLaunch(p->exec, st));
  return 0;
}

}  // namespace

int phc_cmplx_blocked_houseqr(phc_baqr *p, const double *A, double *Q, double *R, void *stream) {
  PHC_REQUIRE(p && A && Q && R, "phc_cmplx_blocked_houseqr: null argument");
  return blocked_houseqr(p, A, Q, R, (hipStream_t)stream, 0);
}

int phc_real_blocked_houseqr(phc_baqr *p, const double *A, double *Q, double *R, void *stream) {
  PHC_REQUIRE(p && A && Q && R, "phc_real_blocked_houseqr: null argument");
  return blocked_houseqr(p, A, Q, R, (hipStream_t)stream, 1);
}

int phc_real_qtb(phc_baqr *p, const double *Q, const double *b, double *y, void *stream) {
  PHC_REQUIRE(p && Q && b && y, "phc_real_qtb: null argument");
  const int n = p->nrows;
  hipStream_t st = (hipStream_t)stream;
  const Op qa{Q, (size_t)n * n, 1, n}, bb{b, (size_t)n, 1, 0};
  if (p->K == 8)
    return cgemm<8, true>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st, p->part,
                          p->pcap, p->cfg);
  return p->K == 2 ? cgemm<2, true>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st,
                                    p->part, p->pcap, p->cfg)
                   : cgemm<4, true>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st,
                                    p->part, p->pcap, p->cfg);
}

int phc_cmplx_qhb(phc_baqr *p, const double *Q, const double *b, double *y, void *stream) {
  PHC_REQUIRE(p && Q && b && y, "phc_cmplx_qhb: null argument");
  const int n = p->nrows;
  hipStream_t st = (hipStream_t)stream;
  const Op qa{Q, (size_t)n * n, 1, n}, bb{b, (size_t)n, 1, 0};
  if (p->K == 8)
    return cgemm<8>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st, p->part, p->pcap,
                    p->cfg);
  return p->K == 2 ? cgemm<2>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st, p->part,
                              p->pcap, p->cfg)
                   : cgemm<4>(true, false, n, 1, n, qa, bb, y, (size_t)n, 1, 0, false, st, p->part,
                              p->pcap, p->cfg);
}

int phc_cmplx_gemm(int K, int conjA, int conjB, int M, int N, int KD, const double *A,
                   size_t aps, long a0, long a1, const double *B, size_t bps, long b0, long b1,
                   double *Cm, size_t cps, long c0, long c1, int accumulate, void *stream) {
  PHC_REQUIRE(K == 2 || K == 4 || K == 8, "phc_cmplx_gemm: K must be 2, 4 or 8");
  PHC_REQUIRE(M >= 0 && N >= 0 && KD >= 0, "phc_cmplx_gemm: negative size");
  PHC_REQUIRE(A && B && Cm, "phc_cmplx_gemm: null array");
  hipStream_t st = (hipStream_t)stream;
  const Op opa{A, aps, a0, a1}, opb{B, bps, b0, b1};
  if (K == 2) return cgemm<2>(conjA, conjB, M, N, KD, opa, opb, Cm, cps, c0, c1, accumulate, st);
  if (K == 4) return cgemm<4>(conjA, conjB, M, N, KD, opa, opb, Cm, cps, c0, c1, accumulate, st);
  return cgemm<8>(conjA, conjB, M, N, KD, opa, opb, Cm, cps, c0, c1, accumulate, st);
}

int phc_cmplx_upper_tiled_solver(int K, int dim, int szt, int nbt, double *U, double *b,
                                 double *x, void *stream) {
  PHC_REQUIRE(K == 2 || K == 4 || K == 8, "phc_cmplx_upper_tiled_solver: K must be 2, 4 or 8");
  PHC_REQUIRE(szt >= 1 && szt <= 256 && nbt >= 1 && dim == szt * nbt,
              "phc_cmplx_upper_tiled_solver: dim must equal szt*nbt, szt <= 256");
  PHC_REQUIRE(U && b && x, "phc_cmplx_upper_tiled_solver: null array");
  hipStream_t st = (hipStream_t)stream;
  double *inv = nullptr;
  const size_t ips = (size_t)nbt * szt * szt;
  PHC_HIP(hipMalloc(&inv, sizeof(double) * 2 * K * ips));
  const int rc = K == 2   ? run_tiled<2>(dim, szt, nbt, U, b, x, inv, st)
                 : K == 4 ? run_tiled<4>(dim, szt, nbt, U, b, x, inv, st)
                          : run_tiled<8>(dim, szt, nbt, U, b, x, inv, st);
  const hipError_t e2 = hipStreamSynchronize(st);
  hipFree(inv);
  if (rc) return rc;
  PHC_HIP(e2);
  return 0;
}

// The same solve with the tile inverses in the library's static device
// scratch: no allocation and no synchronization (the work is queued on the
// stream).  Calls that overlap on one GPU share the scratch and race.
int phc_cmplx_upper_tiled_solver_noalloc(int K, int dim, int szt, int nbt, double *U, double *b,
                                         double *x, void *stream) {
  PHC_REQUIRE(K == 2 || K == 4, "phc_cmplx_upper_tiled_solver_noalloc: K must be 2 or 4");
  PHC_REQUIRE(szt >= 1 && szt <= 256 && nbt >= 1 && dim == szt * nbt,
              "phc_cmplx_upper_tiled_solver_noalloc: dim must equal szt*nbt, szt <= 256");
  PHC_REQUIRE(U && b && x, "phc_cmplx_upper_tiled_solver_noalloc: null array");
  PHC_REQUIRE((size_t)2 * K * dim * szt <= kTiledScratch,
              "phc_cmplx_upper_tiled_solver_noalloc: the tile inverses (2*K*dim*szt doubles) "
              "exceed the 1048576-double (8 MiB) scratch");
  double *inv = nullptr;
  const int rs = tiled_scratch(&inv);
  if (rs) return rs;
  hipStream_t st = (hipStream_t)stream;
  return K == 2 ? run_tiled<2>(dim, szt, nbt, U, b, x, inv, st)
                : run_tiled<4>(dim, szt, nbt, U, b, x, inv, st);
}

}  // extern "C"
