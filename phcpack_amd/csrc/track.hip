// Batched predictor-corrector path tracking of homotopies on gfx950, with
// endpoint Newton refinement and residuals in double, double-double,
// quad-double and octo-double.
//
// Reference behaviour (restated, not translated):
//   step loop        Track_One_Path / Predictor_Corrector_Loop
//                    (Deformations/Curves/predictor_corrector_trackers.adb,
//                     predictor_corrector_loops.adb:16-83)
//   step size        LU_Prediction: pole step pbeta*r with the Fabry radius
//                    (standard_predictor_convolutions.adb:1142-1185;
//                     Math_Lib/Circuits/convergence_radius_estimates.adb:138-163,354-390),
//                    Bounded_Update (three_way_minima.adb:96-109), and the
//                    predictor feedback that halves the step while the
//                    predicted residual exceeds alpha (Predictor_Feedback)
//   predictor        the Taylor polynomial of the series, or its [L/1] or
//                    [L/2] Pade approximant per component (Pade_Vector at
//                    standard_predictor_convolutions.adb:312; Denominator_System
//                    and Pade in standard_rational_approximations.adb)
//   corrector        Newton steps, step halved on failure down to minstep
//   batching         GPU_Path_mult (Path/Path_GPU/path_gpu_mult.cu:84-308)
//                    compacts live paths on the host after every step; here
//                    every wavefront pulls the next path from an atomic queue
//                    when its path ends, so nothing returns to the host until
//                    all paths are done.
//
// Power series of the solution curve.  At the base point (x0, t0) the
// Taylor coefficients x_k of x(t0+s) are computed order by order: with x_k
// set to zero, coefficient k of H(x(s), t0+s) is evaluated through the
// monomial product DAG; it is affine in x_k with the Jacobian A0 = H_x(x0,t0)
// as linear part, so x_k = -A0^{-1} [H]_k.  This is the lower block-triangular
// Toeplitz solve of power-series Newton (Newton/dbl2_bals_host.cpp:228-335,
// x_k = A0^{-1}(b_k - sum_j A_j x_{k-j})) with the tail sum formed by the
// convolutions themselves.  A0 is inverted once per step (Gauss-Jordan with
// partial pivoting), so every order costs one matrix-vector product.
//
// Layout.  A workgroup of W wavefronts stages the compiled program (node
// metadata, terms, coefficients) into LDS once; then each wavefront tracks
// its own paths, one at a time, with all per-path state in its own LDS
// region.  Wavefronts never wait for each other: lanes of one wavefront
// exchange data through LDS with wavefront-scope ordering only, and sums
// across the G lanes that share an equation use butterfly shuffles.  Every
// control decision is computed identically by all lanes of the wavefront.
#include <hip/hip_runtime.h>
#include <vector>
#include "common.hpp"
#include "mdtypes.hpp"
#include "wave.hpp"
#include "../../include/phcmd.h"

using namespace md;

// pointers into LDS: address space 3, so that helpers the compiler does not
// inline still reach LDS with DS instructions (a generic pointer would make
// them FLAT accesses).  The host pass of hipcc only parses the device code
// and has no such address space, so the qualifier is given to the device
// pass alone.
#if defined(__HIP_DEVICE_COMPILE__)
#define LDS __attribute__((address_space(3)))
#else
#define LDS
#endif

namespace {

enum { ST_RUNNING = 0, ST_SUCCESS = 1, ST_MINSTEP = 2, ST_MAXSTEPS = 3, ST_SINGULAR = 4,
       ST_DIVERGED = 5 };

constexpr int JT_TERM_BITS = 16, JT_PAIR_BITS = 12;
// A gfx950 workgroup may allocate the whole 160 KiB of its compute unit's
// LDS (dynamic allocations of 48, 96 and 150 KiB were written and read back
// intact on an MI355X); the device attribute caps it further if lower.
constexpr size_t LDS_LIMIT = 160 * 1024;
constexpr size_t LDS_PER_CU = 160 * 1024;
// wavefronts per workgroup: FP64 tracking may run 12 (a kernel variant
// compiled for three waves per SIMD, at most 168 VGPRs); the others 8
constexpr int TRACK_MAX_WAVES = 12, REFINE_MAX_WAVES = 8;

// device copy of the program (global memory)
struct DevProgram {
  int n, D, ncomp, wdeg, G, nprod, nlev, nslots, nfactor, npair, nterm, njpair, njterm;
  const int *nodes, *lev_ptr, *pairs, *seg_ptr, *term_pair, *jpairs, *jent_ptr, *jterms;
  const double *term_coef;  // [2][nterm]
  const double *weights;    // [2][8][3*ncomp]
  const double *hsig;       // [n]: Hessian singular values, -1: evaluated every step
  const int *hent_ptr, *hterms, *hpairs;
  int hdyn;                 // 1 when some equation's Hessian is evaluated every step
};

struct Params {
  int maxsteps, corsteps, endsteps, fabry_offset;
  double maxstep, minstep, pbeta, alpha, tolcor, tolend, divergence;
  int pnum, pden;  // predictor: numerator degree (>= 1), denominator degree (0, 1 or 2)
  double cbeta;    // curvature step factor
};

// step-size controller (device view of phc_controller)
struct Ctl {
  int mode, ncand, cap;
  double mult[PHC_CTL_MAXCAND];
  const float *w, *mu, *sd;
  float thresh;
  unsigned seed;
  float *log_x;
  int *log_c, *log_y, *log_n;
  double pcap;  // step cap as a fraction of the Fabry radius (0: none)
  // series coefficients computed before the choice (d0) and for the step of
  // candidate c (dc[c] >= d0), at most the program's deg + 1
  int d0;
  int dc[PHC_CTL_MAXCAND];
  double retry;  // mode 1: paths stopped at t < retry are tracked again
};

// log2 of x > 0 from the binary exponent and the series
// ln((1+z)/(1-z)) = 2(z + z^3/3 + z^5/5 + z^7/7), z = (m-1)/(m+1), m in [1,2);
// only IEEE double operations, so host and device agree bit for bit
__host__ __device__ inline double lg2(double x) {
  if (!(x > 0.0)) return -1100.0;
  if (x == INFINITY) return 1100.0;
  int e = 0;
  const double m = 2.0 * frexp(x, &e);  // x = m * 2^(e-1), m in [1,2)
  const double z = (m - 1.0) / (m + 1.0), z2 = z * z;
  const double s = z * (1.0 + z2 * (1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (1.0 / 7.0))));
  return (double)(e - 1) + s * 2.8853900817779268;  // 2/ln 2
}

__host__ __device__ inline unsigned mix_hash(unsigned seed, unsigned path, unsigned step) {
  unsigned h = seed ^ (path * 0x9E3779B1u) ^ (step * 0x85EBCA77u);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

__host__ __device__ inline size_t al16(size_t b) { return (b + 15) & ~(size_t)15; }

// byte offsets of the program sections in LDS
struct ProgLayout {
  size_t nodes, lev, pairs, seg, tpair, tcoef, jpairs, jent, jterms, total;
  __host__ __device__ explicit ProgLayout(const DevProgram &P) {
    size_t o = 0;
    nodes = o; o = al16(o + sizeof(int) * 8 * P.nprod);
    lev = o; o = al16(o + sizeof(int) * (P.nlev + 1));
    pairs = o; o = al16(o + sizeof(int) * 4 * P.npair);
    seg = o; o = al16(o + sizeof(int) * (P.n * P.G + 1));
    tpair = o; o = al16(o + sizeof(int) * P.nterm);
    tcoef = o; o = al16(o + sizeof(double) * 2 * P.nterm);
    jpairs = o; o = al16(o + sizeof(int) * 2 * P.njpair);
    jent = o; o = al16(o + sizeof(int) * (P.n * P.n + 1));
    jterms = o; o = al16(o + sizeof(int) * P.njterm);
    total = o;
  }
};

// scratch of the step rule, in doubles, aliasing the weighted-pair arrays
// u and uj (free outside the residual and the Jacobian): the moduli of the
// nodes for the mixed residual, or the matrices of the singular values
// (Hessian H and Gram power P, and Q where n > 16 keeps no entries in
// registers)
__host__ __device__ inline size_t step_scratch_doubles(const DevProgram &P) {
  const size_t nn = (size_t)P.n * P.n, q = P.n > 16 ? 1 : 0;
  const size_t sv = 2 * nn * ((P.hdyn ? 2 : 1) + q);
  const size_t rn = (size_t)(P.n + 1 + P.nprod);
  return sv > rn ? sv : rn;
}

// per-wavefront path region (bytes)
template <int K> struct PathLayout {
  size_t slots, u, uj, A, rhs, x0, col, W, dnode, perm, hid, total;
  __host__ __device__ explicit PathLayout(const DevProgram &P) {
    const size_t z = sizeof(cplx<K>);
    size_t o = 0;
    slots = o; o += z * P.nslots;
    u = o;
    {
      // u, uj contiguous and together at least the step rule's scratch
      const size_t need = al16(sizeof(double) * step_scratch_doubles(P));
      const size_t have = z * (P.npair + P.njpair);
      o += z * P.npair + (need > have ? need - have : 0);
    }
    uj = o; o += z * P.njpair;
    A = o; o += z * P.n * P.n;
    rhs = o; o += z * P.n;
    x0 = o; o += z * P.n;
    col = o; o += z * P.n;
    W = o; o += z * 3 * P.ncomp;
    dnode = o; o += z * (P.nfactor > 0 ? P.nfactor : 1);
    perm = o; o = al16(o + sizeof(int) * P.n);
    hid = o; o += sizeof(float) * PHC_CTL_HIDDEN;
    total = al16(o);
  }
};

// the program as seen from LDS
struct LProg {
  int n, D, ncomp, wdeg, G, nprod, nlev, npair, nterm, njpair;
  const LDS int *nodes, *lev, *pairs, *seg, *tpair, *jpairs, *jent, *jterms;
  const LDS double2 *tcoef;
  const double *weights;  // global
  const double *hsig;     // global
  const int *hent, *hterms, *hpairs;  // global
};

template <int K> struct Path {
  LDS cplx<K> *slots, *u, *uj, *A, *rhs, *x0, *col, *W, *dnode;
  LDS int *perm;
  LDS float *hid;
};

// Complex operations of the tracker: FP64 tracking (K = 1) forms products
// with fused multiply-adds; K > 1 keeps the multiple-double order of C<K>.
template <int K> struct TC : C<K> {};
template <> struct TC<1> : C<1> {
  typedef cplx<1> Z;
  __device__ static Z mul(const Z &x, const Z &y) {
    Z z;
    z.re = __builtin_fma(x.re, y.re, -(x.im * y.im));
    z.im = __builtin_fma(x.re, y.im, x.im * y.re);
    return z;
  }
  __device__ static void mac(Z &z, const Z &x, const Z &y) {
    z.re = __builtin_fma(x.re, y.re, z.re);
    z.re = __builtin_fma(-x.im, y.im, z.re);
    z.im = __builtin_fma(x.im, y.re, z.im);
    z.im = __builtin_fma(x.re, y.im, z.im);
  }
  __device__ static void msub(Z &z, const Z &x, const Z &y) {
    z.re = __builtin_fma(-x.re, y.re, z.re);
    z.re = __builtin_fma(x.im, y.im, z.re);
    z.im = __builtin_fma(-x.im, y.re, z.im);
    z.im = __builtin_fma(-x.re, y.im, z.im);
  }
};

__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

template <int K> __device__ __forceinline__ double absz(const cplx<K> &z) {
  return fabs(R<K>::hi(z.re)) + fabs(R<K>::hi(z.im));
}

using wave::bfly;
using wave::bfly_b32;

template <int M, int K> __device__ __forceinline__ cplx<K> bfly_z(const cplx<K> &z) {
  cplx<K> r;
  const double *a = reinterpret_cast<const double *>(&z);
  double *b = reinterpret_cast<double *>(&r);
#pragma unroll
  for (int i = 0; i < 2 * K; ++i) b[i] = bfly<M>(a[i]);
  return r;
}

// sum over the aligned group of G lanes (butterfly v + partner per stage;
// identical in every lane of the group)
template <int K> __device__ __forceinline__ cplx<K> group_sum(cplx<K> v, int G) {
  if (G > 1) v = C<K>::add(v, bfly_z<1, K>(v));
  if (G > 2) v = C<K>::add(v, bfly_z<2, K>(v));
  if (G > 4) v = C<K>::add(v, bfly_z<4, K>(v));
  if (G > 8) v = C<K>::add(v, bfly_z<8, K>(v));
  if (G > 16) v = C<K>::add(v, bfly_z<16, K>(v));
  if (G > 32) v = C<K>::add(v, bfly_z<32, K>(v));
  return v;
}

__device__ __forceinline__ double wave_max(double v) { return wave::max64(v); }

__device__ __forceinline__ double wave_min(double v) { return wave::min64(v); }

using wave::argmax_stage;

template <int K> __device__ __forceinline__ cplx<K> coef(const LProg &P, int t) {
  const double2 c = P.tcoef[t];
  return C<K>::from(c.x, c.y);
}

template <int K> __device__ __forceinline__ cplx<K> load_weight(const LProg &P, int j) {
  cplx<K> z;
  z.re = R<K>::load(P.weights, 3 * P.ncomp, j);
  z.im = R<K>::load(P.weights + 8 * 3 * (size_t)P.ncomp, 3 * P.ncomp, j);
  return z;
}

// Taylor coefficients W[r][l] at t of w_r(t) = om0 + om1 t + om2 t^2
template <int K> __device__ void set_weights(const LProg &__restrict__ P, Path<K> &__restrict__ w, double t) {
  typedef TC<K> Cx;
  const int l = lane_id();
  if (l < P.ncomp) {
    cplx<K> o0 = load_weight<K>(P, 3 * l), o1 = load_weight<K>(P, 3 * l + 1),
            o2 = load_weight<K>(P, 3 * l + 2);
    cplx<K> o2t = Cx::scale_d(o2, t);
    w.W[3 * l] = Cx::add(o0, Cx::scale_d(Cx::add(o1, o2t), t));
    w.W[3 * l + 1] = Cx::add(o1, Cx::scale_d(o2t, 2.0));
    w.W[3 * l + 2] = o2;
  }
  wsync();
}

// node_k = sum_q left_q * right_{k-q} for every product node, level by level
template <int K> __device__ void product_pass(const LProg &__restrict__ P, Path<K> &__restrict__ w, int k) {
  typedef TC<K> Cx;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev[lv - 1] + lane_id(); i < P.lev[lv]; i += 64) {
      const LDS int *nd = P.nodes + 8 * i;
      const int aoff = nd[0], boff = nd[2], soff = nd[3], smask = nd[4];
      // terms q = 0..k alternate between two accumulators (even, odd q)
      cplx<K> a0 = Cx::mul(w.slots[aoff], w.slots[boff + k]);
      if (k >= 1) {
        cplx<K> a1 = Cx::mul(w.slots[aoff + 1], w.slots[boff + k - 1]);
        for (int q = 2; q <= k; q += 2) {
          Cx::mac(a0, w.slots[aoff + q], w.slots[boff + k - q]);
          if (q + 1 <= k) Cx::mac(a1, w.slots[aoff + q + 1], w.slots[boff + k - q - 1]);
        }
        a0 = Cx::add(a0, a1);
      }
      w.slots[soff + (k & smask)] = a0;
    }
    wsync();
  }
}

// correction of order k and products of order k+1 in one sweep over the
// levels: a node's corrected order-k value and its order-(k+1) product read
// only factors of lower levels (already final for both orders) and write
// distinct slots (ring storage keeps at least two orders when deg >= 1)
template <int K> __device__ void correct_and_product(const LProg &__restrict__ P, Path<K> &__restrict__ w, int k) {
  typedef TC<K> Cx;
  const int k1 = k + 1;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev[lv - 1] + lane_id(); i < P.lev[lv]; i += 64) {
      const LDS int *nd = P.nodes + 8 * i;
      const int aoff = nd[0], boff = nd[2], soff = nd[3], smask = nd[4], adn = nd[5], sdn = nd[6];
      const cplx<K> da = (adn < 0) ? w.slots[aoff + k] : w.dnode[adn];
      cplx<K> d = Cx::mul(w.slots[aoff], w.slots[boff + k]);
      Cx::mac(d, da, w.slots[boff]);
      if (sdn >= 0) w.dnode[sdn] = d;
      LDS cplx<K> &v = w.slots[soff + (k & smask)];
      v = Cx::add(v, d);
      cplx<K> a0 = Cx::mul(w.slots[aoff], w.slots[boff + k1]);
      cplx<K> a1 = Cx::mul(w.slots[aoff + 1], w.slots[boff + k1 - 1]);
      for (int q = 2; q <= k1; q += 2) {
        Cx::mac(a0, w.slots[aoff + q], w.slots[boff + k1 - q]);
        if (q + 1 <= k1) Cx::mac(a1, w.slots[aoff + q + 1], w.slots[boff + k1 - q - 1]);
      }
      w.slots[soff + (k1 & smask)] = Cx::add(a0, a1);
    }
    wsync();
  }
}

// rhs_i = -[H_i]_k from the weighted component values at order k
template <int K> __device__ void residual(const LProg &__restrict__ P, Path<K> &__restrict__ w, int k) {
  typedef TC<K> Cx;
  const int l = lane_id();
  for (int q = l; q < P.npair; q += 64) {
    const LDS int *pr = P.pairs + 4 * q;
    const int r = pr[0], off = pr[1], mask = pr[2];
    cplx<K> acc = Cx::mul(w.W[3 * r], w.slots[off + (k & mask)]);
    for (int j = 1; j <= P.wdeg && j <= k; ++j)
      Cx::mac(acc, w.W[3 * r + j], w.slots[off + ((k - j) & mask)]);
    w.u[q] = acc;
  }
  wsync();
  cplx<K> acc = Cx::zero();
  if (l < P.n * P.G) {
    // term t of the segment goes to accumulator (t - first) mod 4
    cplx<K> a0 = Cx::zero(), a1 = Cx::zero(), a2 = Cx::zero(), a3 = Cx::zero();
    const int t1 = P.seg[l + 1];
    int t = P.seg[l];
    for (; t + 4 <= t1; t += 4) {
      Cx::mac(a0, coef<K>(P, t), w.u[P.tpair[t]]);
      Cx::mac(a1, coef<K>(P, t + 1), w.u[P.tpair[t + 1]]);
      Cx::mac(a2, coef<K>(P, t + 2), w.u[P.tpair[t + 2]]);
      Cx::mac(a3, coef<K>(P, t + 3), w.u[P.tpair[t + 3]]);
    }
    if (t < t1) Cx::mac(a0, coef<K>(P, t), w.u[P.tpair[t]]);
    if (t + 1 < t1) Cx::mac(a1, coef<K>(P, t + 1), w.u[P.tpair[t + 1]]);
    if (t + 2 < t1) Cx::mac(a2, coef<K>(P, t + 2), w.u[P.tpair[t + 2]]);
    acc = Cx::add(Cx::add(a0, a1), Cx::add(a2, a3));
  }
  acc = group_sum<K>(acc, P.G);
  if (l < P.n * P.G && (l % P.G) == 0) w.rhs[l / P.G] = Cx::neg(acc);
  wsync();
}

// Jacobian A_ij = sum coef * e * W_r0 * node(m/x_j)_0
template <int K> __device__ void jacobian(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  typedef TC<K> Cx;
  const int l = lane_id();
  for (int q = l; q < P.njpair; q += 64) {
    const LDS int *jp = P.jpairs + 2 * q;
    w.uj[q] = Cx::mul(w.W[3 * jp[0]], w.slots[jp[1]]);
  }
  wsync();
  for (int e = l; e < P.n * P.n; e += 64) {
    // term t of the entry goes to accumulator (t - first) mod 4
    cplx<K> a[4] = {Cx::zero(), Cx::zero(), Cx::zero(), Cx::zero()};
    const int t0 = P.jent[e], t1 = P.jent[e + 1];
    for (int t = t0; t < t1; t += 4) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (t + r < t1) {
          const unsigned jt = (unsigned)P.jterms[t + r];
          const int term = jt & ((1u << JT_TERM_BITS) - 1);
          const int q = (jt >> JT_TERM_BITS) & ((1u << JT_PAIR_BITS) - 1);
          const int ex = jt >> (JT_TERM_BITS + JT_PAIR_BITS);
          const cplx<K> v = (ex == 1) ? w.uj[q] : Cx::scale_d(w.uj[q], (double)ex);
          Cx::mac(a[r], coef<K>(P, term), v);
        }
      }
    }
    w.A[e] = Cx::add(Cx::add(a[0], a[1]), Cx::add(a[2], a[3]));
  }
  wsync();
}

// In-place Gauss-Jordan inverse with partial pivoting: A := (P A)^{-1}
// where the row permutation P is recorded in perm ((P r)_i = r[perm[i]]).
// Returns false on a zero or non-finite pivot.  Any n: the row swap, the
// pivot column, the scaled pivot row and the elimination are separate
// phases.
template <int K> __device__ bool gj_invert_wide(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  typedef TC<K> Cx;
  const int n = P.n, l = lane_id();
  if (l < n) w.perm[l] = l;
  wsync();
  for (int j = 0; j < n; ++j) {
    double v = -1.0;
    int idx = l;
    if (l >= j && l < n) {
      v = absz<K>(w.A[l * n + j]);
      if (v != v) v = INFINITY;
    }
    argmax_stage<1>(v, idx);
    argmax_stage<2>(v, idx);
    argmax_stage<4>(v, idx);
    argmax_stage<8>(v, idx);
    argmax_stage<16>(v, idx);
    argmax_stage<32>(v, idx);
    if (!(v > 0.0) || v == INFINITY) return false;
    const int p = idx;
    if (p != j) {
      for (int c = l; c < n; c += 64) {
        const cplx<K> tmp = w.A[j * n + c];
        w.A[j * n + c] = w.A[p * n + c];
        w.A[p * n + c] = tmp;
      }
      if (l == 0) { const int t = w.perm[j]; w.perm[j] = w.perm[p]; w.perm[p] = t; }
      wsync();
    }
    if (l < n) w.col[l] = w.A[l * n + j];
    wsync();
    const cplx<K> inv = Cx::div(Cx::from(1.0, 0.0), w.col[j]);
    for (int c = l; c < n; c += 64) w.A[j * n + c] = (c == j) ? inv : Cx::mul(w.A[j * n + c], inv);
    wsync();
    for (int e = l; e < n * n; e += 64) {
      const int i = e / n, c = e - i * n;
      if (i == j) continue;
      if (c == j) w.A[e] = Cx::neg(Cx::mul(w.col[i], inv));
      else Cx::msub(w.A[e], w.col[i], w.A[j * n + c]);
    }
    wsync();
  }
  return true;
}

// The same inverse for n <= 16 (K <= 2) with the matrix in registers.  Lane (c, g) =
// (l % 16, l / 16) holds column c of rows g, g+4, g+8, g+12.  Per pivot j:
// the lanes of column j find the pivot row p among their rows (then across
// the four of them); the lanes holding row p, row j or column j store those
// entries to LDS; after one wavefront synchronisation every lane reads the
// pivot row, the old row j and the pivot column it needs and updates its
// entries of the swapped, scaled and eliminated matrix in registers.  Each
// entry gets the operations of the phased version in the same order, so
// the result is the same bit for bit.  The matrix is stored once at the end.
template <int K> __device__ bool gj_invert(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  typedef TC<K> Cx;
  const int n = P.n, l = lane_id();
  // quad-double entries would take 64 registers per lane: phased version
  if (n > 16 || K >= 4) return gj_invert_wide<K>(P, w);
  const int c = l & 15, g = l >> 4;
  const bool act = c < n;
  cplx<K> a[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int i = g + 4 * s;
    a[s] = (act && i < n) ? w.A[i * n + c] : Cx::zero();
  }
  if (l < n) w.perm[l] = l;
  for (int j = 0; j < n; ++j) {
    double v = -1.0;
    int idx = 64;
    if (c == j) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = g + 4 * s;
        if (i >= j && i < n) {
          double av = absz<K>(a[s]);
          if (av != av) av = INFINITY;
          if (av > v) { v = av; idx = i; }
        }
      }
    }
    // the four lanes of a column meet (l ^ 16, l ^ 32); column j's result from lane j
    argmax_stage<16>(v, idx);
    argmax_stage<32>(v, idx);
    v = __shfl(v, j, 64);
    idx = __shfl(idx, j, 64);
    if (!(v > 0.0) || v == INFINITY) return false;
    const int p = idx;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int i = g + 4 * s;
      if (act && i < n && (i == p || i == j || c == j)) w.A[i * n + c] = a[s];
    }
    wsync();  // rows p and j and column j of the unswapped matrix in LDS
    const cplx<K> piv = w.A[p * n + j], cjj = w.A[j * n + j];
    const cplx<K> inv = Cx::div(Cx::from(1.0, 0.0), piv);
    if (act) {
      const cplx<K> rc = w.A[p * n + c], aj = w.A[j * n + c];
      const cplx<K> r = (c == j) ? inv : Cx::mul(rc, inv);  // scaled pivot row
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int i = g + 4 * s;
        if (i < n) {
          // entry (i, c) of the swapped matrix: source row p for i = j, j for i = p
          const cplx<K> cs = (i == j) ? piv : ((i == p) ? cjj : w.A[i * n + j]);
          const cplx<K> sv = (i == j) ? rc : ((i == p) ? aj : a[s]);
          cplx<K> nv;
          if (i == j) nv = r;
          else if (c == j) nv = Cx::neg(Cx::mul(cs, inv));
          else {
            nv = sv;
            Cx::msub(nv, cs, r);
          }
          a[s] = nv;
        }
      }
    }
    if (l == 0 && p != j) { const int t = w.perm[j]; w.perm[j] = w.perm[p]; w.perm[p] = t; }
    wsync();  // every read of this pivot before the next pivot's stores
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int i = g + 4 * s;
    if (act && i < n) w.A[i * n + c] = a[s];
  }
  wsync();
  return true;
}

// y = (P A)^{-1} (P rhs): lanes i*G+g sum columns jlo..jhi of row i, then
// the G partial sums meet in a butterfly; row i's result is in the lanes of
// group i.
template <int K> __device__ cplx<K> apply_inverse(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  typedef TC<K> Cx;
  const int n = P.n, G = P.G, l = lane_id();
  cplx<K> acc = Cx::zero();
  if (l < n * G) {
    const int i = l / G, g = l - i * G;
    const int jlo = (g * n) / G, jhi = ((g + 1) * n) / G;
    for (int j = jlo; j < jhi; ++j) Cx::mac(acc, w.A[i * n + j], w.rhs[w.perm[j]]);
  }
  return group_sum<K>(acc, G);
}

template <int K> __device__ void eval_point(const LProg &__restrict__ P, Path<K> &__restrict__ w, bool jac) {
  product_pass<K>(P, w, 0);
  residual<K>(P, w, 0);
  if (jac) jacobian<K>(P, w);
}

template <int K> __device__ double max_abs_rhs(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  const int l = lane_id();
  return wave_max(l < P.n ? absz<K>(w.rhs[l]) : 0.0);
}

template <int K> __device__ double max_abs_x(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  const int l = lane_id();
  return wave_max(l < P.n ? absz<K>(w.slots[l * P.D]) : 0.0);
}

// Newton at fixed t on the order-0 slots (weights already set): returns
// 1 converged, 0 not converged, -1 singular; *dxl = last max |dx|.
// evaluated: the products and the residual at the current point are
// already in place (the corrector's predicted-residual check), so the first
// iteration only forms the Jacobian.
template <int K>
__device__ int newton(const LProg &__restrict__ P, Path<K> &__restrict__ w, int iters, double tol, double *dxl, int *count,
                      double *dxh, bool evaluated) {
  typedef TC<K> Cx;
  const int l = lane_id();
  for (int it = 0; it < iters; ++it) {
    ++*count;
    if (it == 0 && evaluated) jacobian<K>(P, w);
    else eval_point<K>(P, w, true);
    if (!gj_invert<K>(P, w)) return -1;
    const cplx<K> dx = apply_inverse<K>(P, w);
    double mdx = 0.0;
    if (l < P.n * P.G && (l % P.G) == 0) {
      LDS cplx<K> &x = w.slots[(l / P.G) * P.D];
      x = Cx::add(x, dx);
      mdx = absz<K>(dx);
    }
    wsync();
    mdx = wave_max(mdx);
    *dxl = mdx;
    if (dxh) { dxh[3] = dxh[2]; dxh[2] = dxh[1]; dxh[1] = dxh[0]; dxh[0] = mdx; }
    if (mdx <= tol * (1.0 + max_abs_x<K>(P, w))) return 1;
  }
  return 0;
}

// ------------------------------------------------------- the step rule ----

// x^(1/k) for x >= 0 from IEEE double operations only (the oracle's
// otr_root, bit for bit): log2 from the binary exponent and the series of
// ln((1+z)/(1-z)) to z^21, 2^y from the binary exponent and the Taylor
// series of e^g to g^16, then one Newton step on r^k = x
__host__ __device__ inline double kth_root(double x, int k) {
  if (x != x) return x;
  if (!(x > 0.0)) return 0.0;
  if (x == INFINITY || k == 1) return x;
  int e = 0;
  const double m = 2.0 * frexp(x, &e);
  const double z = (m - 1.0) / (m + 1.0), z2 = z * z;
  double s = 1.0 / 21.0;
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
  const double y = ((double)(e - 1) + (z * s) * 2.8853900817779268) / (double)k;
  const double fl = floor(y);
  const double g = (y - fl) * 0.69314718055994531;
  double p = 1.0;
  for (int i = 16; i >= 1; --i) p = 1.0 + (g / (double)i) * p;
  const double r = ldexp(p, (int)fl);
  double rk1 = 1.0;
  for (int i = 1; i < k; ++i) rk1 = rk1 * r;
  return r - (rk1 * r - x) / ((double)k * rk1);
}

// complex double helpers of the step rule (fused multiply-adds in the
// oracle's order)
__device__ __forceinline__ void zmac(double2 &z, const double2 &x, const double2 &y) {
  z.x = __builtin_fma(x.x, y.x, z.x);
  z.x = __builtin_fma(-x.y, y.y, z.x);
  z.y = __builtin_fma(x.y, y.x, z.y);
  z.y = __builtin_fma(x.x, y.y, z.y);
}
__device__ __forceinline__ void zcjmac(double2 &z, const double2 &x, const double2 &y) {
  z.x = __builtin_fma(x.x, y.x, z.x);
  z.x = __builtin_fma(x.y, y.y, z.x);
  z.y = __builtin_fma(-x.y, y.x, z.y);
  z.y = __builtin_fma(x.x, y.y, z.y);
}
__device__ __forceinline__ double2 zmul(const double2 &x, const double2 &y) {
  return make_double2(__builtin_fma(x.x, y.x, -(x.y * y.y)), __builtin_fma(x.x, y.y, x.y * y.x));
}
__device__ __forceinline__ double zrad(double re, double im) { return sqrt(re * re + im * im); }
template <int K> __device__ __forceinline__ double2 lead(const cplx<K> &z) {
  return make_double2(R<K>::hi(z.re), R<K>::hi(z.im));
}

// largest eigenvalue of M^H M (sigma_max(M)^2), M(i, j) the leading limbs
// of an n x n complex matrix: the Gram matrix scaled by its trace and
// squared six times, then the Rayleigh quotient |M v|^2/|v|^2 of the column
// v of largest diagonal entry (the oracle's otr_sigma_max_sq, same
// operations).  The Gram matrix and its powers are Hermitian: lanes own
// upper entries l, l+64, ... and store them with their conjugates; for
// n <= 16 a squaring keeps its new entries in registers until every lane
// has read the old ones, beyond that it goes through Q; Y holds M v (n
// entries).  Every lane returns the value.
// row and column of entry q of the upper triangle (row-major, i <= j)
__device__ __forceinline__ void upper_ij(int n, int q, int &i, int &j) {
  i = 0;
  while (q >= n - i) { q -= n - i; ++i; }
  j = i + q;
}

template <typename MF>
__device__ double sigma_max_sq(int n, MF M, LDS double2 *P, LDS double2 *Q, LDS double2 *Y) {
  const int l = lane_id(), nu = n * (n + 1) / 2;
  // Hermitian: entries i <= j computed, i > j their conjugates
  for (int q = l; q < nu; q += 64) {
    int i, j;
    upper_ij(n, q, i, j);
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < n; ++k) zcjmac(acc, M(k, i), M(k, j));
    P[i * n + j] = acc;
    P[j * n + i] = make_double2(acc.x, i == j ? acc.y : -acc.y);
  }
  wsync();
  double tr = 0.0;
  for (int i = 0; i < n; ++i) tr = tr + P[i * n + i].x;
  if (tr != tr) return tr;
  if (!(tr > 0.0)) return 0.0;
  // P holds the unscaled power; its scaled entries (x*sc) are formed where
  // read, with the roundings of the oracle's scaled copy
  double sc = 1.0 / tr;
  for (int s = 0; s < 6; ++s) {
    auto sp = [&](int e) { const double2 v = P[e]; return make_double2(v.x * sc, v.y * sc); };
    if (n <= 16) {
      // at most 136 upper entries: three per lane
      double2 q[3];
      int qi[3], qj[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int e = l + 64 * r;
        q[r] = make_double2(0.0, 0.0);
        qi[r] = qj[r] = 0;
        if (e < nu) {
          upper_ij(n, e, qi[r], qj[r]);
          for (int k = 0; k < n; ++k) zmac(q[r], sp(qi[r] * n + k), sp(k * n + qj[r]));
        }
      }
      wsync();
#pragma unroll
      for (int r = 0; r < 3; ++r)
        if (l + 64 * r < nu) {
          P[qi[r] * n + qj[r]] = q[r];
          P[qj[r] * n + qi[r]] = make_double2(q[r].x, qi[r] == qj[r] ? q[r].y : -q[r].y);
        }
    } else {
      for (int e = l; e < nu; e += 64) {
        int i, j;
        upper_ij(n, e, i, j);
        double2 acc = make_double2(0.0, 0.0);
        for (int k = 0; k < n; ++k) zmac(acc, sp(i * n + k), sp(k * n + j));
        Q[i * n + j] = acc;
        Q[j * n + i] = make_double2(acc.x, i == j ? acc.y : -acc.y);
      }
      wsync();
      for (int e = l; e < n * n; e += 64) P[e] = Q[e];
    }
    wsync();
    tr = 0.0;
    for (int i = 0; i < n; ++i) tr = tr + P[i * n + i].x;
    sc = 1.0 / tr;
  }
  int jm = 0;
  double best = P[0].x * sc;
  for (int i = 1; i < n; ++i) {
    const double v = P[i * n + i].x * sc;
    if (v > best) { best = v; jm = i; }
  }
  // y = M v in the lanes of the rows, then the sums in row order
  double2 y = make_double2(0.0, 0.0);
  if (l < n)
    for (int k = 0; k < n; ++k) {
      const double2 pv = P[k * n + jm];
      zmac(y, M(l, k), make_double2(pv.x * sc, pv.y * sc));
    }
  wsync();
  if (l < n) Y[l] = y;
  wsync();
  double num = 0.0, den = 0.0;
  for (int i = 0; i < n; ++i) num = num + Y[i].x * Y[i].x + Y[i].y * Y[i].y;
  for (int k = 0; k < n; ++k) {
    const double2 pv = P[k * n + jm];
    const double vx = pv.x * sc, vy = pv.y * sc;
    den = den + vx * vx + vy * vy;
  }
  wsync();
  return num / den;
}

// the mixed residual mean_i |f_i| / (|f|_i + 1) at the point in the order-0
// slots after eval_point (rhs = -f, weights set): |f|_i is equation i with
// the moduli of its coefficients (the components of a monomial added
// first) at the moduli of x (standard_mixed_residuals.adb:472-484), in FP64
// on the leading limbs; node moduli go to the scratch, equation values to col
template <int K> __device__ double mixed_residual(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  const int l = lane_id(), n = P.n, D = P.D, G = P.G;
  LDS double *rn = reinterpret_cast<LDS double *>(w.u);
  LDS double *mr = reinterpret_cast<LDS double *>(w.col);
  for (int j = l; j <= n; j += 64)
    rn[j] = (j < n) ? zrad(R<K>::hi(w.slots[j * D].re), R<K>::hi(w.slots[j * D].im)) : 1.0;
  wsync();
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev[lv - 1] + l; i < P.lev[lv]; i += 64) {
      const LDS int *nd = P.nodes + 8 * i;
      rn[n + 1 + i] = rn[nd[7]] * rn[nd[2] / D];
    }
    wsync();
  }
  double acc = 0.0;
  if (l < n * G) {
    double2 gz = make_double2(0.0, 0.0);
    const int t1 = P.seg[l + 1];
    for (int t = P.seg[l]; t < t1; ++t) {
      const int q = P.tpair[t], m = P.pairs[4 * q + 3], r = P.pairs[4 * q];
      const double2 cw = zmul(P.tcoef[t], lead<K>(w.W[3 * r]));
      gz.x = gz.x + cw.x;
      gz.y = gz.y + cw.y;
      if (t + 1 == t1 || P.pairs[4 * P.tpair[t + 1] + 3] != m) {
        acc = acc + zrad(gz.x, gz.y) * rn[m];
        gz = make_double2(0.0, 0.0);
      }
    }
  }
  if (G > 1) acc = acc + bfly<1>(acc);
  if (G > 2) acc = acc + bfly<2>(acc);
  if (G > 4) acc = acc + bfly<4>(acc);
  if (G > 8) acc = acc + bfly<8>(acc);
  if (G > 16) acc = acc + bfly<16>(acc);
  if (G > 32) acc = acc + bfly<32>(acc);
  if (l < n * G && (l % G) == 0) mr[l / G] = acc;
  wsync();
  double mix = 0.0;
  for (int i = 0; i < n; ++i)
    mix = mix + zrad(R<K>::hi(w.rhs[i].re), R<K>::hi(w.rhs[i].im)) / (mr[i] + 1.0);
  wsync();
  return mix / (double)n;
}

// sum over the equations of sigma_max(Hessian_i)^2 right after the order-0
// evaluation (every node at order 0): the program's constant for equations
// of degree <= 2, the Hessian from its terms otherwise (Distance,
// standard_predictor_convolutions.adb:558-572, with Cached_Singular_Values)
template <int K> __device__ double hessian_sum(const LProg &__restrict__ P, Path<K> &__restrict__ w) {
  const int l = lane_id(), n = P.n, nn = n * (n + 1) / 2;
  LDS double2 *H = reinterpret_cast<LDS double2 *>(w.u), *SP = H + n * n, *SQ = SP + n * n;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = P.hsig[i];
    if (s < 0.0) {
      for (int q = l; q < nn; q += 64) {
        int a = 0, rem = q;
        while (rem >= n - a) { rem -= n - a; ++a; }
        const int b = a + rem, e = i * nn + q;
        double2 z = make_double2(0.0, 0.0);
        for (int h = P.hent[e]; h < P.hent[e + 1]; ++h) {
          const int term = P.hterms[3 * h], hp = P.hterms[3 * h + 1];
          const double mult = (double)P.hterms[3 * h + 2];
          const int r = P.hpairs[2 * hp], off = P.hpairs[2 * hp + 1];
          const double2 v = zmul(zmul(P.tcoef[term], lead<K>(w.W[3 * r])), lead<K>(w.slots[off]));
          z.x = z.x + mult * v.x;
          z.y = z.y + mult * v.y;
        }
        H[a * n + b] = z;
        H[b * n + a] = z;
      }
      wsync();
      s = sqrt(sigma_max_sq(n, [&](int a, int b) { return H[a * n + b]; }, SP, SQ,
                            reinterpret_cast<LDS double2 *>(w.col)));
    }
    acc = acc + s * s;
  }
  return acc;
}

// Newton corrector with the mixed residual (LU_Newton_Steps,
// standard_inlined_newton_circuits.adb:388-422): converged when the mixed
// residual and max|dx| are both <= tol, then one extra step unless the
// residual or the update is zero; stops when both grow.  1 converged, 0
// not, -1 singular.  evaluated: the products and the residual at the point
// are in place, so the first iteration only forms the Jacobian.
template <int K>
__device__ int lu_newton_steps(const LProg &__restrict__ P, Path<K> &__restrict__ w, int maxit, double tol, bool evaluated,
                               double *dxl, int *count, int *nfirst, double *dxh) {
  typedef TC<K> Cx;
  const int l = lane_id();
  double prev_err = 1.0, prev_mix = 1.0;
  int cnt = 0;
  for (int k = 1; k <= maxit + 1; ++k) {
    ++*count;
    if (k == 1 && !evaluated) {
      eval_point<K>(P, w, true);
      ++*nfirst;
    } else {
      jacobian<K>(P, w);
    }
    if (!gj_invert<K>(P, w)) return -1;
    const cplx<K> dx = apply_inverse<K>(P, w);
    double mdx = 0.0;
    if (l < P.n * P.G && (l % P.G) == 0) {
      LDS cplx<K> &x = w.slots[(l / P.G) * P.D];
      x = Cx::add(x, dx);
      mdx = absz<K>(dx);
    }
    wsync();
    const double err = wave_max(mdx);
    *dxl = err;
    if (dxh) { dxh[3] = dxh[2]; dxh[2] = dxh[1]; dxh[1] = dxh[0]; dxh[0] = err; }
    eval_point<K>(P, w, false);
    const double res = max_abs_rhs<K>(P, w);
    const double mix = mixed_residual<K>(P, w);
    if (mix <= tol && err <= tol) {
      if (cnt == 1 || res == 0.0 || err == 0.0) return 1;
      ++cnt;
    } else if (k > 1 && mix > prev_mix && err > prev_err) {
      return 0;
    }
    prev_err = err;
    prev_mix = mix;
  }
  return 0;
}

__device__ void stage_program(const DevProgram &G, const ProgLayout &L, unsigned char *lds) {
  auto cp32 = [&](size_t off, const void *src, size_t bytes) {
    const unsigned *s = (const unsigned *)src;
    unsigned *d = (unsigned *)(lds + off);
    for (size_t i = threadIdx.x; i < bytes / 4; i += blockDim.x) d[i] = s[i];
  };
  cp32(L.nodes, G.nodes, sizeof(int) * 8 * G.nprod);
  cp32(L.lev, G.lev_ptr, sizeof(int) * (G.nlev + 1));
  cp32(L.pairs, G.pairs, sizeof(int) * 4 * G.npair);
  cp32(L.seg, G.seg_ptr, sizeof(int) * (G.n * G.G + 1));
  cp32(L.tpair, G.term_pair, sizeof(int) * G.nterm);
  // coefficients interleaved (re, im): one 16-byte LDS load per term
  double2 *tc = (double2 *)(lds + L.tcoef);
  for (int t = threadIdx.x; t < G.nterm; t += blockDim.x)
    tc[t] = make_double2(G.term_coef[t], G.term_coef[G.nterm + t]);
  cp32(L.jpairs, G.jpairs, sizeof(int) * 2 * G.njpair);
  cp32(L.jent, G.jent_ptr, sizeof(int) * (G.n * G.n + 1));
  cp32(L.jterms, G.jterms, sizeof(int) * G.njterm);
}

__device__ LProg lprog(const DevProgram &G, const ProgLayout &L, unsigned char *lds) {
  LProg P;
  P.n = G.n; P.D = G.D; P.ncomp = G.ncomp; P.wdeg = G.wdeg; P.G = G.G; P.nprod = G.nprod;
  P.nlev = G.nlev; P.npair = G.npair; P.nterm = G.nterm; P.njpair = G.njpair;
  P.nodes = (const LDS int *)(lds + L.nodes);
  P.lev = (const LDS int *)(lds + L.lev);
  P.pairs = (const LDS int *)(lds + L.pairs);
  P.seg = (const LDS int *)(lds + L.seg);
  P.tpair = (const LDS int *)(lds + L.tpair);
  P.tcoef = (const LDS double2 *)(lds + L.tcoef);
  P.jpairs = (const LDS int *)(lds + L.jpairs);
  P.jent = (const LDS int *)(lds + L.jent);
  P.jterms = (const LDS int *)(lds + L.jterms);
  P.weights = G.weights;
  P.hsig = G.hsig;
  P.hent = G.hent_ptr;
  P.hterms = G.hterms;
  P.hpairs = G.hpairs;
  return P;
}

template <int K> __device__ Path<K> carve(const PathLayout<K> &L, unsigned char *base) {
  Path<K> w;
  w.slots = (LDS cplx<K> *)(base + L.slots);
  w.u = (LDS cplx<K> *)(base + L.u);
  w.uj = (LDS cplx<K> *)(base + L.uj);
  w.A = (LDS cplx<K> *)(base + L.A);
  w.rhs = (LDS cplx<K> *)(base + L.rhs);
  w.x0 = (LDS cplx<K> *)(base + L.x0);
  w.col = (LDS cplx<K> *)(base + L.col);
  w.W = (LDS cplx<K> *)(base + L.W);
  w.dnode = (LDS cplx<K> *)(base + L.dnode);
  w.perm = (LDS int *)(base + L.perm);
  w.hid = (LDS float *)(base + L.hid);
  return w;
}

template <int K>
__device__ void load_point(const LProg &__restrict__ P, Path<K> &__restrict__ w, const double *x, int KS, int npaths, int path) {
  const int l = lane_id(), n = P.n;
  if (l < n) {
    cplx<K> z = C<K>::zero();
    double *re = reinterpret_cast<double *>(&z.re), *im = reinterpret_cast<double *>(&z.im);
    for (int i = 0; i < K && i < KS; ++i) {
      re[i] = x[(size_t)i * npaths * n + (size_t)path * n + l];
      im[i] = x[(size_t)(KS + i) * npaths * n + (size_t)path * n + l];
    }
    w.slots[l * P.D] = z;
  }
  // the constant node 1 + 0 s + 0 s^2 + ...
  for (int a = l; a < P.D; a += 64) w.slots[n * P.D + a] = (a == 0) ? C<K>::from(1.0, 0.0) : C<K>::zero();
  wsync();
}

template <int K>
__device__ void store_point(const LProg &__restrict__ P, Path<K> &__restrict__ w, double *x, int npaths, int path) {
  const int l = lane_id(), n = P.n;
  if (l < n) {
    const cplx<K> z = w.slots[l * P.D];
    const double *re = reinterpret_cast<const double *>(&z.re);
    const double *im = reinterpret_cast<const double *>(&z.im);
    for (int i = 0; i < K; ++i) {
      x[(size_t)i * npaths * n + (size_t)path * n + l] = re[i];
      x[(size_t)(K + i) * npaths * n + (size_t)path * n + l] = im[i];
    }
  }
}

// Taylor coefficients x_1..x_deg at the point in the order-0 slots (weights
// set at t0) and *hs2, the Hessian sum at the point; returns false on a
// singular Jacobian, else A holds its inverse
// Taylor coefficients of orders kb..ke-1 (the state after order kb-1 as
// series_orders leaves it: coefficient kb zeroed, its products formed)
template <int K> __device__ void series_orders(const LProg &__restrict__ P, Path<K> &__restrict__ w, int kb, int ke) {
  const int l = lane_id(), n = P.n, D = P.D;
  for (int k = kb; k < ke; ++k) {
    residual<K>(P, w, k);
    const cplx<K> xk = apply_inverse<K>(P, w);
    if (l < n * P.G && (l % P.G) == 0) w.slots[(l / P.G) * D + k] = xk;
    if (k + 1 < D && l < n) w.slots[l * D + k + 1] = TC<K>::zero();
    wsync();
    // the correction of the last order would only touch values no later
    // phase reads
    if (k + 1 < D) correct_and_product<K>(P, w, k);
  }
}

// Fabry's pole radius |c_{kd-1} / c_kd| (leading limbs, FP64) minimized
// over the components whose c_kd is not zero (Is_Zero: both parts vanish
// next to 1), the first of them starting the minimum; *pfail when every
// c_kd is zero.  mr: n entries of LDS scratch.
template <int K>
__device__ __forceinline__ void fabry_radius(const LProg &__restrict__ P, Path<K> &__restrict__ w, int kd,
                                             LDS double2 *mr, bool *pfail, double *r) {
  const int l = lane_id(), n = P.n, D = P.D;
  if (l < n) {
    const double2 a = lead<K>(w.slots[l * D + kd - 1]), b = lead<K>(w.slots[l * D + kd]);
    double rad = -1.0;
    if (!(b.x + 1.0 == 1.0 && b.y + 1.0 == 1.0)) {
      const double nrm = b.x * b.x + b.y * b.y;
      const double zr = (a.x * b.x + a.y * b.y) / nrm, zi = (a.y * b.x - a.x * b.y) / nrm;
      rad = zrad(zr, zi);
    }
    mr[l].x = rad;
  }
  wsync();
  bool fail = true;
  double m = 0.0;
  for (int j = 0; j < n; ++j) {
    const double v = mr[j].x;
    if (v == -1.0) continue;
    if (fail) { m = v; fail = false; }
    else if (v < m) m = v;
  }
  wsync();
  *pfail = fail;
  *r = m;
}

// the point's value, Jacobian (inverted), Hessian sum, and the Taylor
// coefficients of orders 1..ke-1 (ke <= D; series_orders continues them)
template <int K> __device__ bool predictor_series(const LProg &__restrict__ P, Path<K> &__restrict__ w, double *hs2, int ke) {
  const int l = lane_id(), n = P.n, D = P.D;
  eval_point<K>(P, w, true);
  *hs2 = hessian_sum<K>(P, w);
  if (!gj_invert<K>(P, w)) return false;
  if (l < n) w.slots[l * D + 1] = TC<K>::zero();
  wsync();
  product_pass<K>(P, w, 1);
  series_orders<K>(P, w, 1, ke);
  return true;
}

// features of the current step (identical in every lane)
template <int K>
__device__ void step_features(const LProg &__restrict__ P, Path<K> &__restrict__ w, double rad, double maxstep, double hprev,
                              double t, const double *dxh, int rejprev, int Dk, double *f) {
  const int l = lane_id(), n = P.n, D = P.D;
  f[0] = lg2(rad);
  f[1] = lg2(hprev);
  f[2] = t;
  for (int k = 1; k <= 8; ++k) {
    double c = 0.0;
    if (k < Dk && l < n) c = absz<K>(w.slots[l * D + k]);
    f[2 + k] = lg2(wave_max(c) + 1e-300);
  }
  for (int i = 0; i < 4; ++i) f[11 + i] = lg2(dxh[i] + 1e-300);
  f[15] = (double)rejprev;
}

// candidate chosen by the MLP: the largest c with logit >= threshold
template <int K> __device__ int mlp_choose(const Ctl &ctl, Path<K> &w, const double *f) {
  const int j = lane_id();  // hidden unit j
  constexpr int F = PHC_CTL_NFEAT, H = PHC_CTL_HIDDEN;
  const float *W1 = ctl.w, *b1 = W1 + F * H, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H;
  const float *b3 = W3 + H * ctl.ncand;
  float h = b1[j];
  for (int i = 0; i < F; ++i) h += W1[i * H + j] * (((float)f[i] - ctl.mu[i]) / ctl.sd[i]);
  h = h < 0.0f ? 0.0f : h;
  w.hid[j] = h;
  wsync();
  float g = b2[j];
  for (int i = 0; i < H; ++i) g += W2[i * H + j] * w.hid[i];
  g = g < 0.0f ? 0.0f : g;
  wsync();
  int best = 0;
  for (int c = 0; c < ctl.ncand; ++c) {
    float v = W3[j * ctl.ncand + c] * g;
    v += bfly<1>(v);
    v += bfly<2>(v);
    v += bfly<4>(v);
    v += bfly<8>(v);
    v += bfly<16>(v);
    v += bfly<32>(v);
    if (v + b3[c] >= ctl.thresh) best = c;
  }
  return best;
}

// [L/2] Pade coefficients of one component's series c[0..L+2] in place
// (the reference's Denominator_System, lufac and lusolve for a 2x2 system):
// a = [c_{L-1} c_L; c_L c_{L+1}] (c_{L-1} taken as zero when L = 1),
// b = -(c_{L+1}, c_{L+2}); partial pivoting on |re| + |im| of the leading
// limbs, multiplier -a_10/a_00; the solution reversed is (d_1, d_2), the
// numerator c_i + d_1 c_{i-1} + d_2 c_{i-2}; d_1, d_2 take the places of
// c_{L+1}, c_{L+2}.  A zero pivot leaves the denominator 1 and a zero
// numerator, or the Taylor numerator when b is zero.
template <int K> __device__ __forceinline__ void pade2_coefficients(LDS cplx<K> *c, int L, LDS cplx<K> *x0) {
  typedef TC<K> Cx;
  cplx<K> a00 = (L >= 2) ? c[L - 1] : Cx::zero();
  cplx<K> a01 = c[L], a10 = c[L], a11 = c[L + 1];
  cplx<K> b0 = Cx::neg(c[L + 1]), b1 = Cx::neg(c[L + 2]);
  const double m0 = absz<K>(a00), m1 = absz<K>(a10);
  const bool sw = m1 > m0;
  bool sing = (sw ? m1 : m0) == 0.0;
  if (!sing) {
    if (sw) { const cplx<K> t = a10; a10 = a00; a00 = t; }
    const cplx<K> t = Cx::neg(Cx::div(Cx::from(1.0, 0.0), a00));
    a10 = Cx::mul(t, a10);
    const cplx<K> tj = sw ? a11 : a01;
    if (sw) { a11 = a01; a01 = tj; }
    a11 = Cx::add(a11, Cx::mul(tj, a10));
  }
  if (R<K>::hi(a11.re) == 0.0 && R<K>::hi(a11.im) == 0.0) sing = true;
  if (sing) {
    const bool rhs0 = R<K>::hi(b0.re) == 0.0 && R<K>::hi(b0.im) == 0.0 &&
                      R<K>::hi(b1.re) == 0.0 && R<K>::hi(b1.im) == 0.0;
    if (!rhs0) {
      for (int i = 1; i <= L; ++i) c[i] = Cx::zero();
      *x0 = Cx::zero();
    }
    c[L + 1] = Cx::zero();
    c[L + 2] = Cx::zero();
    return;
  }
  if (sw) { const cplx<K> t = b0; b0 = b1; b1 = t; }
  b1 = Cx::add(b1, Cx::mul(b0, a10));
  b1 = Cx::div(b1, a11);
  b0 = Cx::add(b0, Cx::mul(Cx::neg(b1), a01));
  b0 = Cx::div(b0, a00);
  for (int i = L; i >= 1; --i) {
    cplx<K> v = Cx::add(c[i], Cx::mul(b1, c[i - 1]));
    if (i >= 2) v = Cx::add(v, Cx::mul(b0, c[i - 2]));
    c[i] = v;
  }
  c[L + 1] = b1;
  c[L + 2] = b0;
}

// [L/PD] Pade coefficients of one component in place (PD = 0 keeps the
// Taylor polynomial) and the error estimate of Solution_Error_Estimate
// (homotopy_pade_approximants.adb:558-568,597-622) from the series before
// it is replaced: top = the largest i <= L+PD+2 with |c_i| > 1e-5 (c_0 =
// x0), est = c_top + sum_{i<=min(PD,top)} d_i c_{top-i}, est = num_top - est
// when top <= L, est = 1 when top < 0 or |est| < 1e-5 (leading limbs)
template <int K, int PD>
__device__ double2 pade_with_error(LDS cplx<K> *c, int L, LDS cplx<K> *x0) {
  typedef TC<K> Cx;
  int top = -1;
  for (int i = L + PD + 2; i >= 0; --i) {
    if (absz<K>(i == 0 ? *x0 : c[i]) > 1e-5) { top = i; break; }
  }
  cplx<K> o0 = Cx::zero(), o1 = Cx::zero(), o2 = Cx::zero();
  if (top >= 0) o0 = (top == 0) ? *x0 : c[top];
  if (PD >= 1 && top >= 1) o1 = (top == 1) ? *x0 : c[top - 1];
  if (PD >= 2 && top >= 2) o2 = (top == 2) ? *x0 : c[top - 2];
  if (PD == 1) {
    // 1x1 denominator system c_L d = -c_{L+1}: numerator c_i + d c_{i-1}
    // in place, d in place of c_{L+1}; a zero c_L leaves the denominator 1
    // and the Taylor numerator when c_{L+1} is zero too, else a zero
    // numerator (Pade, lufac info /= 0)
    const cplx<K> cl = c[L], cn = c[L + 1];
    if (R<K>::hi(cl.re) == 0.0 && R<K>::hi(cl.im) == 0.0) {
      if (R<K>::hi(cn.re) != 0.0 || R<K>::hi(cn.im) != 0.0) {
        for (int i = 1; i <= L; ++i) c[i] = Cx::zero();
        *x0 = Cx::zero();
      }
      c[L + 1] = Cx::zero();
    } else {
      const cplx<K> d = Cx::div(Cx::neg(cn), cl);
      for (int i = L; i >= 1; --i) c[i] = Cx::add(c[i], Cx::mul(d, c[i - 1]));
      c[L + 1] = d;
    }
  }
  if (PD == 2) pade2_coefficients<K>(c, L, x0);
  cplx<K> res;
  if (top < 0) {
    res = Cx::from(1.0, 0.0);
  } else {
    res = o0;
    if (PD >= 1 && top >= 1) res = Cx::add(res, Cx::mul(c[L + 1], o1));
    if (PD >= 2 && top >= 2) res = Cx::add(res, Cx::mul(c[L + 2], o2));
    if (top <= L) res = Cx::sub(top == 0 ? *x0 : c[top], res);
  }
  if (absz<K>(res) < 1e-5) res = Cx::from(1.0, 0.0);
  return lead<K>(res);
}

// the predicted component at step s: the numerator (Taylor polynomial of
// degree L when PD = 0) by Horner with x0 added last, over 1 + d_1 s + ...
template <int K, int PD>
__device__ cplx<K> pade_value(const LDS cplx<K> *c, int L, const cplx<K> x0, double s) {
  typedef TC<K> Cx;
  cplx<K> acc = c[L];
  for (int k = L - 1; k >= 1; --k) acc = Cx::add(Cx::scale_d(acc, s), c[k]);
  cplx<K> p = Cx::add(x0, Cx::scale_d(acc, s));
  if (PD >= 1) {
    cplx<K> den = c[L + PD];
    for (int k = PD - 1; k >= 1; --k) den = Cx::add(Cx::scale_d(den, s), c[L + k]);
    p = Cx::div(p, Cx::add(Cx::scale_d(den, s), Cx::from(1.0, 0.0)));
  }
  return p;
}

// PD: denominator degree of the predictor (0: Taylor polynomial, 1 or 2:
// [L/PD] Pade approximant), a template parameter so that the Taylor variant
// keeps its registers.  One step (LU_Prediction,
// standard_predictor_convolutions.adb:1142-1185, and
// Predictor_Corrector_Loop, predictor_corrector_loops.adb:16-83): series,
// pole step from Fabry's ratio, curvature step from the singular values and
// the Pade error estimates, Minimum, Bounded_Update, the feedback halving
// the step while the mixed residual of the prediction is >= alpha, then the
// corrector, halving the step on failure.
template <int K, int MW, int PD>
__global__ void __launch_bounds__(64 * MW) track_kernel(DevProgram G, const Params *__restrict__ parp, int npaths,
                                                   const double *__restrict__ start, int KS,
                                                   const double *__restrict__ tstart,
                                                   double *__restrict__ endp, double *tend,
                                                   int *status, int *nsteps_out, int *nrej_out,
                                                   double *dx_out, int *work,
                                                   const Ctl *__restrict__ ctlp, int *counter) {
  // the step parameters and the controller come through device memory and
  // are read where used (scalar loads), not held in registers for the whole
  // kernel as by-value arguments would be
  const Params &par = *parp;
  const Ctl &ctl = *ctlp;
  typedef TC<K> Cx;
  extern __shared__ __align__(16) unsigned char lds[];
  const ProgLayout PL(G);
  const PathLayout<K> WL(G);
  stage_program(G, PL, lds);
  __syncthreads();
  const LProg P = lprog(G, PL, lds);
  Path<K> w = carve<K>(WL, lds + PL.total + (threadIdx.x >> 6) * WL.total);
  const int l = lane_id(), n = P.n, D = P.D;
  LDS double2 *mr = reinterpret_cast<LDS double2 *>(w.col);
  for (;;) {
    int path = 0;
    if (l == 0) path = atomicAdd(counter, 1);
    path = __shfl(path, 0, 64);
    if (path >= npaths) break;
    double t = 0.0, dxl = 0.0, hprev = par.maxstep;
    double dxh[4] = {1.0, 1.0, 1.0, 1.0};
    int st = ST_RUNNING, steps = 0, rej = 0, npred = 0, nnewt = 0, nfeval = 0, rejprev = 0;
    int ncorr = 0, nbpole = 0, nbhess = 0, nbmaxm = 0, nfirst = 0;
    int nlog = 0, pmode = ctl.mode;
    // a path the learned controller (mode 1) does not finish, stopped at a
    // t below ctl.retry, is tracked again from its start with the plain
    // rule; the work counters add up both attempts, the other outputs are
    // the last attempt's
    for (int attempt = 0; attempt < 2; ++attempt) {
    pmode = (attempt == 0) ? ctl.mode : 0;
    t = tstart ? tstart[path] : 0.0;
    dxl = 0.0; hprev = par.maxstep;
    dxh[0] = dxh[1] = dxh[2] = dxh[3] = 1.0;
    st = ST_RUNNING; steps = 0; rej = 0; rejprev = 0;
    load_point<K>(P, w, start, KS, npaths, path);
    while (st == ST_RUNNING && t < 1.0) {
      // ---- predictor: Taylor series of x(t+s) ----
      set_weights<K>(P, w, t);
      ++npred;
      double hs2;
      // the controller chooses after d0 coefficients; its candidate's
      // series degree may add orders (the plain rule: all D at once)
      int Dk = (pmode != 0) ? ctl.d0 : D;
      if (!predictor_series<K>(P, w, &hs2, Dk)) { st = ST_SINGULAR; break; }
      // ---- pole step: Fabry's ratio, minimized over the components whose
      // c_kd is not zero (Is_Zero), the first of them starting the minimum
      bool pfail = true;
      double r = 0.0;
      fabry_radius<K>(P, w, Dk - 1 - par.fabry_offset, mr, &pfail, &r);
      // ---- step-size and predictor-order controller (reads the Taylor series) ----
      int cand = 0;
      if (pmode != 0) {
        double feat[PHC_CTL_NFEAT];
        step_features<K>(P, w, pfail ? par.maxstep : r, par.maxstep, hprev, t, dxh, rejprev, Dk, feat);
        if (pmode == 1) cand = mlp_choose<K>(ctl, w, feat);
        else cand = (int)(mix_hash(ctl.seed, (unsigned)path, (unsigned)(npred - 1)) % (unsigned)ctl.ncand);
        // the features and the choice are logged now (the outcome after the
        // corrector), so that they are not kept live across it
        if (ctl.mode == 2 && nlog < ctl.cap && l == 0) {
#pragma unroll
          for (int i = 0; i < PHC_CTL_NFEAT; ++i)
            ctl.log_x[((size_t)path * ctl.cap + nlog) * PHC_CTL_NFEAT + i] = (float)feat[i];
          ctl.log_c[(size_t)path * ctl.cap + nlog] = cand;
        }
        if (ctl.dc[cand] > Dk) {
          series_orders<K>(P, w, Dk, ctl.dc[cand]);
          Dk = ctl.dc[cand];
          pfail = true;
          r = 0.0;
          fabry_radius<K>(P, w, Dk - 1 - par.fabry_offset, mr, &pfail, &r);
        }
      }
      const double pole = pfail ? par.maxstep : par.pbeta * r;
      // the [L/PD] predictor of a series of Dk coefficients
      const int pnum = par.pnum - (D - Dk);
      const int rej0 = rej, nnewt0 = nnewt;
      // ---- predictor coefficients and their error estimates ----
      if (l < n) {
        w.x0[l] = w.slots[l * D];
        mr[l] = pade_with_error<K, PD>(w.slots + l * D, pnum, w.x0 + l);
      }
      wsync();
      double nrm = 0.0;
      for (int j = 0; j < n; ++j) nrm = nrm + mr[j].x * mr[j].x + mr[j].y * mr[j].y;
      nrm = sqrt(nrm);
      wsync();
      // ---- curvature step: sigma_min(J) = 1 / sigma_max(J^{-1}) ----
      double step;
      {
        const double lam = sigma_max_sq(
            n, [&](int a, int b) { return lead<K>(w.A[a * n + b]); },
            reinterpret_cast<LDS double2 *>(w.u), reinterpret_cast<LDS double2 *>(w.u) + n * n, mr);
        const double sig1 = 1.0 / sqrt(lam);
        const double eta = (2.0 * sig1) / sqrt(hs2);
        const double curv = kth_root((par.cbeta * eta) / nrm, Dk - 1);
        // Minimum (three_way_minima.adb:49-62)
        if (pole < curv) {
          if (pole < par.maxstep) { step = pole; ++nbpole; }
          else { step = par.maxstep; ++nbmaxm; }
        } else if (curv < par.maxstep) { step = curv; ++nbhess; }
        else { step = par.maxstep; ++nbmaxm; }
      }
      if (pmode != 0) {
        step *= ctl.mult[cand];
        if (ctl.pcap > 0.0 && !pfail && r > 0.0) step = fmin(step, ctl.pcap * r);
      }
      // ---- Bounded_Update (three_way_minima.adb:96-109) ----
      double tn;
      if (t + step >= 1.0 - par.minstep) { step = 1.0 - t; tn = 1.0; }
      else tn = t + step;
      // ---- Predictor_Feedback (standard_predictor_convolutions.adb:958-993) ----
      bool evaluated = false;
      for (;;) {
        if (l < n) w.slots[l * D] = pade_value<K, PD>(w.slots + l * D, pnum, w.x0[l], step);
        wsync();
        set_weights<K>(P, w, tn);
        ++nfeval;
        eval_point<K>(P, w, false);
        const double mix = mixed_residual<K>(P, w);
        if (mix < par.alpha) { evaluated = true; break; }
        step = step / 2.0;
        ++rej;
        tn = t + step;
        if (step < par.minstep) break;
      }
      // ---- corrector, halving the step on failure ----
      for (;;) {
        if (!evaluated) set_weights<K>(P, w, tn);
        const int nr =
            lu_newton_steps<K>(P, w, par.corsteps, par.tolcor, evaluated, &dxl, &nnewt, &nfirst, dxh);
        ++ncorr;
        if (nr < 0) {
          st = ST_SINGULAR;
        } else if (nr == 1) {
          t = tn; hprev = step; break;
        } else {
          step = step / 2.0;
          ++rej;
          if (step < par.minstep) st = ST_MINSTEP;
        }
        if (st != ST_RUNNING) {
          // the path ends at its last accepted point (t, x0)
          if (l < n) w.slots[l * D] = w.x0[l];
          wsync();
          break;
        }
        tn = t + step;
        if (l < n) w.slots[l * D] = pade_value<K, PD>(w.slots + l * D, pnum, w.x0[l], step);
        wsync();
        evaluated = false;
      }
      rejprev = rej - rej0;
      if (ctl.mode == 2 && nlog < ctl.cap) {
        // accepted at the first attempt: the Newton iterations it took (>= 1)
        if (l == 0)
          ctl.log_y[(size_t)path * ctl.cap + nlog] = (rej == rej0 && st == ST_RUNNING) ? nnewt - nnewt0 : 0;
        ++nlog;
      }
      if (st != ST_RUNNING) break;
      ++steps;
      if (max_abs_x<K>(P, w) > par.divergence) { st = ST_DIVERGED; break; }
      if (steps >= par.maxsteps && t < 1.0) { st = ST_MAXSTEPS; break; }
    }
    if (st == ST_RUNNING) {
      set_weights<K>(P, w, 1.0);
      const int nr =
          lu_newton_steps<K>(P, w, par.endsteps, par.tolend, false, &dxl, &nnewt, &nfirst, nullptr);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
    }
    if (!(pmode == 1 && st != ST_SUCCESS && t < ctl.retry)) break;
    }
    store_point<K>(P, w, endp, npaths, path);
    if (l == 0) {
      tend[path] = t;
      status[path] = st;
      nsteps_out[path] = steps;
      nrej_out[path] = rej;
      dx_out[path] = dxl;
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
      if (ctl.mode == 2) ctl.log_n[path] = nlog;
    }
    wsync();
  }
}

// Newton at t = tval in K-limb precision from KS-limb points; writes the
// refined point, the last |dx|, the max residual |H(x,t)| after the last
// update (same precision) and the status.
template <int K>
__global__ void __launch_bounds__(64 * REFINE_MAX_WAVES) refine_kernel(DevProgram G, int npaths, double tval,
                                                    const double *__restrict__ xin, int KS,
                                                    double *__restrict__ xout, int iters,
                                                    double tol, double *dx_out, double *res_out,
                                                    int *status, const int *only, int *counter) {
  extern __shared__ __align__(16) unsigned char lds[];
  const ProgLayout PL(G);
  const PathLayout<K> WL(G);
  stage_program(G, PL, lds);
  __syncthreads();
  const LProg P = lprog(G, PL, lds);
  Path<K> w = carve<K>(WL, lds + PL.total + (threadIdx.x >> 6) * WL.total);
  const int l = lane_id();
  for (;;) {
    int path = 0;
    if (l == 0) path = atomicAdd(counter, 1);
    path = __shfl(path, 0, 64);
    if (path >= npaths) break;
    load_point<K>(P, w, xin, KS, npaths, path);
    if (only && only[path] != ST_SUCCESS) {
      // not a tracked endpoint: copied through, status = the tracker's
      store_point<K>(P, w, xout, npaths, path);
      if (l == 0) {
        dx_out[path] = INFINITY;
        res_out[path] = INFINITY;
        status[path] = only[path];
      }
      wsync();
      continue;
    }
    set_weights<K>(P, w, tval);
    double dxl = 0.0;
    int nr = 1, its = 0;
    if (iters > 0) nr = newton<K>(P, w, iters, tol, &dxl, &its, nullptr, false);
    eval_point<K>(P, w, false);
    const double res = max_abs_rhs<K>(P, w);
    store_point<K>(P, w, xout, npaths, path);
    if (l == 0) {
      dx_out[path] = dxl;
      res_out[path] = res;
      status[path] = nr;
    }
    wsync();
  }
}

// Where a path goes as t -> 1 (the question of the endgames of
// Deformations/End_Games), from the Taylor series of x(t+s) at the path's
// last point: ratio method on the last three coefficients of each component
// (FP64, leading limbs; q_k = c_{k+1}/c_k at k = D-2): the nearest
// singularity at s* = 1/((k+1) q_k - k q_{k-1}), the component like
// (s* - s)^alpha with alpha = k - (k+1) s* q_k.  The component whose last
// coefficient weighs most at the distance 1 - t gives alpha = Re and
// rho = Re s* (+inf when every component has a zero coefficient); rco =
// 1/(sigma_max(J) sigma_max(J^{-1})) (0 when J is singular), xmax =
// max |x_i|.  The oracle's otr_endgame, same operations.
__device__ __forceinline__ bool is_zero2(double2 z) { return z.x + 1.0 == 1.0 && z.y + 1.0 == 1.0; }

__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  const double nrm = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / nrm, (a.y * b.x - a.x * b.y) / nrm);
}

template <int K>
__global__ void __launch_bounds__(64 * REFINE_MAX_WAVES) endgame_kernel(DevProgram G, int npaths,
                                                     const double *__restrict__ x, int KS,
                                                     const double *__restrict__ tend,
                                                     double *alpha_out, double *rho_out,
                                                     double *rco_out, double *xmax_out,
                                                     int *counter) {
  extern __shared__ __align__(16) unsigned char lds[];
  const ProgLayout PL(G);
  const PathLayout<K> WL(G);
  stage_program(G, PL, lds);
  __syncthreads();
  const LProg P = lprog(G, PL, lds);
  Path<K> w = carve<K>(WL, lds + PL.total + (threadIdx.x >> 6) * WL.total);
  const int l = lane_id(), n = P.n, D = P.D, k = D - 2;
  LDS double2 *U = reinterpret_cast<LDS double2 *>(w.u), *Y = reinterpret_cast<LDS double2 *>(w.col);
  for (;;) {
    int path = 0;
    if (l == 0) path = atomicAdd(counter, 1);
    path = __shfl(path, 0, 64);
    if (path >= npaths) break;
    load_point<K>(P, w, x, KS, npaths, path);
    const double t = tend[path], dist = 1.0 - t;
    set_weights<K>(P, w, t);
    eval_point<K>(P, w, true);
    const double lamJ =
        sigma_max_sq(n, [&](int a, int b) { return lead<K>(w.A[a * n + b]); }, U, U + n * n, Y);
    double hs2, rco = 0.0;
    if (predictor_series<K>(P, w, &hs2, D)) {
      const double lamI =
          sigma_max_sq(n, [&](int a, int b) { return lead<K>(w.A[a * n + b]); }, U, U + n * n, Y);
      rco = 1.0 / sqrt(lamJ * lamI);
    }
    double xm = 0.0, alpha = INFINITY, rho = INFINITY, best = -1.0;
    for (int j = 0; j < n; ++j) {
      const double2 z = lead<K>(w.slots[j * D]);
      const double r = zrad(z.x, z.y);
      if (r > xm) xm = r;
    }
    if (rco != 0.0 && k >= 2) {
      for (int j = 0; j < n; ++j) {
        const double2 a0 = lead<K>(w.slots[j * D + k - 1]), a1 = lead<K>(w.slots[j * D + k]),
                      a2 = lead<K>(w.slots[j * D + k + 1]);
        if (is_zero2(a0) || is_zero2(a1) || is_zero2(a2)) continue;
        const double2 q0 = zdiv(a1, a0), q1 = zdiv(a2, a1);
        const double2 den = make_double2((double)(k + 1) * q1.x - (double)k * q0.x,
                                         (double)(k + 1) * q1.y - (double)k * q0.y);
        if (den.x == 0.0 && den.y == 0.0) continue;
        const double2 ss = zdiv(make_double2(1.0, 0.0), den), sq = zmul(ss, q1);
        double pw = 1.0;
        for (int i = 0; i <= k; ++i) pw = pw * dist;
        const double wt = zrad(a2.x, a2.y) * pw;
        if (wt > best) {
          best = wt;
          alpha = (double)k - (double)(k + 1) * sq.x;
          rho = ss.x;
        }
      }
    }
    if (l == 0) {
      alpha_out[path] = alpha;
      rho_out[path] = rho;
      rco_out[path] = rco;
      xmax_out[path] = xm;
    }
    wsync();
  }
}

}  // namespace

// ================================================================ C-ABI ====

struct phc_tracker {
  DevProgram P;
  std::vector<void *> bufs;
  int *counter = nullptr;
  void *args = nullptr;  // the kernel's Params and Ctl (one block per tracker)
  // the last call's stream and an event after its kernels: a call on another
  // stream first waits for it, so calls on one tracker never overlap
  hipEvent_t done = nullptr;
  void *last_stream = nullptr;
  bool used = false;
};

namespace {
// orders a call on `stream` after the tracker's previous call
int tracker_enter(phc_tracker *tr, void *stream) {
  if (tr->used && stream != tr->last_stream)
    PHC_HIP(hipStreamWaitEvent((hipStream_t)stream, tr->done, 0));
  return 0;
}
int tracker_leave(phc_tracker *tr, void *stream) {
  PHC_HIP(hipEventRecord(tr->done, (hipStream_t)stream));
  tr->last_stream = stream;
  tr->used = true;
  return 0;
}
}  // namespace

namespace {
struct TrackArgs {
  Params par;
  Ctl ctl;
};

// writes the argument block in stream order; the block travels as a kernel
// argument, captured at launch, so the host copy need not outlive the call
__global__ void store_args_kernel(TrackArgs a, TrackArgs *dst) {
  if (threadIdx.x == 0) *dst = a;
}
}  // namespace

namespace {

template <typename T> int upload(phc_tracker *tr, const T *src, size_t count, const T **dst) {
  void *p = nullptr;
  PHC_HIP(hipMalloc(&p, sizeof(T) * (count > 0 ? count : 1)));
  tr->bufs.push_back(p);
  if (count > 0) PHC_HIP(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = (const T *)p;
  return 0;
}

size_t path_bytes(int K, const DevProgram &P) {
  if (K == 1) return PathLayout<1>(P).total;
  if (K == 2) return PathLayout<2>(P).total;
  if (K == 4) return PathLayout<4>(P).total;
  return PathLayout<8>(P).total;
}

// launch geometry: one workgroup per CU with as many wavefronts as LDS holds
int geometry(const DevProgram &P, int K, int max_waves, int kmax, int npaths, int *grid, int *waves,
             size_t *lds) {
  int dev = 0, ncu = 0, ldsmax = 0;
  PHC_HIP(hipGetDevice(&dev));
  PHC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PHC_HIP(hipDeviceGetAttribute(&ldsmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t cap = (size_t)ldsmax < LDS_LIMIT ? (size_t)ldsmax : LDS_LIMIT;
  const size_t prog = ProgLayout(P).total, pb = path_bytes(K, P);
  PHC_REQUIRE(prog + pb <= cap, "tracker program and one path do not fit in LDS");
  int wv = (int)((cap - prog) / pb);
  if (wv > max_waves) wv = max_waves;
  if (wv > kmax) wv = kmax;
  const size_t bytes = prog + (size_t)wv * pb;
  long per_cu = (long)(LDS_PER_CU / bytes);
  if (per_cu < 1) per_cu = 1;
  long g = (long)ncu * per_cu;
  const long need = (npaths + wv - 1) / wv;
  if (g > need) g = need;
  *grid = (int)(g < 1 ? 1 : g);
  *waves = wv;
  *lds = bytes;
  return 0;
}

}  // namespace

extern "C" {

int phc_program_validate(const phc_program *d) {
  PHC_REQUIRE(d != nullptr, "phc_program_validate: null program");
  PHC_REQUIRE(d->n >= 1 && d->n <= 64, "phc_program: n must be in 1..64");
  PHC_REQUIRE(d->deg >= 0 && d->deg <= 32, "phc_program: deg must be in 0..32");
  PHC_REQUIRE(d->wdeg >= 0 && d->wdeg <= 2, "phc_program: weight degree must be <= 2");
  PHC_REQUIRE(d->G >= 1 && (d->G & (d->G - 1)) == 0 && d->n * d->G <= 64,
              "phc_program: G must be a power of two with n*G <= 64");
  PHC_REQUIRE(d->ncomp >= 1 && d->ncomp <= 64, "phc_program: ncomp must be in 1..64");
  PHC_REQUIRE(d->nprod >= 0 && d->nlev >= 0 && d->npair >= 0 && d->njpair >= 0 &&
              d->nterm >= 0 && d->njterm >= 0 && d->nfactor >= 0 && d->nslots >= 0,
              "phc_program: negative count");
  PHC_REQUIRE(d->njpair <= (1 << JT_PAIR_BITS) && d->nterm <= (1 << JT_TERM_BITS),
              "phc_program: too many terms for the packed Jacobian");
  const int D = d->deg + 1;
  const int need = d->wdeg < d->deg ? d->wdeg : d->deg;
  auto storage_ok = [&](int off, int mask) {
    if (off < 0) return false;
    if (mask == -1) return off + D <= d->nslots;
    // a ring keeps at least two orders when deg >= 1 (the predictor corrects
    // order k and forms order k+1 of a node in one sweep)
    return mask >= need && (d->deg == 0 || mask >= 1) && ((mask + 1) & mask) == 0 &&
           off + mask + 1 <= d->nslots;
  };
  PHC_REQUIRE((d->n + 1) * D <= d->nslots, "phc_program: no room for the variables and the constant");
  PHC_REQUIRE(d->lev_ptr[0] == 0 && d->lev_ptr[d->nlev] == d->nprod, "phc_program: bad levels");
  for (int lv = 1; lv <= d->nlev; ++lv)
    PHC_REQUIRE(d->lev_ptr[lv] >= d->lev_ptr[lv - 1], "phc_program: bad levels");
  for (int i = 0; i < d->nprod; ++i) {
    const int *nd = d->nodes + 8 * i;
    PHC_REQUIRE(storage_ok(nd[0], -1) && nd[1] == -1, "phc_program: left factor needs full storage");
    PHC_REQUIRE(nd[2] >= 0 && nd[2] % D == 0 && nd[2] / D < d->n,
                "phc_program: right factor must be a variable");
    PHC_REQUIRE(storage_ok(nd[3], nd[4]), "phc_program: node storage out of range");
    PHC_REQUIRE(nd[5] >= -1 && nd[5] < d->nfactor && nd[6] >= -1 && nd[6] < d->nfactor,
                "phc_program: bad correction index");
    PHC_REQUIRE(nd[5] >= 0 || (nd[0] % D == 0 && nd[0] / D < d->n),
                "phc_program: a left factor without correction index must be a variable");
    PHC_REQUIRE(nd[6] < 0 || nd[4] == -1, "phc_program: factor nodes need full storage");
  }
  // a product's left factor is a variable or a node of the previous level
  for (int lv = 1; lv <= d->nlev; ++lv)
    for (int i = d->lev_ptr[lv - 1]; i < d->lev_ptr[lv]; ++i) {
      const int adn = d->nodes[8 * i + 5];
      if (adn < 0) {
        PHC_REQUIRE(lv == 1, "phc_program: a product of two variables must be on level 1");
        continue;
      }
      bool found = false;
      for (int q = (lv >= 2 ? d->lev_ptr[lv - 2] : 0); q < d->lev_ptr[lv - 1]; ++q)
        if (d->nodes[8 * q + 6] == adn && d->nodes[8 * q + 3] == d->nodes[8 * i]) found = true;
      PHC_REQUIRE(found, "phc_program: left factor not on the previous level");
    }
  for (int q = 0; q < d->npair; ++q) {
    const int *pr = d->pairs + 4 * q;
    PHC_REQUIRE(pr[0] >= 0 && pr[0] < d->ncomp && storage_ok(pr[1], pr[2]),
                "phc_program: bad component pair");
  }
  for (int q = 0; q < d->njpair; ++q) {
    const int *jp = d->jpairs + 2 * q;
    PHC_REQUIRE(jp[0] >= 0 && jp[0] < d->ncomp && jp[1] >= 0 && jp[1] < d->nslots,
                "phc_program: bad Jacobian pair");
  }
  const int nseg = d->n * d->G;
  PHC_REQUIRE(d->seg_ptr[0] == 0 && d->seg_ptr[nseg] == d->nterm, "phc_program: bad residual segments");
  for (int s = 0; s < nseg; ++s)
    PHC_REQUIRE(d->seg_ptr[s + 1] >= d->seg_ptr[s], "phc_program: segments not increasing");
  for (int t = 0; t < d->nterm; ++t)
    PHC_REQUIRE(d->term_pair[t] >= 0 && d->term_pair[t] < d->npair, "phc_program: bad term");
  PHC_REQUIRE(d->jent_ptr[0] == 0 && d->jent_ptr[d->n * d->n] == d->njterm,
              "phc_program: bad Jacobian entries");
  for (int e = 0; e < d->n * d->n; ++e)
    PHC_REQUIRE(d->jent_ptr[e + 1] >= d->jent_ptr[e], "phc_program: Jacobian entries not increasing");
  for (int t = 0; t < d->njterm; ++t) {
    const unsigned jt = (unsigned)d->jterms[t];
    const int term = jt & ((1u << JT_TERM_BITS) - 1);
    const int q = (jt >> JT_TERM_BITS) & ((1u << JT_PAIR_BITS) - 1);
    const int ex = jt >> (JT_TERM_BITS + JT_PAIR_BITS);
    PHC_REQUIRE(term < d->nterm && q < d->njpair && ex >= 1, "phc_program: bad Jacobian term");
  }
  // node numbers: the left factor's and every pair's, against their storage
  auto node_off = [&](int id) {
    if (id < 0 || id > d->n + d->nprod) return -1;
    return id <= d->n ? id * D : d->nodes[8 * (id - d->n - 1) + 3];
  };
  for (int i = 0; i < d->nprod; ++i) {
    const int *nd = d->nodes + 8 * i;
    PHC_REQUIRE(nd[7] < d->n + 1 + i && node_off(nd[7]) == nd[0],
                "phc_program: the left factor's node number does not match its storage");
  }
  for (int q = 0; q < d->npair; ++q)
    PHC_REQUIRE(node_off(d->pairs[4 * q + 3]) == d->pairs[4 * q + 1],
                "phc_program: a pair's node number does not match its storage");
  // Hessians
  PHC_REQUIRE(d->hsig && d->hent_ptr && d->nhterm >= 0 && d->nhpair >= 0,
              "phc_program: missing Hessian data");
  const int nn = d->n * (d->n + 1) / 2;
  for (int i = 0; i < d->n; ++i)
    PHC_REQUIRE(d->hsig[i] >= 0.0 || d->hsig[i] == -1.0,
                "phc_program: hsig must be >= 0 or -1");
  PHC_REQUIRE(d->hent_ptr[0] == 0 && d->hent_ptr[d->n * nn] == d->nhterm,
              "phc_program: bad Hessian entries");
  for (int e = 0; e < d->n * nn; ++e)
    PHC_REQUIRE(d->hent_ptr[e + 1] >= d->hent_ptr[e], "phc_program: Hessian entries not increasing");
  for (int h = 0; h < d->nhterm; ++h)
    PHC_REQUIRE(d->hterms[3 * h] >= 0 && d->hterms[3 * h] < d->nterm && d->hterms[3 * h + 1] >= 0 &&
                    d->hterms[3 * h + 1] < d->nhpair && d->hterms[3 * h + 2] >= 1,
                "phc_program: bad Hessian term");
  for (int q = 0; q < d->nhpair; ++q)
    PHC_REQUIRE(d->hpairs[2 * q] >= 0 && d->hpairs[2 * q] < d->ncomp && d->hpairs[2 * q + 1] >= 0 &&
                    d->hpairs[2 * q + 1] < d->nslots,
                "phc_program: bad Hessian pair");
  return 0;
}

int phc_tracker_create(phc_tracker **out, const phc_program *d) {
  PHC_REQUIRE(out && d, "phc_tracker_create: null argument");
  *out = nullptr;
  if (phc_program_validate(d) != 0) return -2;
  phc_tracker *tr = new phc_tracker();
  DevProgram &P = tr->P;
  P.n = d->n; P.D = d->deg + 1; P.ncomp = d->ncomp; P.wdeg = d->wdeg; P.G = d->G;
  P.nprod = d->nprod; P.nlev = d->nlev; P.nslots = d->nslots; P.nfactor = d->nfactor;
  P.npair = d->npair; P.nterm = d->nterm; P.njpair = d->njpair; P.njterm = d->njterm;
  int rc = 0;
  rc |= upload<int>(tr, d->nodes, (size_t)8 * d->nprod, &P.nodes);
  rc |= upload<int>(tr, d->lev_ptr, d->nlev + 1, &P.lev_ptr);
  rc |= upload<int>(tr, d->pairs, (size_t)4 * d->npair, &P.pairs);
  rc |= upload<int>(tr, d->seg_ptr, (size_t)d->n * d->G + 1, &P.seg_ptr);
  rc |= upload<int>(tr, d->term_pair, d->nterm, &P.term_pair);
  rc |= upload<double>(tr, d->term_coef, (size_t)2 * d->nterm, &P.term_coef);
  rc |= upload<int>(tr, d->jpairs, (size_t)2 * d->njpair, &P.jpairs);
  rc |= upload<int>(tr, d->jent_ptr, (size_t)d->n * d->n + 1, &P.jent_ptr);
  rc |= upload<int>(tr, d->jterms, d->njterm, &P.jterms);
  rc |= upload<double>(tr, d->weights, (size_t)16 * 3 * d->ncomp, &P.weights);
  rc |= upload<double>(tr, d->hsig, d->n, &P.hsig);
  rc |= upload<int>(tr, d->hent_ptr, (size_t)d->n * (d->n * (d->n + 1) / 2) + 1, &P.hent_ptr);
  rc |= upload<int>(tr, d->hterms, (size_t)3 * d->nhterm, &P.hterms);
  rc |= upload<int>(tr, d->hpairs, (size_t)2 * d->nhpair, &P.hpairs);
  P.hdyn = 0;
  for (int i = 0; i < d->n; ++i)
    if (d->hsig[i] < 0.0) P.hdyn = 1;
  void *c = nullptr, *a = nullptr;
  if (rc == 0 && hipMalloc(&a, sizeof(TrackArgs)) == hipSuccess) tr->bufs.push_back(a);
  else a = nullptr;
  if (rc != 0 || !a || hipMalloc(&c, sizeof(int)) != hipSuccess) {
    phc_tracker_destroy(tr);
    if (rc == 0) phc::set_error("phc_tracker_create: device allocation failed");
    return -1;
  }
  tr->counter = (int *)c;
  tr->args = a;
  if (hipEventCreateWithFlags(&tr->done, hipEventDisableTiming) != hipSuccess) {
    tr->done = nullptr;
    phc_tracker_destroy(tr);
    phc::set_error("phc_tracker_create: event creation failed");
    return -1;
  }
  *out = tr;
  return 0;
}

int phc_tracker_destroy(phc_tracker *tr) {
  if (!tr) return 0;
  for (void *p : tr->bufs) (void)hipFree(p);
  (void)hipFree(tr->counter);
  if (tr->done) (void)hipEventDestroy(tr->done);
  delete tr;
  return 0;
}

int phc_tracker_lds_bytes(const phc_tracker *tr, int K, size_t *program_bytes, size_t *path_bytes_out) {
  PHC_REQUIRE(tr && program_bytes && path_bytes_out, "phc_tracker_lds_bytes: null argument");
  PHC_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, "phc_tracker_lds_bytes: K must be 1, 2, 4, 8");
  *program_bytes = ProgLayout(tr->P).total;
  *path_bytes_out = path_bytes(K, tr->P);
  return 0;
}

int phc_track(phc_tracker *tr, int K, int npaths, const double *start, int KS, double *endp,
              double *tend, int *status, int *nsteps, int *nrej, double *dxlast, int *work,
              const phc_track_params *prm, const phc_controller *ctlp, int waves, void *stream) {
  return phc_track_from(tr, K, npaths, start, KS, nullptr, endp, tend, status, nsteps, nrej,
                        dxlast, work, prm, ctlp, waves, stream);
}

int phc_track_from(phc_tracker *tr, int K, int npaths, const double *start, int KS,
                   const double *tstart, double *endp, double *tend, int *status, int *nsteps,
                   int *nrej, double *dxlast, int *work, const phc_track_params *prm,
                   const phc_controller *ctlp, int waves, void *stream) {
  PHC_REQUIRE(tr && prm, "phc_track: null argument");
  PHC_REQUIRE(K == 1 || K == 2, "phc_track: tracking precision K must be 1 or 2");
  PHC_REQUIRE(KS >= 1 && KS <= 8, "phc_track: start limbs KS must be in 1..8");
  PHC_REQUIRE(npaths >= 0, "phc_track: npaths < 0");
  PHC_REQUIRE(waves >= 1 && waves <= TRACK_MAX_WAVES, "phc_track: waves per workgroup must be in 1..12");
  PHC_REQUIRE(tr->P.D >= 2, "phc_track: the tracker program needs series degree >= 1");
  PHC_REQUIRE(prm->fabry_offset >= 0 && prm->fabry_offset < tr->P.D - 1,
              "phc_track: fabry_offset must be below the series degree");
  PHC_REQUIRE(prm->maxstep > 0 && prm->minstep > 0 && prm->maxsteps > 0 && prm->corsteps > 0 &&
              prm->endsteps >= 0, "phc_track: step parameters must be positive");
  PHC_REQUIRE(prm->pade_dendeg >= 0 && prm->pade_dendeg <= 2,
              "phc_track: pade_dendeg must be 0, 1 or 2");
  PHC_REQUIRE(tr->P.D >= 4, "phc_track: the step rule needs series degree >= 3");
  PHC_REQUIRE(prm->pade_dendeg == 0
                  ? (prm->pade_numdeg >= 0 && prm->pade_numdeg <= tr->P.D - 3)
                  : (prm->pade_numdeg >= 1 && prm->pade_numdeg + prm->pade_dendeg <= tr->P.D - 3),
              "phc_track: pade_numdeg + pade_dendeg + 2 must not exceed the series degree "
              "(the error estimate reads coefficient L+M+2)");
  PHC_REQUIRE(prm->cbeta >= 0.0, "phc_track: cbeta must be >= 0");
  if (npaths == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int grid = 0, wv = 0;
  size_t lds = 0;
  if (geometry(tr->P, K, waves, K == 1 ? TRACK_MAX_WAVES : 8, npaths, &grid, &wv, &lds) != 0)
    return -2;
  Params par;
  par.maxsteps = prm->maxsteps; par.corsteps = prm->corsteps; par.endsteps = prm->endsteps;
  par.fabry_offset = prm->fabry_offset; par.maxstep = prm->maxstep; par.minstep = prm->minstep;
  par.pbeta = prm->pbeta; par.alpha = prm->alpha; par.tolcor = prm->tolcor;
  par.tolend = prm->tolend; par.divergence = prm->divergence;
  par.pden = prm->pade_dendeg;
  par.pnum = (prm->pade_dendeg == 0 && prm->pade_numdeg == 0) ? tr->P.D - 3 : prm->pade_numdeg;
  par.cbeta = prm->cbeta;
  Ctl ctl;
  ctl.mode = 0; ctl.ncand = 1; ctl.cap = 0; ctl.w = ctl.mu = ctl.sd = nullptr; ctl.thresh = 0.0f;
  ctl.seed = 0; ctl.log_x = nullptr; ctl.log_c = ctl.log_y = ctl.log_n = nullptr; ctl.pcap = 0.0;
  for (int c = 0; c < PHC_CTL_MAXCAND; ++c) {
    ctl.mult[c] = 1.0;
    ctl.dc[c] = tr->P.D;
  }
  ctl.d0 = tr->P.D;
  ctl.retry = 0.99;
  if (ctlp && ctlp->mode != 0) {
    PHC_REQUIRE(ctlp->mode == 1 || ctlp->mode == 2, "phc_track: controller mode must be 0, 1 or 2");
    PHC_REQUIRE(ctlp->ncand >= 1 && ctlp->ncand <= PHC_CTL_MAXCAND, "phc_track: bad ncand");
    PHC_REQUIRE(tr->P.D >= 9, "phc_track: the controller features need series degree >= 8");
    for (int c = 0; c < ctlp->ncand; ++c)
      PHC_REQUIRE(ctlp->mult[c] > 0.0, "phc_track: candidate multipliers must be positive");
    PHC_REQUIRE(ctlp->mode != 1 || (ctlp->weights && ctlp->mu && ctlp->sd),
                "phc_track: learned mode needs weights, mu and sd");
    PHC_REQUIRE(ctlp->mode != 2 || (ctlp->log_capacity >= 1 && ctlp->log_x && ctlp->log_c &&
                                    ctlp->log_y && ctlp->log_n),
                "phc_track: explore mode needs the log buffers");
    ctl.mode = ctlp->mode; ctl.ncand = ctlp->ncand; ctl.cap = ctlp->log_capacity;
    for (int c = 0; c < ctlp->ncand; ++c) ctl.mult[c] = ctlp->mult[c];
    ctl.w = ctlp->weights; ctl.mu = ctlp->mu; ctl.sd = ctlp->sd; ctl.thresh = ctlp->logit_threshold;
    ctl.seed = ctlp->seed; ctl.log_x = ctlp->log_x; ctl.log_c = ctlp->log_c;
    ctl.log_y = ctlp->log_y; ctl.log_n = ctlp->log_n;
    PHC_REQUIRE(ctlp->pole_cap >= 0.0, "phc_track: pole_cap must be >= 0");
    ctl.pcap = ctlp->pole_cap;
    PHC_REQUIRE(ctlp->retry_below >= 0.0, "phc_track: retry_below must be >= 0");
    if (ctlp->retry_below > 0.0) ctl.retry = ctlp->retry_below;
    // series degrees: before the choice and per candidate (0: the program's)
    const int deg = tr->P.D - 1;
    auto usable = [&](int d) {
      return d >= 1 && d <= deg && d - par.fabry_offset >= 1 && par.pnum - (deg - d) >= 1;
    };
    const int d0 = ctlp->deg0 ? ctlp->deg0 : deg;
    PHC_REQUIRE(usable(d0), "phc_track: controller deg0 out of range (Fabry offset, Pade degrees)");
    ctl.d0 = d0 + 1;
    for (int c = 0; c < ctlp->ncand; ++c) {
      const int dcand = ctlp->degree[c] ? ctlp->degree[c] : deg;
      PHC_REQUIRE(usable(dcand) && dcand >= d0,
                  "phc_track: candidate series degree out of range (deg0..program degree)");
      ctl.dc[c] = dcand + 1;
    }
  }
  if (tracker_enter(tr, stream) != 0) return -1;
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  // the argument block is written in stream order, like the queue counter
  // (tracker_enter orders this call after the previous one on the tracker)
  TrackArgs blk;
  blk.par = par;
  blk.ctl = ctl;
  hipLaunchKernelGGL(store_args_kernel, dim3(1), dim3(64), 0, st, blk, (TrackArgs *)tr->args);
  PHC_HIP(hipGetLastError());
  const Params *dpar = (const Params *)((char *)tr->args + offsetof(TrackArgs, par));
  const Ctl *dctl = (const Ctl *)((char *)tr->args + offsetof(TrackArgs, ctl));
  // up to 8 waves: the variant with 256 VGPRs per lane; 9 to 12 waves (FP64
  // tracking only): the variant compiled for three waves per SIMD
#define TK(KK, MW, PD)                                                                        \
  do {                                                                                        \
    PHC_HIP(hipFuncSetAttribute((const void *)track_kernel<KK, MW, PD>,                      \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));       \
    hipLaunchKernelGGL((track_kernel<KK, MW, PD>), dim3(grid), dim3(64 * wv), lds, st, tr->P, \
                       dpar, npaths, start, KS, tstart, endp, tend, status, nsteps, nrej, dxlast, work, \
                       dctl, tr->counter);                                                    \
  } while (0)
  if (par.pden == 0) {
    if (K == 1 && wv > 8) TK(1, TRACK_MAX_WAVES, 0);
    else if (K == 1) TK(1, 8, 0);
    else TK(2, 8, 0);
  } else if (par.pden == 1) {
    if (K == 1 && wv > 8) TK(1, TRACK_MAX_WAVES, 1);
    else if (K == 1) TK(1, 8, 1);
    else TK(2, 8, 1);
  } else {
    if (K == 1 && wv > 8) TK(1, TRACK_MAX_WAVES, 2);
    else if (K == 1) TK(1, 8, 2);
    else TK(2, 8, 2);
  }
#undef TK
  PHC_HIP(hipGetLastError());
  return tracker_leave(tr, stream);
}

int phc_refine(phc_tracker *tr, int K, int npaths, double t, const double *xin, int KS,
               double *xout, int iters, double tol, double *dxlast, double *res, int *status,
               const int *only, int waves, void *stream) {
  PHC_REQUIRE(tr, "phc_refine: null tracker");
  PHC_REQUIRE(K == 1 || K == 2 || K == 4, "phc_refine: K must be 1, 2 or 4 (octo-double device "
              "refinement is switched off: its kernel stalled a device run and the cause is not "
              "established)");
  PHC_REQUIRE(KS >= 1 && KS <= 8, "phc_refine: KS must be in 1..8");
  PHC_REQUIRE(npaths >= 0 && iters >= 0, "phc_refine: negative size");
  PHC_REQUIRE(waves >= 1 && waves <= REFINE_MAX_WAVES, "phc_refine: waves per workgroup must be in 1..8");
  if (npaths == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int grid = 0, wv = 0;
  size_t lds = 0;
  if (geometry(tr->P, K, waves, REFINE_MAX_WAVES, npaths, &grid, &wv, &lds) != 0) return -2;
  if (tracker_enter(tr, stream) != 0) return -1;
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  const void *fn = K == 1 ? (const void *)refine_kernel<1> : K == 2 ? (const void *)refine_kernel<2>
                                                           : (const void *)refine_kernel<4>;
  PHC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
#define RK(KK) hipLaunchKernelGGL(refine_kernel<KK>, dim3(grid), dim3(64 * wv), lds, st, tr->P, npaths, t, xin, KS, xout, iters, tol, dxlast, res, status, only, tr->counter)
  if (K == 1) RK(1); else if (K == 2) RK(2); else RK(4);
#undef RK
  PHC_HIP(hipGetLastError());
  return tracker_leave(tr, stream);
}

int phc_endgame(phc_tracker *tr, int K, int npaths, const double *x, int KS, const double *tend,
                double *alpha, double *rho, double *rco, double *xmax, int waves, void *stream) {
  PHC_REQUIRE(tr, "phc_endgame: null tracker");
  PHC_REQUIRE(K == 1 || K == 2, "phc_endgame: K must be 1 or 2");
  PHC_REQUIRE(KS >= 1 && KS <= 8, "phc_endgame: KS must be in 1..8");
  PHC_REQUIRE(npaths >= 0, "phc_endgame: npaths < 0");
  PHC_REQUIRE(waves >= 1 && waves <= REFINE_MAX_WAVES, "phc_endgame: waves per workgroup must be in 1..8");
  PHC_REQUIRE(tr->P.D >= 4, "phc_endgame: the ratio method needs series degree >= 3");
  if (npaths == 0) return 0;
  PHC_REQUIRE(x && tend && alpha && rho && rco && xmax, "phc_endgame: null array");
  hipStream_t st = (hipStream_t)stream;
  int grid = 0, wv = 0;
  size_t lds = 0;
  if (geometry(tr->P, K, waves, REFINE_MAX_WAVES, npaths, &grid, &wv, &lds) != 0) return -2;
  if (tracker_enter(tr, stream) != 0) return -1;
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  const void *fn = K == 1 ? (const void *)endgame_kernel<1> : (const void *)endgame_kernel<2>;
  PHC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (K == 1)
    hipLaunchKernelGGL(endgame_kernel<1>, dim3(grid), dim3(64 * wv), lds, st, tr->P, npaths, x, KS,
                       tend, alpha, rho, rco, xmax, tr->counter);
  else
    hipLaunchKernelGGL(endgame_kernel<2>, dim3(grid), dim3(64 * wv), lds, st, tr->P, npaths, x, KS,
                       tend, alpha, rho, rco, xmax, tr->counter);
  PHC_HIP(hipGetLastError());
  return tracker_leave(tr, stream);
}

}  // extern "C"
