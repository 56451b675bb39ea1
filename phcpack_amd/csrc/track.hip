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
};

struct Params {
  int maxsteps, corsteps, endsteps, fabry_offset;
  double maxstep, minstep, pbeta, alpha, tolcor, tolend, divergence;
  int pnum, pden;  // predictor: numerator degree (>= 1), denominator degree (0, 1 or 2)
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

// per-wavefront path region (bytes)
template <int K> struct PathLayout {
  size_t slots, u, uj, A, rhs, x0, col, W, dnode, perm, hid, total;
  __host__ __device__ explicit PathLayout(const DevProgram &P) {
    const size_t z = sizeof(cplx<K>);
    size_t o = 0;
    slots = o; o += z * P.nslots;
    u = o; o += z * P.npair;
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
  const int *nodes, *lev, *pairs, *seg, *tpair, *jpairs, *jent, *jterms;
  const double2 *tcoef;
  const double *weights;  // global
};

template <int K> struct Path {
  cplx<K> *slots, *u, *uj, *A, *rhs, *x0, *col, *W, *dnode;
  int *perm;
  float *hid;
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
template <int K> __device__ void set_weights(const LProg &P, Path<K> &w, double t) {
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
template <int K> __device__ void product_pass(const LProg &P, Path<K> &w, int k) {
  typedef TC<K> Cx;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev[lv - 1] + lane_id(); i < P.lev[lv]; i += 64) {
      const int *nd = P.nodes + 8 * i;
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
template <int K> __device__ void correct_and_product(const LProg &P, Path<K> &w, int k) {
  typedef TC<K> Cx;
  const int k1 = k + 1;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev[lv - 1] + lane_id(); i < P.lev[lv]; i += 64) {
      const int *nd = P.nodes + 8 * i;
      const int aoff = nd[0], boff = nd[2], soff = nd[3], smask = nd[4], adn = nd[5], sdn = nd[6];
      const cplx<K> da = (adn < 0) ? w.slots[aoff + k] : w.dnode[adn];
      cplx<K> d = Cx::mul(w.slots[aoff], w.slots[boff + k]);
      Cx::mac(d, da, w.slots[boff]);
      if (sdn >= 0) w.dnode[sdn] = d;
      cplx<K> &v = w.slots[soff + (k & smask)];
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
template <int K> __device__ void residual(const LProg &P, Path<K> &w, int k) {
  typedef TC<K> Cx;
  const int l = lane_id();
  for (int q = l; q < P.npair; q += 64) {
    const int *pr = P.pairs + 4 * q;
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
template <int K> __device__ void jacobian(const LProg &P, Path<K> &w) {
  typedef TC<K> Cx;
  const int l = lane_id();
  for (int q = l; q < P.njpair; q += 64) {
    const int *jp = P.jpairs + 2 * q;
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
template <int K> __device__ bool gj_invert_wide(const LProg &P, Path<K> &w) {
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
template <int K> __device__ bool gj_invert(const LProg &P, Path<K> &w) {
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
template <int K> __device__ cplx<K> apply_inverse(const LProg &P, Path<K> &w) {
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

template <int K> __device__ void eval_point(const LProg &P, Path<K> &w, bool jac) {
  product_pass<K>(P, w, 0);
  residual<K>(P, w, 0);
  if (jac) jacobian<K>(P, w);
}

template <int K> __device__ double max_abs_rhs(const LProg &P, Path<K> &w) {
  const int l = lane_id();
  return wave_max(l < P.n ? absz<K>(w.rhs[l]) : 0.0);
}

template <int K> __device__ double max_abs_x(const LProg &P, Path<K> &w) {
  const int l = lane_id();
  return wave_max(l < P.n ? absz<K>(w.slots[l * P.D]) : 0.0);
}

// Newton at fixed t on the order-0 slots (weights already set): returns
// 1 converged, 0 not converged, -1 singular; *dxl = last max |dx|.
// evaluated: the products and the residual at the current point are
// already in place (the corrector's predicted-residual check), so the first
// iteration only forms the Jacobian.
template <int K>
__device__ int newton(const LProg &P, Path<K> &w, int iters, double tol, double *dxl, int *count,
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
      cplx<K> &x = w.slots[(l / P.G) * P.D];
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
  P.nodes = (const int *)(lds + L.nodes);
  P.lev = (const int *)(lds + L.lev);
  P.pairs = (const int *)(lds + L.pairs);
  P.seg = (const int *)(lds + L.seg);
  P.tpair = (const int *)(lds + L.tpair);
  P.tcoef = (const double2 *)(lds + L.tcoef);
  P.jpairs = (const int *)(lds + L.jpairs);
  P.jent = (const int *)(lds + L.jent);
  P.jterms = (const int *)(lds + L.jterms);
  P.weights = G.weights;
  return P;
}

template <int K> __device__ Path<K> carve(const PathLayout<K> &L, unsigned char *base) {
  Path<K> w;
  w.slots = (cplx<K> *)(base + L.slots);
  w.u = (cplx<K> *)(base + L.u);
  w.uj = (cplx<K> *)(base + L.uj);
  w.A = (cplx<K> *)(base + L.A);
  w.rhs = (cplx<K> *)(base + L.rhs);
  w.x0 = (cplx<K> *)(base + L.x0);
  w.col = (cplx<K> *)(base + L.col);
  w.W = (cplx<K> *)(base + L.W);
  w.dnode = (cplx<K> *)(base + L.dnode);
  w.perm = (int *)(base + L.perm);
  w.hid = (float *)(base + L.hid);
  return w;
}

template <int K>
__device__ void load_point(const LProg &P, Path<K> &w, const double *x, int KS, int npaths, int path) {
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
__device__ void store_point(const LProg &P, Path<K> &w, double *x, int npaths, int path) {
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
// set at t0); returns false on a singular Jacobian
template <int K> __device__ bool predictor_series(const LProg &P, Path<K> &w) {
  const int l = lane_id(), n = P.n, D = P.D;
  eval_point<K>(P, w, true);
  if (!gj_invert<K>(P, w)) return false;
  if (l < n) w.slots[l * D + 1] = TC<K>::zero();
  wsync();
  product_pass<K>(P, w, 1);
  for (int k = 1; k < D; ++k) {
    residual<K>(P, w, k);
    const cplx<K> xk = apply_inverse<K>(P, w);
    if (l < n * P.G && (l % P.G) == 0) w.slots[(l / P.G) * D + k] = xk;
    if (k + 1 < D && l < n) w.slots[l * D + k + 1] = TC<K>::zero();
    wsync();
    // the correction of the last order would only touch values no later
    // phase reads
    if (k + 1 < D) correct_and_product<K>(P, w, k);
  }
  return true;
}

// features of the current step (identical in every lane)
template <int K>
__device__ void step_features(const LProg &P, Path<K> &w, double rad, double maxstep, double hprev,
                              double t, const double *dxh, int rejprev, double *f) {
  const int l = lane_id(), n = P.n, D = P.D;
  f[0] = lg2((rad < INFINITY && rad > 0.0) ? rad : maxstep);
  f[1] = lg2(hprev);
  f[2] = t;
  for (int k = 1; k <= 8; ++k) {
    double c = 0.0;
    if (k < D && l < n) c = absz<K>(w.slots[l * D + k]);
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
template <int K> __device__ __forceinline__ void pade2_coefficients(cplx<K> *c, int L, cplx<K> *x0) {
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

// PD: denominator degree of the predictor (0: Taylor polynomial, 1 or 2:
// [L/PD] Pade approximant), a template parameter so that the Taylor variant
// keeps its registers
template <int K, int MW, int PD>
__global__ void __launch_bounds__(64 * MW) track_kernel(DevProgram G, const Params *__restrict__ parp, int npaths,
                                                   const double *__restrict__ start, int KS,
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
  for (;;) {
    int path = 0;
    if (l == 0) path = atomicAdd(counter, 1);
    path = __shfl(path, 0, 64);
    if (path >= npaths) break;
    double t = 0.0, h = par.maxstep, dxl = 0.0, hprev = par.maxstep;
    double dxh[4] = {1.0, 1.0, 1.0, 1.0};
    int st = ST_RUNNING, steps = 0, rej = 0, npred = 0, nnewt = 0, nfeval = 0, rejprev = 0;
    int ncorr = 0;
    int nlog = 0, pmode = ctl.mode;
    // a path the learned controller (mode 1) does not finish is tracked
    // again from its start with the plain pole-step rule; the work counters
    // add up both attempts, the other outputs are the last attempt's
    for (int attempt = 0; attempt < 2; ++attempt) {
    pmode = (attempt == 0) ? ctl.mode : 0;
    t = 0.0; h = par.maxstep; dxl = 0.0; hprev = par.maxstep;
    dxh[0] = dxh[1] = dxh[2] = dxh[3] = 1.0;
    st = ST_RUNNING; steps = 0; rej = 0; rejprev = 0;
    load_point<K>(P, w, start, KS, npaths, path);
    while (st == ST_RUNNING && t < 1.0) {
      // ---- predictor: Taylor series of x(t+s) ----
      set_weights<K>(P, w, t);
      ++npred;
      if (!predictor_series<K>(P, w)) { st = ST_SINGULAR; break; }
      // ---- pole step from the Fabry ratio ----
      const int kd = D - 1 - par.fabry_offset;
      double r = INFINITY;
      if (kd >= 1 && l < n) {
        const double den = absz<K>(w.slots[l * D + kd]);
        if (den > 0.0) r = absz<K>(w.slots[l * D + kd - 1]) / den;
      }
      r = wave_min(r);
      h = (r < INFINITY && r > 0.0) ? fmin(par.pbeta * r, par.maxstep) : par.maxstep;
      // ---- step-size controller ----
      int cand = 0;
      if (pmode != 0) {
        double feat[PHC_CTL_NFEAT];
        step_features<K>(P, w, r, par.maxstep, hprev, t, dxh, rejprev, feat);
        if (pmode == 1) cand = mlp_choose<K>(ctl, w, feat);
        else cand = (int)(mix_hash(ctl.seed, (unsigned)path, (unsigned)(npred - 1)) % (unsigned)ctl.ncand);
        h *= ctl.mult[cand];
        if (ctl.pcap > 0.0 && r < INFINITY && r > 0.0) h = fmin(h, ctl.pcap * r);
        // the features and the choice are logged now (the outcome after the
        // corrector), so that they are not kept live across it
        if (ctl.mode == 2 && nlog < ctl.cap && l == 0) {
#pragma unroll
          for (int i = 0; i < PHC_CTL_NFEAT; ++i)
            ctl.log_x[((size_t)path * ctl.cap + nlog) * PHC_CTL_NFEAT + i] = (float)feat[i];
          ctl.log_c[(size_t)path * ctl.cap + nlog] = cand;
        }
      }
      const int rej0 = rej, nnewt0 = nnewt;
      if (l < n) w.x0[l] = w.slots[l * D];
      // ---- [L/1] Pade coefficients per component ----
      // The 1x1 denominator system c_L d = -c_{L+1} gives the denominator
      // 1 + d s; the numerator coefficients c_i + d c_{i-1} (i = 1..L)
      // replace the series in place and d takes the place of c_{L+1}.  A
      // zero c_L leaves the denominator 1 and the Taylor numerator when
      // c_{L+1} is zero too, else a zero numerator (Pade, lufac info /= 0).
      // ---- [L/1] Pade coefficients per component ----
      // The 1x1 denominator system c_L d = -c_{L+1} gives the denominator
      // 1 + d s; the numerator coefficients c_i + d c_{i-1} (i = 1..L)
      // replace the series in place and d takes the place of c_{L+1}.  A
      // zero c_L leaves the denominator 1 and the Taylor numerator when
      // c_{L+1} is zero too, else a zero numerator (Pade, lufac info /= 0).
      if (PD == 1 && l < n) {
        const int L = par.pnum;
        cplx<K> *c = w.slots + l * D;
        const cplx<K> cl = c[L], cn = c[L + 1];
        if (R<K>::hi(cl.re) == 0.0 && R<K>::hi(cl.im) == 0.0) {
          if (R<K>::hi(cn.re) != 0.0 || R<K>::hi(cn.im) != 0.0) {
            for (int i = 1; i <= L; ++i) c[i] = Cx::zero();
            w.x0[l] = Cx::zero();
          }
          c[L + 1] = Cx::zero();
        } else {
          const cplx<K> d = Cx::div(Cx::neg(cn), cl);
          for (int i = L; i >= 1; --i) c[i] = Cx::add(c[i], Cx::mul(d, c[i - 1]));
          c[L + 1] = d;
        }
      }
      if (PD == 2 && l < n) pade2_coefficients<K>(w.slots + l * D, par.pnum, w.x0 + l);
      wsync();
      // ---- feedback and corrector; halve the step on failure ----
      bool accepted = false;
      while (!accepted) {
        double tn = t + h;
        if (tn >= 1.0 - par.minstep) tn = 1.0;
        const double hs = tn - t;
        if (l < n) {
          // numerator (the Taylor polynomial of degree L when pden = 0) by
          // Horner, then the quotient by the denominator 1 + d s
          const int L = par.pnum;
          cplx<K> acc = w.slots[l * D + L];
          for (int k = L - 1; k >= 1; --k) acc = Cx::add(Cx::scale_d(acc, hs), w.slots[l * D + k]);
          cplx<K> p = Cx::add(w.x0[l], Cx::scale_d(acc, hs));
          if (PD >= 1) {
            cplx<K> den = w.slots[l * D + L + PD];
            for (int k = PD - 1; k >= 1; --k) den = Cx::add(Cx::scale_d(den, hs), w.slots[l * D + L + k]);
            p = Cx::div(p, Cx::add(Cx::scale_d(den, hs), Cx::from(1.0, 0.0)));
          }
          w.slots[l * D] = p;
        }
        wsync();
        set_weights<K>(P, w, tn);
        ++nfeval;
        eval_point<K>(P, w, false);
        const double res = max_abs_rhs<K>(P, w);
        bool ok = res <= par.alpha * (1.0 + max_abs_x<K>(P, w));
        if (ok) {
          const int nr = newton<K>(P, w, par.corsteps, par.tolcor, &dxl, &nnewt, dxh, true);
          ++ncorr;
          if (nr < 0) { st = ST_SINGULAR; break; }
          ok = (nr == 1);
        }
        if (ok) { t = tn; accepted = true; hprev = hs; }
        else {
          ++rej;
          h *= 0.5;
          if (h < par.minstep) { st = ST_MINSTEP; break; }
        }
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
      const int nr = newton<K>(P, w, par.endsteps, par.tolend, &dxl, &nnewt, nullptr, false);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
    }
    if (!(pmode == 1 && st != ST_SUCCESS)) break;
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

}  // namespace

// ================================================================ C-ABI ====

struct phc_tracker {
  DevProgram P;
  std::vector<void *> bufs;
  int *counter = nullptr;
  void *args = nullptr;  // the kernel's Params and Ctl (one block per tracker)
};

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
  *out = tr;
  return 0;
}

int phc_tracker_destroy(phc_tracker *tr) {
  if (!tr) return 0;
  for (void *p : tr->bufs) (void)hipFree(p);
  (void)hipFree(tr->counter);
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
  PHC_REQUIRE(prm->pade_dendeg == 0
                  ? (prm->pade_numdeg >= 0 && prm->pade_numdeg <= tr->P.D - 1)
                  : (prm->pade_numdeg >= 1 && prm->pade_numdeg <= tr->P.D - 1 - prm->pade_dendeg),
              "phc_track: pade_numdeg must be in 0..deg (Taylor) or 1..deg-M (Pade [L/M])");
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
  par.pnum = (prm->pade_dendeg == 0 && prm->pade_numdeg == 0) ? tr->P.D - 1 : prm->pade_numdeg;
  Ctl ctl;
  ctl.mode = 0; ctl.ncand = 1; ctl.cap = 0; ctl.w = ctl.mu = ctl.sd = nullptr; ctl.thresh = 0.0f;
  ctl.seed = 0; ctl.log_x = nullptr; ctl.log_c = ctl.log_y = ctl.log_n = nullptr; ctl.pcap = 0.0;
  for (int c = 0; c < PHC_CTL_MAXCAND; ++c) ctl.mult[c] = 1.0;
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
  }
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  // the argument block is written in stream order, like the queue counter:
  // calls on one tracker must not overlap (the same stream, or synchronized)
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
                       dpar, npaths, start, KS, endp, tend, status, nsteps, nrej, dxlast, work, \
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
  return 0;
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
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  const void *fn = K == 1 ? (const void *)refine_kernel<1> : K == 2 ? (const void *)refine_kernel<2>
                                                           : (const void *)refine_kernel<4>;
  PHC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
#define RK(KK) hipLaunchKernelGGL(refine_kernel<KK>, dim3(grid), dim3(64 * wv), lds, st, tr->P, npaths, t, xin, KS, xout, iters, tol, dxlast, res, status, only, tr->counter)
  if (K == 1) RK(1); else if (K == 2) RK(2); else RK(4);
#undef RK
  PHC_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
