// Batched predictor-corrector path tracking of homotopies on gfx950, with
// endpoint Newton refinement and residuals in any of double, double-double,
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
//   corrector        LU Newton steps, step halved on failure down to minstep
//   batching         GPU_Path_mult (Path/Path_GPU/path_gpu_mult.cu:84-308)
//                    compacts live paths on the host after every step; here
//                    each workgroup pulls the next path from an atomic queue
//                    when its path ends, so no path waits and nothing returns
//                    to the host until all paths are done.
//
// Power series of the solution curve.  At the base point (x0, t0) the
// Taylor coefficients x_k of x(t0+s) are computed order by order: with the
// coefficient x_k set to zero, coefficient k of H(x(s), t0+s) is evaluated
// through the monomial product DAG; it is affine in x_k with the Jacobian
// A0 = H_x(x0,t0) as linear part, so x_k = -A0^{-1} [H]_k.  This is the lower
// block-triangular Toeplitz solve of power-series Newton
// (Newton/dbl2_bals_host.cpp:228-335, x_k = A0^{-1}(b_k - sum_j A_j x_{k-j}))
// with the tail sum formed implicitly by the convolutions, one LU of A0 per
// step and no stored Jacobian series.
//
// Layout: one workgroup per path; every per-path array lives in LDS
// (node series, weighted component values, Jacobian LU, vectors).  Lanes
// split the work of each phase (nodes of a DAG level, residual segments,
// Jacobian entries, matrix rows); all control decisions are computed by
// every lane from LDS values, so branches and barriers stay uniform.
#include <hip/hip_runtime.h>
#include <vector>
#include "common.hpp"
#include "mdtypes.hpp"
#include "../../include/phcmd.h"

using namespace md;

namespace {

enum { ST_RUNNING = 0, ST_SUCCESS = 1, ST_MINSTEP = 2, ST_MAXSTEPS = 3, ST_SINGULAR = 4,
       ST_DIVERGED = 5 };

struct DevProgram {
  int n, D, nprod, nlev, nslots, ring, wdeg, ncomp;
  const int *left, *right, *lev_ptr, *off, *slen;
  int npair;
  const int *pair_comp, *pair_node;
  int nseg;
  const int *seg_ptr, *eq_seg, *term_pair;
  int nterm;
  const double *term_coef;
  int njpair;
  const int *jpair_comp, *jpair_node, *jent_ptr, *jterm_pair;
  int njterm;
  const double *jterm_coef;
  const double *weights;
};

struct Params {
  int maxsteps, corsteps, endsteps, fabry_offset;
  double maxstep, minstep, pbeta, alpha, tolcor, tolend, divergence;
};

// complex coefficient j of an array of m coefficients stored as octo-double
// planes [2][8][m]; the leading K limbs are used
template <int K>
__device__ __forceinline__ cplx<K> load_coef(const double *base, int m, int j) {
  cplx<K> z;
  z.re = R<K>::load(base, m, j);
  z.im = R<K>::load(base + 8 * (size_t)m, m, j);
  return z;
}

template <int K> struct Work {
  typedef cplx<K> Z;
  Z *slots, *u, *uj, *A, *part, *rhs, *x0, *W, *dnode, *invd;
  int *piv;
  double *red;  // reduction scratch (one double per lane)
};

template <int K> __device__ __forceinline__ double absz(const cplx<K> &z) {
  return fabs(R<K>::hi(z.re)) + fabs(R<K>::hi(z.im));
}

// node m coefficient a (a >= 0)
template <int K>
__device__ __forceinline__ cplx<K> &nodec(const DevProgram &P, Work<K> &w, int m, int a) {
  return w.slots[P.off[m] + (a % P.slen[m])];
}

// Taylor coefficients W[r][l] at t of w_r(t) = om0 + om1 t + om2 t^2
template <int K> __device__ void set_weights(const DevProgram &P, Work<K> &w, double t) {
  typedef C<K> Cx;
  for (int r = threadIdx.x; r < P.ncomp; r += blockDim.x) {
    cplx<K> o0 = load_coef<K>(P.weights, P.ncomp * 3, 3 * r);
    cplx<K> o1 = load_coef<K>(P.weights, P.ncomp * 3, 3 * r + 1);
    cplx<K> o2 = load_coef<K>(P.weights, P.ncomp * 3, 3 * r + 2);
    cplx<K> o2t = Cx::scale_d(o2, t);
    w.W[3 * r] = Cx::add(o0, Cx::scale_d(Cx::add(o1, o2t), t));
    w.W[3 * r + 1] = Cx::add(o1, Cx::scale_d(o2t, 2.0));
    w.W[3 * r + 2] = o2;
  }
}

// products of the DAG at order k: node_k = sum_a left_a * right_{k-a}
template <int K> __device__ void product_pass(const DevProgram &P, Work<K> &w, int k) {
  typedef C<K> Cx;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev_ptr[lv - 1] + threadIdx.x; i < P.lev_ptr[lv]; i += blockDim.x) {
      const int a = P.left[i], b = P.right[i];
      cplx<K> acc = Cx::mul(nodec(P, w, a, 0), nodec(P, w, b, k));
      for (int q = 1; q <= k; ++q) Cx::mac(acc, nodec(P, w, a, q), nodec(P, w, b, k - q));
      nodec(P, w, P.n + 1 + i, k) = acc;
    }
    __syncthreads();
  }
}

// node_k += left_0*dright_k + dleft_k*right_0 once x_k is known
template <int K> __device__ void correction_pass(const DevProgram &P, Work<K> &w, int k) {
  typedef C<K> Cx;
  for (int lv = 1; lv <= P.nlev; ++lv) {
    for (int i = P.lev_ptr[lv - 1] + threadIdx.x; i < P.lev_ptr[lv]; i += blockDim.x) {
      const int a = P.left[i], b = P.right[i];  // b is a variable
      cplx<K> db = nodec(P, w, b, k);
      cplx<K> da = (a < P.n) ? nodec(P, w, a, k) : (a == P.n ? Cx::zero() : w.dnode[a - P.n - 1]);
      cplx<K> d = Cx::mul(nodec(P, w, a, 0), db);
      Cx::mac(d, da, nodec(P, w, b, 0));
      w.dnode[i] = d;
      cplx<K> &v = nodec(P, w, P.n + 1 + i, k);
      v = Cx::add(v, d);
    }
    __syncthreads();
  }
}

// residual rhs_i = [H_i]_k from weighted component values at order k
template <int K> __device__ void residual(const DevProgram &P, Work<K> &w, int k) {
  typedef C<K> Cx;
  for (int q = threadIdx.x; q < P.npair; q += blockDim.x) {
    const int r = P.pair_comp[q], m = P.pair_node[q];
    cplx<K> acc = Cx::mul(w.W[3 * r], nodec(P, w, m, k));
    for (int l = 1; l <= P.wdeg && l <= k; ++l) Cx::mac(acc, w.W[3 * r + l], nodec(P, w, m, k - l));
    w.u[q] = acc;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < P.nseg; s += blockDim.x) {
    cplx<K> acc = Cx::zero();
    for (int t = P.seg_ptr[s]; t < P.seg_ptr[s + 1]; ++t)
      Cx::mac(acc, load_coef<K>(P.term_coef, P.nterm, t), w.u[P.term_pair[t]]);
    w.part[s] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P.n; i += blockDim.x) {
    cplx<K> acc = w.part[P.eq_seg[i]];
    for (int s = P.eq_seg[i] + 1; s < P.eq_seg[i + 1]; ++s) acc = Cx::add(acc, w.part[s]);
    w.rhs[i] = acc;
  }
  __syncthreads();
}

// Jacobian A_ij = sum c * e * W_r0 * node(m/x_j)_0
template <int K> __device__ void jacobian(const DevProgram &P, Work<K> &w) {
  typedef C<K> Cx;
  for (int q = threadIdx.x; q < P.njpair; q += blockDim.x)
    w.uj[q] = Cx::mul(w.W[3 * P.jpair_comp[q]], nodec(P, w, P.jpair_node[q], 0));
  __syncthreads();
  const int nn = P.n * P.n;
  for (int e = threadIdx.x; e < nn; e += blockDim.x) {
    cplx<K> acc = Cx::zero();
    for (int t = P.jent_ptr[e]; t < P.jent_ptr[e + 1]; ++t)
      Cx::mac(acc, load_coef<K>(P.jterm_coef, P.njterm, t), w.uj[P.jterm_pair[t]]);
    w.A[e] = acc;
  }
  __syncthreads();
}

// in-place LU with partial pivoting; returns false if a pivot is zero
template <int K> __device__ bool lu_factor(const DevProgram &P, Work<K> &w) {
  typedef C<K> Cx;
  const int n = P.n;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double amax = absz<K>(w.A[j * n + j]);
    for (int i = j + 1; i < n; ++i) {
      double v = absz<K>(w.A[i * n + j]);
      if (v > amax) { amax = v; p = i; }
    }
    if (!(amax > 0.0)) return false;
    __syncthreads();
    if (p != j)
      for (int c = threadIdx.x; c < n; c += blockDim.x) {
        cplx<K> tmp = w.A[j * n + c];
        w.A[j * n + c] = w.A[p * n + c];
        w.A[p * n + c] = tmp;
      }
    if (threadIdx.x == 0) w.piv[j] = p;
    __syncthreads();
    const cplx<K> inv = Cx::div(Cx::from(1.0, 0.0), w.A[j * n + j]);
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) w.A[i * n + j] = Cx::mul(w.A[i * n + j], inv);
    if (threadIdx.x == 0) w.invd[j] = inv;
    __syncthreads();
    const int m = n - j - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      Cx::msub(w.A[i * n + c], w.A[i * n + j], w.A[j * n + c]);
    }
    __syncthreads();
  }
  return true;
}

// rhs := A^{-1} rhs with the factors of lu_factor
template <int K> __device__ void lu_solve(const DevProgram &P, Work<K> &w) {
  typedef C<K> Cx;
  const int n = P.n;
  if (threadIdx.x == 0)
    for (int j = 0; j < n; ++j) {
      const int p = w.piv[j];
      if (p != j) { cplx<K> t = w.rhs[j]; w.rhs[j] = w.rhs[p]; w.rhs[p] = t; }
    }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) Cx::msub(w.rhs[i], w.A[i * n + j], w.rhs[j]);
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; --j) {
    if (threadIdx.x == 0) w.rhs[j] = Cx::mul(w.rhs[j], w.invd[j]);
    __syncthreads();
    for (int i = threadIdx.x; i < j; i += blockDim.x) Cx::msub(w.rhs[i], w.A[i * n + j], w.rhs[j]);
    __syncthreads();
  }
}

// value (and optionally Jacobian) at the point held in the order-0 slots
template <int K> __device__ void eval_point(const DevProgram &P, Work<K> &w, bool jac) {
  product_pass<K>(P, w, 0);
  residual<K>(P, w, 0);
  if (jac) jacobian<K>(P, w);
}

template <int K> __device__ double max_abs(const cplx<K> *v, int n) {
  double m = 0.0;
  for (int i = 0; i < n; ++i) m = fmax(m, absz<K>(v[i]));
  return m;
}

template <int K> __device__ double max_abs_x(const DevProgram &P, Work<K> &w) {
  double m = 0.0;
  for (int j = 0; j < P.n; ++j) m = fmax(m, absz<K>(nodec(P, w, j, 0)));
  return m;
}

// Newton steps at fixed t on the order-0 slots; returns status
// (1 converged, 0 not converged, -1 singular); *dxl = last max |dx|
template <int K>
__device__ int newton(const DevProgram &P, Work<K> &w, int iters, double tol, double *dxl,
                      int *count) {
  typedef C<K> Cx;
  for (int it = 0; it < iters; ++it) {
    ++*count;
    eval_point<K>(P, w, true);
    if (!lu_factor<K>(P, w)) return -1;
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) w.rhs[i] = Cx::neg(w.rhs[i]);
    __syncthreads();
    lu_solve<K>(P, w);
    const double dx = max_abs<K>(w.rhs, P.n);
    for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
      cplx<K> &x = nodec(P, w, j, 0);
      x = Cx::add(x, w.rhs[j]);
    }
    __syncthreads();
    *dxl = dx;
    if (dx <= tol * (1.0 + max_abs_x<K>(P, w))) return 1;
  }
  return 0;
}

template <int K>
__device__ void carve(Work<K> &w, const DevProgram &P, double *lds) {
  typedef cplx<K> Z;
  Z *z = reinterpret_cast<Z *>(lds);
  w.slots = z; z += P.nslots;
  w.u = z; z += P.npair;
  w.uj = z; z += P.njpair;
  w.A = z; z += P.n * P.n;
  w.part = z; z += P.nseg;
  w.rhs = z; z += P.n;
  w.x0 = z; z += P.n;
  w.W = z; z += 3 * P.ncomp;
  w.dnode = z; z += (P.nprod > 0 ? P.nprod : 1);
  w.invd = z; z += P.n;
  w.red = reinterpret_cast<double *>(z);
  w.piv = reinterpret_cast<int *>(w.red + 8);
}

template <int K> size_t lds_bytes(const DevProgram &P) {
  size_t z = (size_t)P.nslots + P.npair + P.njpair + P.n * P.n + P.nseg + 3 * P.n +
             3 * P.ncomp + (P.nprod > 0 ? P.nprod : 1);
  return z * sizeof(cplx<K>) + 8 * sizeof(double) + 16 * sizeof(int) + P.n * sizeof(int);
}

// ----------------------------------------------------------- tracking ------

template <int K>
__global__ void __launch_bounds__(256) track_kernel(DevProgram P, Params par, int npaths,
                                                   const double *__restrict__ start, int KS,
                                                   double *__restrict__ endp, double *tend,
                                                   int *status, int *nsteps_out, int *nrej_out,
                                                   double *dx_out, int *work, int *counter) {
  typedef C<K> Cx;
  extern __shared__ double lds[];
  __shared__ int s_path;
  Work<K> w;
  carve<K>(w, P, lds);
  const int n = P.n, D = P.D;
  for (;;) {
    if (threadIdx.x == 0) s_path = atomicAdd(counter, 1);
    __syncthreads();
    const int path = s_path;
    __syncthreads();
    if (path >= npaths) break;
    // start point: KS limbs -> K limbs (extra limbs zero)
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      cplx<K> z = Cx::zero();
      for (int i = 0; i < K && i < KS; ++i) {
        reinterpret_cast<double *>(&z.re)[i] = start[(size_t)i * npaths * n + (size_t)path * n + j];
        reinterpret_cast<double *>(&z.im)[i] = start[(size_t)(KS + i) * npaths * n + (size_t)path * n + j];
      }
      nodec(P, w, j, 0) = z;
    }
    // the constant node ONE = 1 + 0 s + ...
    for (int a = threadIdx.x; a < D; a += blockDim.x)
      w.slots[P.off[n] + a] = (a == 0) ? Cx::from(1.0, 0.0) : Cx::zero();
    __syncthreads();
    double t = 0.0, h = par.maxstep, dxl = 0.0;
    int st = ST_RUNNING, steps = 0, rej = 0, npred = 0, nnewt = 0, nfeval = 0;
    while (st == ST_RUNNING && t < 1.0) {
      // ---- predictor: Taylor series of x(t+s) ----
      set_weights<K>(P, w, t);
      __syncthreads();
      ++npred;
      eval_point<K>(P, w, true);
      if (!lu_factor<K>(P, w)) { st = ST_SINGULAR; break; }
      for (int k = 1; k < D; ++k) {
        for (int j = threadIdx.x; j < n; j += blockDim.x) nodec(P, w, j, k) = Cx::zero();
        __syncthreads();
        product_pass<K>(P, w, k);
        residual<K>(P, w, k);
        for (int i = threadIdx.x; i < n; i += blockDim.x) w.rhs[i] = Cx::neg(w.rhs[i]);
        __syncthreads();
        lu_solve<K>(P, w);
        for (int j = threadIdx.x; j < n; j += blockDim.x) nodec(P, w, j, k) = w.rhs[j];
        __syncthreads();
        correction_pass<K>(P, w, k);
      }
      // ---- pole step from the Fabry ratio (offset par.fabry_offset) ----
      const int kd = D - 1 - par.fabry_offset;
      double rad = -1.0;
      if (kd >= 1)
        for (int j = 0; j < n; ++j) {
          const double den = absz<K>(nodec(P, w, j, kd));
          if (den > 0.0) {
            const double r = absz<K>(nodec(P, w, j, kd - 1)) / den;
            if (rad < 0.0 || r < rad) rad = r;
          }
        }
      h = (rad > 0.0) ? fmin(par.pbeta * rad, par.maxstep) : par.maxstep;
      for (int j = threadIdx.x; j < n; j += blockDim.x) w.x0[j] = nodec(P, w, j, 0);
      __syncthreads();
      // ---- feedback and corrector; halve the step on failure ----
      bool accepted = false;
      while (!accepted) {
        double tn = t + h;
        if (tn >= 1.0 - par.minstep) tn = 1.0;
        const double hs = tn - t;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
          cplx<K> acc = nodec(P, w, j, D - 1);
          for (int k = D - 2; k >= 1; --k) acc = Cx::add(Cx::scale_d(acc, hs), nodec(P, w, j, k));
          nodec(P, w, j, 0) = Cx::add(w.x0[j], Cx::scale_d(acc, hs));
        }
        set_weights<K>(P, w, tn);
        __syncthreads();
        ++nfeval;
        eval_point<K>(P, w, false);
        const double res = max_abs<K>(w.rhs, n);
        bool ok = res <= par.alpha * (1.0 + max_abs_x<K>(P, w));
        if (ok) {
          const int nr = newton<K>(P, w, par.corsteps, par.tolcor, &dxl, &nnewt);
          if (nr < 0) { st = ST_SINGULAR; break; }
          ok = (nr == 1);
        }
        if (ok) { t = tn; accepted = true; }
        else {
          ++rej;
          h *= 0.5;
          if (h < par.minstep) { st = ST_MINSTEP; break; }
        }
      }
      if (st != ST_RUNNING) break;
      ++steps;
      if (max_abs_x<K>(P, w) > par.divergence) { st = ST_DIVERGED; break; }
      if (steps >= par.maxsteps && t < 1.0) { st = ST_MAXSTEPS; break; }
    }
    if (st == ST_RUNNING) {
      // endpoint: Newton at t = 1
      set_weights<K>(P, w, 1.0);
      __syncthreads();
      const int nr = newton<K>(P, w, par.endsteps, par.tolend, &dxl, &nnewt);
      st = (nr < 0) ? ST_SINGULAR : ST_SUCCESS;
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const cplx<K> x = nodec(P, w, j, 0);
      for (int i = 0; i < K; ++i) {
        endp[(size_t)i * npaths * n + (size_t)path * n + j] = reinterpret_cast<const double *>(&x.re)[i];
        endp[(size_t)(K + i) * npaths * n + (size_t)path * n + j] = reinterpret_cast<const double *>(&x.im)[i];
      }
    }
    if (threadIdx.x == 0) {
      tend[path] = t;
      status[path] = st;
      nsteps_out[path] = steps;
      nrej_out[path] = rej;
      dx_out[path] = dxl;
      if (work) {
        work[path] = npred;
        work[npaths + path] = nnewt;
        work[2 * npaths + path] = nfeval;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------- endpoint refinement / residual --

// Newton at t = tval in K-limb precision from KS-limb points; writes the
// refined point, the last |dx| and the max residual |H(x,t)| after the last
// update (evaluated in the same precision).
template <int K>
__global__ void __launch_bounds__(256) refine_kernel(DevProgram P, int npaths, double tval,
                                                    const double *__restrict__ xin, int KS,
                                                    double *__restrict__ xout, int iters,
                                                    double tol, double *dx_out, double *res_out,
                                                    int *status, int *counter) {
  typedef C<K> Cx;
  extern __shared__ double lds[];
  __shared__ int s_path;
  Work<K> w;
  carve<K>(w, P, lds);
  const int n = P.n;
  for (;;) {
    if (threadIdx.x == 0) s_path = atomicAdd(counter, 1);
    __syncthreads();
    const int path = s_path;
    __syncthreads();
    if (path >= npaths) break;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      cplx<K> z = Cx::zero();
      for (int i = 0; i < K && i < KS; ++i) {
        reinterpret_cast<double *>(&z.re)[i] = xin[(size_t)i * npaths * n + (size_t)path * n + j];
        reinterpret_cast<double *>(&z.im)[i] = xin[(size_t)(KS + i) * npaths * n + (size_t)path * n + j];
      }
      nodec(P, w, j, 0) = z;
    }
    if (threadIdx.x == 0) w.slots[P.off[n]] = Cx::from(1.0, 0.0);
    set_weights<K>(P, w, tval);
    __syncthreads();
    double dxl = 0.0;
    int nr = 1;
    int its = 0;
    if (iters > 0) nr = newton<K>(P, w, iters, tol, &dxl, &its);
    eval_point<K>(P, w, false);
    const double res = max_abs<K>(w.rhs, n);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const cplx<K> x = nodec(P, w, j, 0);
      for (int i = 0; i < K; ++i) {
        xout[(size_t)i * npaths * n + (size_t)path * n + j] = reinterpret_cast<const double *>(&x.re)[i];
        xout[(size_t)(K + i) * npaths * n + (size_t)path * n + j] = reinterpret_cast<const double *>(&x.im)[i];
      }
    }
    if (threadIdx.x == 0) {
      dx_out[path] = dxl;
      res_out[path] = res;
      status[path] = nr;
    }
    __syncthreads();
  }
}

}  // namespace

// ================================================================ C-ABI ====

struct phc_tracker {
  DevProgram P;
  std::vector<void *> bufs;
  int *counter = nullptr;
};

namespace {

template <typename T> int upload(phc_tracker *tr, const T *src, size_t count, const T **dst) {
  void *p = nullptr;
  size_t bytes = sizeof(T) * (count > 0 ? count : 1);
  PHC_HIP(hipMalloc(&p, bytes));
  tr->bufs.push_back(p);
  if (count > 0) PHC_HIP(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = (const T *)p;
  return 0;
}

size_t lds_for(int K, const DevProgram &P) {
  if (K == 1) return lds_bytes<1>(P);
  if (K == 2) return lds_bytes<2>(P);
  if (K == 4) return lds_bytes<4>(P);
  return lds_bytes<8>(P);
}

int occupancy_grid(const void *fn, int block, size_t lds, int npaths) {
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, lds) != hipSuccess) return -1;
  if (per < 1) per = 1;
  long g = (long)ncu * per;
  if (g > npaths) g = npaths;
  return (int)(g > 0 ? g : 1);
}

}  // namespace

extern "C" {

int phc_program_validate(const phc_program *d) {
  PHC_REQUIRE(d != nullptr, "phc_program_validate: null program");
  PHC_REQUIRE(d->n >= 1 && d->n <= 64, "phc_program: n must be in 1..64");
  PHC_REQUIRE(d->deg >= 0 && d->deg <= 32, "phc_program: deg must be in 0..32");
  PHC_REQUIRE(d->wdeg >= 0 && d->wdeg <= 2, "phc_program: weight degree must be <= 2");
  PHC_REQUIRE(d->ncomp >= 1 && d->nprod >= 0 && d->nlev >= 0 && d->npair >= 0 && d->njpair >= 0,
              "phc_program: negative count");
  PHC_REQUIRE(d->nseg >= d->n && d->nterm >= 0 && d->njterm >= 0, "phc_program: bad counts");
  const int nn = d->n + 1 + d->nprod;
  const int need = d->wdeg < d->deg ? d->wdeg : d->deg;  // orders k-l kept besides k
  for (int i = 0; i < d->nprod; ++i)
    PHC_REQUIRE(d->left[i] >= 0 && d->left[i] < d->n + 1 + i && d->left[i] != d->n &&
                d->right[i] >= 0 && d->right[i] < d->n,
                "phc_program: product node factors must be earlier nodes, right a variable");
  for (int i = 0; i < nn; ++i)
    PHC_REQUIRE(d->off[i] >= 0 && d->slen[i] >= 1 && d->off[i] + d->slen[i] <= d->nslots,
                "phc_program: node storage out of range");
  for (int i = 0; i <= d->n; ++i)
    PHC_REQUIRE(d->slen[i] == d->deg + 1, "phc_program: variables need full series storage");
  for (int i = 0; i < d->nprod; ++i)
    PHC_REQUIRE(d->slen[d->left[i]] == d->deg + 1, "phc_program: factor nodes need full series storage");
  for (int i = 0; i < d->npair; ++i)
    PHC_REQUIRE(d->pair_comp[i] >= 0 && d->pair_comp[i] < d->ncomp && d->pair_node[i] >= 0 &&
                d->pair_node[i] < nn && d->slen[d->pair_node[i]] > need,
                "phc_program: bad component pair");
  for (int i = 0; i < d->njpair; ++i)
    PHC_REQUIRE(d->jpair_comp[i] >= 0 && d->jpair_comp[i] < d->ncomp && d->jpair_node[i] >= 0 &&
                d->jpair_node[i] < nn, "phc_program: bad Jacobian pair");
  PHC_REQUIRE(d->seg_ptr[0] == 0 && d->seg_ptr[d->nseg] == d->nterm && d->eq_seg[0] == 0 &&
              d->eq_seg[d->n] == d->nseg, "phc_program: bad residual segments");
  for (int s = 0; s < d->nseg; ++s)
    PHC_REQUIRE(d->seg_ptr[s + 1] >= d->seg_ptr[s], "phc_program: segments not increasing");
  for (int i = 0; i < d->n; ++i)
    PHC_REQUIRE(d->eq_seg[i + 1] > d->eq_seg[i], "phc_program: equation without segments");
  for (int i = 0; i < d->nterm; ++i)
    PHC_REQUIRE(d->term_pair[i] >= 0 && d->term_pair[i] < d->npair, "phc_program: bad term");
  PHC_REQUIRE(d->jent_ptr[0] == 0 && d->jent_ptr[d->n * d->n] == d->njterm,
              "phc_program: bad Jacobian entries");
  for (int e = 0; e < d->n * d->n; ++e)
    PHC_REQUIRE(d->jent_ptr[e + 1] >= d->jent_ptr[e], "phc_program: Jacobian entries not increasing");
  for (int i = 0; i < d->njterm; ++i)
    PHC_REQUIRE(d->jterm_pair[i] >= 0 && d->jterm_pair[i] < d->njpair,
                "phc_program: bad Jacobian term");
  PHC_REQUIRE(d->lev_ptr[0] == 0 && d->lev_ptr[d->nlev] == d->nprod, "phc_program: bad levels");
  for (int l = 1; l <= d->nlev; ++l) {
    PHC_REQUIRE(d->lev_ptr[l] >= d->lev_ptr[l - 1], "phc_program: bad levels");
    // a product's left factor lies on the previous level (variables on level 0)
    for (int i = d->lev_ptr[l - 1]; i < d->lev_ptr[l]; ++i) {
      const int a = d->left[i];
      const int la = (a < d->n) ? 0 : -1;
      if (la == 0) PHC_REQUIRE(l == 1, "phc_program: level of a product of two variables must be 1");
      else
        PHC_REQUIRE(a - d->n - 1 >= (l >= 2 ? d->lev_ptr[l - 2] : 0) && a - d->n - 1 < d->lev_ptr[l - 1],
                    "phc_program: left factor not on the previous level");
    }
  }
  return 0;
}

int phc_tracker_create(phc_tracker **out, const phc_program *d) {
  PHC_REQUIRE(out && d, "phc_tracker_create: null argument");
  *out = nullptr;
  if (phc_program_validate(d) != 0) return -2;
  const int nn = d->n + 1 + d->nprod;
  phc_tracker *tr = new phc_tracker();
  DevProgram &P = tr->P;
  P.n = d->n; P.D = d->deg + 1; P.nprod = d->nprod; P.nlev = d->nlev; P.nslots = d->nslots;
  P.ring = 0; P.wdeg = d->wdeg; P.ncomp = d->ncomp; P.npair = d->npair; P.nseg = d->nseg;
  P.nterm = d->nterm; P.njpair = d->njpair; P.njterm = d->njterm;
  int rc = 0;
  rc |= upload<int>(tr, d->left, d->nprod, &P.left);
  rc |= upload<int>(tr, d->right, d->nprod, &P.right);
  rc |= upload<int>(tr, d->lev_ptr, d->nlev + 1, &P.lev_ptr);
  rc |= upload<int>(tr, d->off, nn, &P.off);
  rc |= upload<int>(tr, d->slen, nn, &P.slen);
  rc |= upload<int>(tr, d->pair_comp, d->npair, &P.pair_comp);
  rc |= upload<int>(tr, d->pair_node, d->npair, &P.pair_node);
  rc |= upload<int>(tr, d->seg_ptr, d->nseg + 1, &P.seg_ptr);
  rc |= upload<int>(tr, d->eq_seg, d->n + 1, &P.eq_seg);
  rc |= upload<int>(tr, d->term_pair, d->nterm, &P.term_pair);
  rc |= upload<double>(tr, d->term_coef, (size_t)16 * d->nterm, &P.term_coef);
  rc |= upload<int>(tr, d->jpair_comp, d->njpair, &P.jpair_comp);
  rc |= upload<int>(tr, d->jpair_node, d->njpair, &P.jpair_node);
  rc |= upload<int>(tr, d->jent_ptr, (size_t)d->n * d->n + 1, &P.jent_ptr);
  rc |= upload<int>(tr, d->jterm_pair, d->njterm, &P.jterm_pair);
  rc |= upload<double>(tr, d->jterm_coef, (size_t)16 * d->njterm, &P.jterm_coef);
  rc |= upload<double>(tr, d->weights, (size_t)16 * 3 * d->ncomp, &P.weights);
  void *c = nullptr;
  if (rc != 0 || hipMalloc(&c, sizeof(int)) != hipSuccess) {
    phc_tracker_destroy(tr);
    if (rc == 0) phc::set_error("phc_tracker_create: device allocation failed");
    return -1;
  }
  tr->counter = (int *)c;
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

int phc_tracker_lds_bytes(const phc_tracker *tr, int K, size_t *bytes) {
  PHC_REQUIRE(tr && bytes, "phc_tracker_lds_bytes: null argument");
  PHC_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, "phc_tracker_lds_bytes: K must be 1, 2, 4, 8");
  *bytes = lds_for(K, tr->P);
  return 0;
}

int phc_track(phc_tracker *tr, int K, int npaths, const double *start, int KS, double *endp,
              double *tend, int *status, int *nsteps, int *nrej, double *dxlast, int *work,
              const phc_track_params *prm, int block, void *stream) {
  PHC_REQUIRE(tr && prm, "phc_track: null argument");
  PHC_REQUIRE(K == 1 || K == 2, "phc_track: tracking precision K must be 1 or 2");
  PHC_REQUIRE(KS >= 1 && KS <= 8, "phc_track: start limbs KS must be in 1..8");
  PHC_REQUIRE(npaths >= 0, "phc_track: npaths < 0");
  PHC_REQUIRE(block == 64 || block == 128 || block == 256, "phc_track: block must be 64, 128 or 256");
  PHC_REQUIRE(tr->P.D >= 2, "phc_track: the tracker program needs series degree >= 1");
  PHC_REQUIRE(prm->fabry_offset >= 0 && prm->fabry_offset < tr->P.D - 1,
              "phc_track: fabry_offset must be below the series degree");
  PHC_REQUIRE(prm->maxstep > 0 && prm->minstep > 0 && prm->maxsteps > 0 && prm->corsteps > 0,
              "phc_track: step parameters must be positive");
  if (npaths == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = lds_for(K, tr->P);
  PHC_REQUIRE(lds <= 160 * 1024, "phc_track: program too large for LDS");
  Params par;
  par.maxsteps = prm->maxsteps; par.corsteps = prm->corsteps; par.endsteps = prm->endsteps;
  par.fabry_offset = prm->fabry_offset; par.maxstep = prm->maxstep; par.minstep = prm->minstep;
  par.pbeta = prm->pbeta; par.alpha = prm->alpha; par.tolcor = prm->tolcor;
  par.tolend = prm->tolend; par.divergence = prm->divergence;
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  const void *fn = (K == 1) ? (const void *)track_kernel<1> : (const void *)track_kernel<2>;
  if (lds > 64 * 1024) PHC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = occupancy_grid(fn, block, lds, npaths);
  PHC_REQUIRE(grid > 0, "phc_track: occupancy query failed");
  if (K == 1)
    hipLaunchKernelGGL(track_kernel<1>, dim3(grid), dim3(block), lds, st, tr->P, par, npaths, start,
                       KS, endp, tend, status, nsteps, nrej, dxlast, work, tr->counter);
  else
    hipLaunchKernelGGL(track_kernel<2>, dim3(grid), dim3(block), lds, st, tr->P, par, npaths, start,
                       KS, endp, tend, status, nsteps, nrej, dxlast, work, tr->counter);
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_refine(phc_tracker *tr, int K, int npaths, double t, const double *xin, int KS,
               double *xout, int iters, double tol, double *dxlast, double *res, int *status,
               int block, void *stream) {
  PHC_REQUIRE(tr, "phc_refine: null tracker");
  PHC_REQUIRE(K == 1 || K == 2 || K == 4 || K == 8, "phc_refine: K must be 1, 2, 4 or 8");
  PHC_REQUIRE(KS >= 1 && KS <= 8, "phc_refine: KS must be in 1..8");
  PHC_REQUIRE(npaths >= 0 && iters >= 0, "phc_refine: negative size");
  PHC_REQUIRE(block == 64 || block == 128 || block == 256, "phc_refine: block must be 64, 128 or 256");
  if (npaths == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = lds_for(K, tr->P);
  PHC_REQUIRE(lds <= 160 * 1024, "phc_refine: program too large for LDS");
  PHC_HIP(hipMemsetAsync(tr->counter, 0, sizeof(int), st));
  const void *fn = K == 1 ? (const void *)refine_kernel<1> : K == 2 ? (const void *)refine_kernel<2>
                 : K == 4 ? (const void *)refine_kernel<4> : (const void *)refine_kernel<8>;
  if (lds > 64 * 1024) PHC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = occupancy_grid(fn, block, lds, npaths);
  PHC_REQUIRE(grid > 0, "phc_refine: occupancy query failed");
#define RK(KK) hipLaunchKernelGGL(refine_kernel<KK>, dim3(grid), dim3(block), lds, st, tr->P, npaths, t, xin, KS, xout, iters, tol, dxlast, res, status, tr->counter)
  if (K == 1) RK(1); else if (K == 2) RK(2); else if (K == 4) RK(4); else RK(8);
#undef RK
  PHC_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
