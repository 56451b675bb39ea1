// Small FP32 multilayer perceptron for the learned step-size controller:
// F inputs -> H (ReLU) -> H (ReLU) -> C logits, trained with minibatch Adam
// on the tracker's own step logs.  Forward and backward passes are tiled
// FP32 matrix products staged through LDS plus fused elementwise kernels.
//
// The reference has no learned component; the step rule it keeps as judge
// and fallback is the pole step of LU_Prediction
// (Deformations/Curves/standard_predictor_convolutions.adb:1142-1185).
//
// Parameter vector layout (floats): W1[F][H], b1[H], W2[H][H], b2[H],
// W3[H][C], b3[C].  Loss: mean over the batch of the binary cross-entropy of
// the logit of the candidate that was tried, sigmoid(o[cand]) vs accepted.
#include <hip/hip_runtime.h>
#include <vector>
#include "common.hpp"
#include "../../include/phcmd.h"

namespace {

constexpr int TILE = 16;

// C[M][N] = op(A)[M][Kd] * op(B)[Kd][N] (+ C when accumulate)
// op(A)[m][k] = ta ? A[k*lda+m] : A[m*lda+k]; op(B)[k][n] = tb ? B[n*ldb+k] : B[k*ldb+n]
__global__ void gemm_kernel(int M, int N, int Kd, const float *__restrict__ A, int lda, int ta,
                            const float *__restrict__ B, int ldb, int tb, float *__restrict__ C,
                            int ldc, int accumulate) {
  __shared__ float As[TILE][TILE + 1], Bs[TILE][TILE + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int row = blockIdx.y * TILE + ty, col = blockIdx.x * TILE + tx;
  float acc = 0.0f;
  for (int k0 = 0; k0 < Kd; k0 += TILE) {
    const int ka = k0 + tx, kb = k0 + ty;
    float a = 0.0f, b = 0.0f;
    if (row < M && ka < Kd) a = ta ? A[(size_t)ka * lda + row] : A[(size_t)row * lda + ka];
    if (col < N && kb < Kd) b = tb ? B[(size_t)col * ldb + kb] : B[(size_t)kb * ldb + col];
    As[ty][tx] = a;
    Bs[ty][tx] = b;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TILE; ++k) acc += As[ty][k] * Bs[k][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    float *c = C + (size_t)row * ldc + col;
    *c = accumulate ? *c + acc : acc;
  }
}

// H = relu(Z + b) (relu optional), Z, H [M][N]
__global__ void bias_act_kernel(int M, int N, const float *Z, const float *b, float *H, int relu) {
  const size_t n = (size_t)M * N;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float v = Z[i] + b[i % N];
    H[i] = (relu && v < 0.0f) ? 0.0f : v;
  }
}

// dO[i][c] = (c == cand_i) ? (sigmoid(O[i][c]) - y_i) * scale : 0; loss += scale * BCE
__global__ void loss_grad_kernel(int B, int C, const float *O, const int *cand, const float *y,
                                 float scale, float *dO, float *loss) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B; i += gridDim.x * blockDim.x) {
    const int c = cand[i];
    for (int j = 0; j < C; ++j) dO[(size_t)i * C + j] = 0.0f;
    if (c < 0 || c >= C) continue;
    const float o = O[(size_t)i * C + c];
    const float p = 1.0f / (1.0f + expf(-o));
    dO[(size_t)i * C + c] = (p - y[i]) * scale;
    // log(1 + e^o) - y*o, computed stably
    const float l = fmaxf(o, 0.0f) - o * y[i] + log1pf(expf(-fabsf(o)));
    if (loss) atomicAdd(loss, l * scale);
  }
}

// dZ = dH * (H > 0)
__global__ void relu_back_kernel(size_t n, const float *H, const float *dH, float *dZ) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dZ[i] = H[i] > 0.0f ? dH[i] : 0.0f;
}

// db[j] = sum_i dZ[i][j]; one block per column
__global__ void colsum_kernel(int M, int N, const float *dZ, float *db) {
  __shared__ float part[256];
  const int j = blockIdx.x;
  float s = 0.0f;
  for (int i = threadIdx.x; i < M; i += blockDim.x) s += dZ[(size_t)i * N + j];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if (threadIdx.x < m) part[threadIdx.x] += part[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = part[0];
}

__global__ void adam_kernel(int n, float *w, const float *g, float *m, float *v, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i];
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    w[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
  }
}

struct Dims {
  int F, H, C;
  size_t nW1() const { return (size_t)F * H; }
  size_t oW1() const { return 0; }
  size_t ob1() const { return nW1(); }
  size_t oW2() const { return ob1() + H; }
  size_t ob2() const { return oW2() + (size_t)H * H; }
  size_t oW3() const { return ob2() + H; }
  size_t ob3() const { return oW3() + (size_t)H * C; }
  size_t total() const { return ob3() + C; }
};

void gemm(hipStream_t st, int M, int N, int Kd, const float *A, int lda, int ta, const float *B,
          int ldb, int tb, float *C, int ldc, int acc) {
  dim3 grid((N + TILE - 1) / TILE, (M + TILE - 1) / TILE), block(TILE, TILE);
  hipLaunchKernelGGL(gemm_kernel, grid, block, 0, st, M, N, Kd, A, lda, ta, B, ldb, tb, C, ldc, acc);
}

// forward on B samples; Z1,H1,Z2,H2 [B][H], O [B][C]
void forward(hipStream_t st, const Dims &d, const float *w, int B, const float *X, float *Z1,
             float *H1, float *Z2, float *H2, float *O) {
  const int g = phc::grid_for((size_t)B * d.H, 256);
  gemm(st, B, d.H, d.F, X, d.F, 0, w + d.oW1(), d.H, 0, Z1, d.H, 0);
  hipLaunchKernelGGL(bias_act_kernel, dim3(g), dim3(256), 0, st, B, d.H, Z1, w + d.ob1(), H1, 1);
  gemm(st, B, d.H, d.H, H1, d.H, 0, w + d.oW2(), d.H, 0, Z2, d.H, 0);
  hipLaunchKernelGGL(bias_act_kernel, dim3(g), dim3(256), 0, st, B, d.H, Z2, w + d.ob2(), H2, 1);
  gemm(st, B, d.C, d.H, H2, d.H, 0, w + d.oW3(), d.C, 0, O, d.C, 0);
  hipLaunchKernelGGL(bias_act_kernel, dim3(phc::grid_for((size_t)B * d.C, 256)), dim3(256), 0, st,
                     B, d.C, O, w + d.ob3(), O, 0);
}

// gradient of the mean loss over B samples into grad (same layout as w)
void backward(hipStream_t st, const Dims &d, const float *w, int B, const float *X, const int *cand,
              const float *y, const float *H1, const float *H2, const float *O, float *dO,
              float *dH, float *dZ, float *grad, float *loss) {
  const size_t nh = (size_t)B * d.H;
  hipLaunchKernelGGL(loss_grad_kernel, dim3(phc::grid_for(B, 256)), dim3(256), 0, st, B, d.C, O,
                     cand, y, 1.0f / (float)B, dO, loss);
  // layer 3: dW3 = H2^T dO, db3 = colsum dO, dH2 = dO W3^T
  gemm(st, d.H, d.C, B, H2, d.H, 1, dO, d.C, 0, grad + d.oW3(), d.C, 0);
  hipLaunchKernelGGL(colsum_kernel, dim3(d.C), dim3(256), 0, st, B, d.C, dO, grad + d.ob3());
  gemm(st, B, d.H, d.C, dO, d.C, 0, w + d.oW3(), d.C, 1, dH, d.H, 0);
  hipLaunchKernelGGL(relu_back_kernel, dim3(phc::grid_for(nh, 256)), dim3(256), 0, st, nh, H2, dH, dZ);
  // layer 2
  gemm(st, d.H, d.H, B, H1, d.H, 1, dZ, d.H, 0, grad + d.oW2(), d.H, 0);
  hipLaunchKernelGGL(colsum_kernel, dim3(d.H), dim3(256), 0, st, B, d.H, dZ, grad + d.ob2());
  gemm(st, B, d.H, d.H, dZ, d.H, 0, w + d.oW2(), d.H, 1, dH, d.H, 0);
  hipLaunchKernelGGL(relu_back_kernel, dim3(phc::grid_for(nh, 256)), dim3(256), 0, st, nh, H1, dH, dZ);
  // layer 1
  gemm(st, d.F, d.H, B, X, d.F, 1, dZ, d.H, 0, grad + d.oW1(), d.H, 0);
  hipLaunchKernelGGL(colsum_kernel, dim3(d.H), dim3(256), 0, st, B, d.H, dZ, grad + d.ob1());
}

struct Scratch {
  float *Z1 = nullptr, *H1, *Z2, *H2, *O, *dO, *dH, *dZ;
  int alloc(int B, const Dims &d) {
    const size_t nh = (size_t)B * d.H, nc = (size_t)B * d.C;
    float *base = nullptr;
    PHC_HIP(hipMalloc(&base, sizeof(float) * (6 * nh + 2 * nc)));
    Z1 = base; H1 = Z1 + nh; Z2 = H1 + nh; H2 = Z2 + nh; dH = H2 + nh; dZ = dH + nh;
    O = dZ + nh; dO = O + nc;
    return 0;
  }
  void release() { if (Z1) (void)hipFree(Z1); Z1 = nullptr; }
};

}  // namespace

extern "C" {

int phc_mlp_size(int F, int H, int C, int *nparams) {
  PHC_REQUIRE(nparams && F >= 1 && H >= 1 && C >= 1, "phc_mlp_size: bad sizes");
  *nparams = (int)Dims{F, H, C}.total();
  return 0;
}

int phc_mlp_forward(int F, int H, int C, const float *w, int B, const float *X, float *logits,
                    void *stream) {
  PHC_REQUIRE(F >= 1 && H >= 1 && C >= 1 && B >= 0, "phc_mlp_forward: bad sizes");
  if (B == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  Dims d{F, H, C};
  Scratch s;
  if (s.alloc(B, d) != 0) return -1;
  forward(st, d, w, B, X, s.Z1, s.H1, s.Z2, s.H2, s.O);
  PHC_HIP(hipMemcpyAsync(logits, s.O, sizeof(float) * B * C, hipMemcpyDeviceToDevice, st));
  PHC_HIP(hipStreamSynchronize(st));
  s.release();
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_mlp_grad(int F, int H, int C, const float *w, int B, const float *X, const int *cand,
                 const float *y, float *grad, float *loss, void *stream) {
  PHC_REQUIRE(F >= 1 && H >= 1 && C >= 1 && B >= 1, "phc_mlp_grad: bad sizes");
  hipStream_t st = (hipStream_t)stream;
  Dims d{F, H, C};
  Scratch s;
  if (s.alloc(B, d) != 0) return -1;
  if (loss) PHC_HIP(hipMemsetAsync(loss, 0, sizeof(float), st));
  forward(st, d, w, B, X, s.Z1, s.H1, s.Z2, s.H2, s.O);
  backward(st, d, w, B, X, cand, y, s.H1, s.H2, s.O, s.dO, s.dH, s.dZ, grad, loss);
  PHC_HIP(hipStreamSynchronize(st));
  s.release();
  PHC_HIP(hipGetLastError());
  return 0;
}

int phc_mlp_train(int F, int H, int C, float *w, int N, const float *X, const int *cand,
                  const float *y, int epochs, int batch, float lr, unsigned seed, float *losses,
                  void *stream) {
  PHC_REQUIRE(F >= 1 && H >= 1 && C >= 1 && N >= 1 && epochs >= 0 && batch >= 1,
              "phc_mlp_train: bad sizes");
  hipStream_t st = (hipStream_t)stream;
  Dims d{F, H, C};
  const int B = batch < N ? batch : N;
  const size_t np = d.total();
  Scratch s;
  if (s.alloc(B, d) != 0) return -1;
  float *opt = nullptr, *loss = nullptr;
  PHC_HIP(hipMalloc(&opt, sizeof(float) * (3 * np + 1)));
  float *grad = opt, *m = opt + np, *v = m + np;
  loss = v + np;
  PHC_HIP(hipMemsetAsync(opt, 0, sizeof(float) * (3 * np + 1), st));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.0f, c2 = 1.0f;
  // minibatches are consecutive slices of an epoch-dependent rotation of the
  // data (deterministic given the seed; the caller shuffles rows once)
  const int nb = N / B > 0 ? N / B : 1;
  for (int e = 0; e < epochs; ++e) {
    PHC_HIP(hipMemsetAsync(loss, 0, sizeof(float), st));
    const int rot = (int)((seed + 7919u * (unsigned)e) % (unsigned)nb);
    for (int q = 0; q < nb; ++q) {
      const size_t off = (size_t)((q + rot) % nb) * B;
      forward(st, d, w, B, X + off * F, s.Z1, s.H1, s.Z2, s.H2, s.O);
      backward(st, d, w, B, X + off * F, cand + off, y + off, s.H1, s.H2, s.O, s.dO, s.dH, s.dZ,
               grad, loss);
      c1 *= b1;
      c2 *= b2;
      hipLaunchKernelGGL(adam_kernel, dim3(phc::grid_for(np, 256)), dim3(256), 0, st, (int)np, w,
                         grad, m, v, lr, b1, b2, eps, 1.0f - c1, 1.0f - c2);
    }
    if (losses) {
      float l = 0.0f;
      PHC_HIP(hipMemcpyAsync(&l, loss, sizeof(float), hipMemcpyDeviceToHost, st));
      PHC_HIP(hipStreamSynchronize(st));
      losses[e] = l / (float)nb;
    }
  }
  PHC_HIP(hipStreamSynchronize(st));
  (void)hipFree(opt);
  s.release();
  PHC_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
