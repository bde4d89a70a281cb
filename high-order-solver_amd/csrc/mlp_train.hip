// mlp_train.hip — training of the troubled-interface classifier on the GPU
// (FP32): per-sample forward and backward pass, a deterministic reduction of
// the parameter gradients over the mini-batch, and Adam.  Layout of the
// parameter vector as in include/gks_learn.h.
#include <hip/hip_runtime.h>

#include "mlp_train.hpp"

namespace mlp {

constexpr int NIN = 16, NH = 32;
constexpr int OW1 = 0, OB1 = 512, OW2 = 544, OB2 = 1568, OW3 = 1600, OB3 = 1632, NP = 1633;
constexpr int kB = 256;

// per sample: activations h1, h2 and the back-propagated pre-activation
// gradients g1, g2 (32 each) and the logit gradient gz, stored SoA
// [k][batch] for coalesced reads in the reduction; the loss term goes to L
__global__ void __launch_bounds__(kB) k_fwd_bwd(const float* __restrict__ P, int nb, const int* __restrict__ idx,
                                                const float* __restrict__ X, const float* __restrict__ Y,
                                                const float* __restrict__ W, float inv_b, float* __restrict__ H1,
                                                float* __restrict__ H2, float* __restrict__ G1,
                                                float* __restrict__ G2, float* __restrict__ GZ,
                                                float* __restrict__ L) {
  const int b = blockIdx.x * kB + threadIdx.x;
  if (b >= nb) return;
  const long long s = idx ? idx[b] : b;
  float x[NIN], a1[NH], a2[NH];
#pragma unroll
  for (int k = 0; k < NIN; k++) x[k] = X[s * NIN + k];
#pragma unroll
  for (int j = 0; j < NH; j++) {
    float acc = P[OB1 + j];
#pragma unroll
    for (int k = 0; k < NIN; k++) acc += x[k] * P[OW1 + k * NH + j];
    a1[j] = acc;
  }
#pragma unroll
  for (int j = 0; j < NH; j++) {
    float acc = P[OB2 + j];
#pragma unroll
    for (int k = 0; k < NH; k++) acc += fmaxf(a1[k], 0.0f) * P[OW2 + k * NH + j];
    a2[j] = acc;
  }
  float z = P[OB3];
#pragma unroll
  for (int k = 0; k < NH; k++) z += fmaxf(a2[k], 0.0f) * P[OW3 + k];
  const float y = Y[s], w = W ? W[s] : 1.0f;
  // weighted binary cross-entropy with logits, averaged over the batch
  L[b] = w * (fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)))) * inv_b;
  const float gz = w * (1.0f / (1.0f + expf(-z)) - y) * inv_b;
  GZ[b] = gz;
  float g2[NH];
#pragma unroll
  for (int j = 0; j < NH; j++) {
    g2[j] = a2[j] > 0.0f ? gz * P[OW3 + j] : 0.0f;
    H2[(long long)j * nb + b] = fmaxf(a2[j], 0.0f);
    G2[(long long)j * nb + b] = g2[j];
  }
#pragma unroll
  for (int k = 0; k < NH; k++) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < NH; j++) acc += g2[j] * P[OW2 + k * NH + j];
    H1[(long long)k * nb + b] = fmaxf(a1[k], 0.0f);
    G1[(long long)k * nb + b] = a1[k] > 0.0f ? acc : 0.0f;
  }
}

// the fixed tree over a workgroup's kB partial sums (lane 0 ends with the
// total); the pairing depends only on kB
__device__ __forceinline__ double block_sum(double acc, double* s_acc) {
  const int t = threadIdx.x;
  s_acc[t] = acc;
  __syncthreads();
  for (int w = kB / 2; w > 0; w >>= 1) {
    if (t < w) s_acc[t] += s_acc[t + w];
    __syncthreads();
  }
  return s_acc[0];
}

// gradient of parameter p = blockIdx.x over the mini-batch: lane t sums the
// samples t, t + kB, ... in order, then the fixed tree over the lanes, in
// double; the summation order depends only on nb, so training is
// reproducible run to run
__global__ void __launch_bounds__(kB) k_reduce(int nb, const int* __restrict__ idx, const float* __restrict__ X,
                                               const float* __restrict__ H1, const float* __restrict__ H2,
                                               const float* __restrict__ G1, const float* __restrict__ G2,
                                               const float* __restrict__ GZ, float* __restrict__ grad) {
  __shared__ double s_acc[kB];
  const int p = blockIdx.x;
  double acc = 0.0;
  if (p < OB1) {  // W1[k][j] = sum x_k g1_j
    const int k = p / NH, j = p % NH;
    for (int b = threadIdx.x; b < nb; b += kB) {
      const long long s = idx ? idx[b] : b;
      acc += (double)X[s * NIN + k] * G1[(long long)j * nb + b];
    }
  } else if (p < OW2) {
    const int j = p - OB1;
    for (int b = threadIdx.x; b < nb; b += kB) acc += G1[(long long)j * nb + b];
  } else if (p < OB2) {  // W2[k][j] = sum h1_k g2_j
    const int k = (p - OW2) / NH, j = (p - OW2) % NH;
    for (int b = threadIdx.x; b < nb; b += kB) acc += (double)H1[(long long)k * nb + b] * G2[(long long)j * nb + b];
  } else if (p < OW3) {
    const int j = p - OB2;
    for (int b = threadIdx.x; b < nb; b += kB) acc += G2[(long long)j * nb + b];
  } else if (p < OB3) {  // W3[k] = sum h2_k gz
    const int k = p - OW3;
    for (int b = threadIdx.x; b < nb; b += kB) acc += (double)H2[(long long)k * nb + b] * GZ[b];
  } else {
    for (int b = threadIdx.x; b < nb; b += kB) acc += GZ[b];
  }
  const double tot = block_sum(acc, s_acc);
  if (threadIdx.x == 0) grad[p] = (float)tot;
}

// the batch loss: one workgroup, the same fixed-order scheme
__global__ void __launch_bounds__(kB) k_sum(int n, const float* __restrict__ v, double* __restrict__ out) {
  __shared__ double s_acc[kB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kB) acc += v[i];
  const double tot = block_sum(acc, s_acc);
  if (threadIdx.x == 0) *out = tot;
}

__global__ void __launch_bounds__(kB) k_adam(float* __restrict__ P, const float* __restrict__ g,
                                             float* __restrict__ m, float* __restrict__ v, float lr, float bc1,
                                             float bc2) {
  const int p = blockIdx.x * kB + threadIdx.x;
  if (p >= NP) return;
  const float gp = g[p];
  const float mp = 0.9f * m[p] + 0.1f * gp;
  const float vp = 0.999f * v[p] + 0.001f * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + 1e-8f);
}

hipError_t fwd_bwd_reduce(const Work& w, const float* P, int nb, const int* idx, const float* X, const float* Y,
                          const float* W, float* grad, hipStream_t s) {
  const float inv_b = 1.0f / (float)nb;
  hipLaunchKernelGGL(k_fwd_bwd, dim3((nb + kB - 1) / kB), dim3(kB), 0, s, P, nb, idx, X, Y, W, inv_b, w.h1, w.h2,
                     w.g1, w.g2, w.gz, w.loss);
  hipLaunchKernelGGL(k_reduce, dim3(NP), dim3(kB), 0, s, nb, idx, X, w.h1, w.h2, w.g1, w.g2, w.gz, grad);
  return hipGetLastError();
}
hipError_t loss_sum(const Work& w, int nb, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kB), 0, s, nb, (const float*)w.loss, out);
  return hipGetLastError();
}
hipError_t adam(float* P, const float* g, float* m, float* v, float lr, float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((NP + kB - 1) / kB), dim3(kB), 0, s, P, g, m, v, lr, bc1, bc2);
  return hipGetLastError();
}

}  // namespace mlp
