// mlp_train.hpp — host interface of the classifier's training kernels
// (mlp_train.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace mlp {

// device scratch for one mini-batch of up to `cap` samples
struct Work {
  int cap = 0;
  float* h1 = nullptr;  // [32][cap]
  float* h2 = nullptr;
  float* g1 = nullptr;
  float* g2 = nullptr;
  float* gz = nullptr;    // [cap]
  float* loss = nullptr;  // [cap]
};

// per-sample forward/backward of nb samples (rows idx[b] of X, or b when
// idx is null) and the batch-mean gradient of the weighted BCE in grad
hipError_t fwd_bwd_reduce(const Work& w, const float* P, int nb, const int* idx, const float* X, const float* Y,
                          const float* W, float* grad, hipStream_t s);
// sum of the per-sample loss terms of the last fwd_bwd_reduce
hipError_t loss_sum(const Work& w, int nb, double* out, hipStream_t s);
// one Adam step (bias corrections bc1 = 1 - 0.9^t, bc2 = 1 - 0.999^t)
hipError_t adam(float* P, const float* g, float* m, float* v, float lr, float bc1, float bc2, hipStream_t s);

}  // namespace mlp
