// learn_api.cpp — extern "C" training and inference entry points of the
// troubled-interface classifier (include/gks_learn.h); the solver-side hooks
// (samples, set_classifier, stats) live with the 2D context in gks2d_api.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../include/gks_learn.h"
#include "gks2d_launch.hpp"
#include "mlp_pack.hpp"
#include "mlp_train.hpp"

namespace {

thread_local std::string g_err_l;

int fail(int code, const char* what) {
  g_err_l = what;
  return code;
}

// device buffers freed on scope exit
struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T> hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 16);
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T*>(q);
    return e;
  }
};

hipError_t make_work(DevBufs& b, mlp::Work& w, int cap) {
  w.cap = cap;
  hipError_t e = b.alloc(&w.h1, (size_t)32 * cap);
  if (e == hipSuccess) e = b.alloc(&w.h2, (size_t)32 * cap);
  if (e == hipSuccess) e = b.alloc(&w.g1, (size_t)32 * cap);
  if (e == hipSuccess) e = b.alloc(&w.g2, (size_t)32 * cap);
  if (e == hipSuccess) e = b.alloc(&w.gz, (size_t)cap);
  if (e == hipSuccess) e = b.alloc(&w.loss, (size_t)cap);
  return e;
}

int hip_fail(hipError_t e, const char* where) {
  g_err_l = std::string(where) + ": " + hipGetErrorString(e);
  return GKS_ERR_HIP;
}
#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

}  // namespace

extern "C" {

const char* gks_mlp_last_error(void) { return g_err_l.c_str(); }

int gks_mlp_loss_grad(const float* params, int64_t n, const float* X, const float* y, const float* w, float* grad,
                      double* loss) {
  if (!params || !X || !y || !grad || !loss) return fail(GKS_ERR_ARG, "null");
  if (n <= 0 || n > (1 << 24)) return fail(GKS_ERR_ARG, "n must be in [1, 2^24]");
  DevBufs b;
  float *P, *dX, *dY, *dW = nullptr, *dG;
  double* dL;
  mlp::Work wk;
  HIPCHK(b.alloc(&P, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dX, (size_t)n * 16));
  HIPCHK(b.alloc(&dY, (size_t)n));
  if (w) HIPCHK(b.alloc(&dW, (size_t)n));
  HIPCHK(b.alloc(&dG, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dL, 1));
  HIPCHK(make_work(b, wk, (int)n));
  HIPCHK(hipMemcpy(P, params, GKS_MLP_NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dX, X, (size_t)n * 64, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dY, y, (size_t)n * 4, hipMemcpyHostToDevice));
  if (w) HIPCHK(hipMemcpy(dW, w, (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(mlp::fwd_bwd_reduce(wk, P, (int)n, nullptr, dX, dY, dW, dG, nullptr));
  HIPCHK(mlp::loss_sum(wk, (int)n, dL, nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(grad, dG, GKS_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(loss, dL, 8, hipMemcpyDeviceToHost));
  return GKS_OK;
}

int gks_mlp_train(float* params, int64_t n, const float* X, const float* y, const float* w, int32_t batch,
                  int32_t epochs, float lr, uint64_t seed, float* loss_per_epoch) {
  if (!params || !X || !y) return fail(GKS_ERR_ARG, "null");
  if (n <= 0 || n > (1 << 26)) return fail(GKS_ERR_ARG, "n must be in [1, 2^26]");
  if (batch <= 0 || batch > 65536 || epochs < 0 || !(lr > 0.0f)) return fail(GKS_ERR_ARG, "batch / epochs / lr");
  DevBufs b;
  float *P, *dX, *dY, *dW = nullptr, *dG, *dm, *dv;
  int* dIdx;
  double* dL;
  mlp::Work wk;
  HIPCHK(b.alloc(&P, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dX, (size_t)n * 16));
  HIPCHK(b.alloc(&dY, (size_t)n));
  if (w) HIPCHK(b.alloc(&dW, (size_t)n));
  HIPCHK(b.alloc(&dG, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dm, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dv, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dIdx, (size_t)n));
  HIPCHK(b.alloc(&dL, 1));
  HIPCHK(make_work(b, wk, batch));
  HIPCHK(hipMemcpy(P, params, GKS_MLP_NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dm, 0, GKS_MLP_NPARAM * 4));
  HIPCHK(hipMemset(dv, 0, GKS_MLP_NPARAM * 4));
  HIPCHK(hipMemcpy(dX, X, (size_t)n * 64, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dY, y, (size_t)n * 4, hipMemcpyHostToDevice));
  if (w) HIPCHK(hipMemcpy(dW, w, (size_t)n * 4, hipMemcpyHostToDevice));
  std::vector<int> order((size_t)n);
  for (int64_t k = 0; k < n; k++) order[k] = (int)k;
  std::mt19937_64 rng(seed);
  long long t = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(order.begin(), order.end(), rng);
    HIPCHK(hipMemcpy(dIdx, order.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    double ep_loss = 0.0;
    for (int64_t off = 0; off < n; off += batch) {
      const int nb = (int)std::min<int64_t>(batch, n - off);
      HIPCHK(mlp::fwd_bwd_reduce(wk, P, nb, dIdx + off, dX, dY, dW, dG, nullptr));
      t++;
      const float bc1 = 1.0f - (float)std::pow(0.9, (double)t), bc2 = 1.0f - (float)std::pow(0.999, (double)t);
      HIPCHK(mlp::adam(P, dG, dm, dv, lr, bc1, bc2, nullptr));
      if (loss_per_epoch) {
        double l = 0.0;
        HIPCHK(mlp::loss_sum(wk, nb, dL, nullptr));
        HIPCHK(hipMemcpy(&l, dL, 8, hipMemcpyDeviceToHost));
        ep_loss += l * nb;
      }
    }
    if (loss_per_epoch) loss_per_epoch[ep] = (float)(ep_loss / (double)n);
  }
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(params, P, GKS_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  return GKS_OK;
}

int gks_mlp_infer_bf16(const float* params, int64_t n, const float* X, float* logits) {
  if (!params || !X || !logits) return fail(GKS_ERR_ARG, "null");
  if (n <= 0 || n > (1 << 26)) return fail(GKS_ERR_ARG, "n must be in [1, 2^26]");
  DevBufs b;
  float *P, *dX, *dO, *dW;
  std::vector<uint16_t> packed(mlp::kPackHalves);
  mlp::pack_bf16(params, packed.data());
  HIPCHK(b.alloc(&P, GKS_MLP_NPARAM));
  HIPCHK(b.alloc(&dW, mlp::kPackHalves / 2));
  HIPCHK(b.alloc(&dX, (size_t)n * 16));
  HIPCHK(b.alloc(&dO, (size_t)n));
  HIPCHK(hipMemcpy(P, params, GKS_MLP_NPARAM * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dW, packed.data(), packed.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dX, X, (size_t)n * 64, hipMemcpyHostToDevice));
  HIPCHK(fast::table.infer(P, reinterpret_cast<const unsigned short*>(dW), n, dX, dO, nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(logits, dO, (size_t)n * 4, hipMemcpyDeviceToHost));
  return GKS_OK;
}

}  // extern "C"
