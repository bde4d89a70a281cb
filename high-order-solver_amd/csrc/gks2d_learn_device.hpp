// gks2d_learn_device.hpp — device side of the learned troubled-interface
// classifier (include/gks_learn.h): stencil features, the training label
// (WENO5-AO-DF against the linear weights, from the solver's own
// reconstruction leaves), the linear-weight reconstructions, and the MLP on
// the matrix cores.
#pragma once
#include "gks2d_device.hpp"

namespace gksd {
inline namespace GKS_VARIANT {

// MLP parameter offsets in the packed float vector (gks_learn.h)
constexpr int kMlpW1 = 0, kMlpB1 = 512, kMlpW2 = 544, kMlpB2 = 1568, kMlpW3 = 1600, kMlpB3 = 1632;

// 16 features of one stencil from its five (rho, p) pairs, in FP32: for
// each variable the values centred on their mean and scaled by their spread
// (in [-1, 1]), then log10 of the spread relative to the mean, / 16; the
// last four inputs are zero padding
GKS_DI void stencil_features_f(float* x, const float (*q)[5]) {
#pragma unroll
  for (int v = 0; v < 2; v++) {
    const float* f = q[v];
    const float m = ((((f[0] + f[1]) + f[2]) + f[3]) + f[4]) * 0.2f;
    float mx = f[0], mn = f[0];
#pragma unroll
    for (int s = 1; s < 5; s++) {
      mx = fmaxf(mx, f[s]);
      mn = fminf(mn, f[s]);
    }
    const float spread = mx - mn;
    const float inv = spread > 0.0f ? 1.0f / spread : 0.0f;
#pragma unroll
    for (int s = 0; s < 5; s++) x[5 * v + s] = (f[s] - m) * inv;
    x[10 + v] = __log2f(spread / (fabsf(m) + 1e-30f) + 1e-16f) * (0.30102999566398120f * 0.0625f);
  }
#pragma unroll
  for (int k = 12; k < 16; k++) x[k] = 0.0f;
}
GKS_DI void stencil_features(float* x, const double (*q)[5]) {
  float f[2][5];
#pragma unroll
  for (int v = 0; v < 2; v++)
#pragma unroll
    for (int s = 0; s < 5; s++) f[v][s] = (float)q[v][s];
  stencil_features_f(x, f);
}

// (rho, p) of five conservative states (x-like frame)
GKS_DI void stencil_rho_p(double (*q)[5], const double (*w)[4], const Gas& g) {
#pragma unroll
  for (int s = 0; s < 5; s++) {
    double p[4];
    convar_to_primvar(p, w[s], g);
    q[0][s] = p[0];
    q[1][s] = p[3];
  }
}

// the fifth-order candidate alone (AO weights at their linear values, DF
// factor 1): value and derivative at the cell's left (side 0) / right
// (side 1) face, the p[3] / px[3] of weno_5th_ao_with_df_left/right
GKS_DI void weno5_linear_face(int side, double& var, double& der1, double wn2, double wn1, double w0, double wp1,
                              double wp2, double h, double ih) {
  (void)h;
  (void)ih;
  if (side == 0) {
    var = w0 + ((1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2) - w0);
#if GKS_FAST
    der1 = (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) * (ih * (1.0 / 12.0));
#else
    der1 = (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) / (12.0 * h);
#endif
  } else {
    var = w0 + ((1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2) - w0);
#if GKS_FAST
    der1 = (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) * (ih * (1.0 / 12.0));
#else
    der1 = (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) / (12.0 * h);
#endif
  }
}

// the fifth-order candidate at the two tangential Gauss points (p[3],
// px[3] of weno_5th_ao_with_df_2gauss with DF factor 1)
GKS_DI void weno5_linear_2gauss(double& g1, double& g1x, double& g2, double& g2x, double wn2, double wn1, double w0,
                                double wp1, double wp2, double h, double ih) {
  (void)h;
  (void)ih;
#if GKS_FAST
  {
    // the coefficients of each stencil value combined at compile time
    constexpr double s3 = 1.732050807568877294, q = 1.0 / 4320;
    constexpr double e0 = 4314 * q, e1 = (4 + 500 * s3) * q, e2 = (-1 - 70 * s3) * q, e3 = (4 - 500 * s3) * q,
                     e4 = (-1 + 70 * s3) * q;
    constexpr double t0 = 48 * s3, t1 = -72 - 26 * s3, t2 = 9 + 2 * s3, t3 = 72 - 26 * s3, t4 = -9 + 2 * s3;
    const double r108 = ih * (1.0 / 108.0);
    g1 = w0 + ((e0 * w0 + e1 * wn1 + e2 * wn2 + e3 * wp1 + e4 * wp2) - w0);
    g2 = w0 + ((e0 * w0 + e3 * wn1 + e4 * wn2 + e1 * wp1 + e2 * wp2) - w0);
    g1x = (t0 * w0 + t1 * wn1 + t2 * wn2 + t3 * wp1 + t4 * wp2) * r108;
    g2x = -((t0 * w0 + t3 * wn1 + t4 * wn2 + t1 * wp1 + t2 * wp2) * r108);
    return;
  }
#endif
  const double sqrt3 = 1.732050807568877294;
  g1 = w0 + (GKS_CDIV(4314 * w0 + (4 + 500 * sqrt3) * wn1 - wn2 - 70 * sqrt3 * wn2 + 4 * wp1 - 500 * sqrt3 * wp1 - wp2 +
                          70 * sqrt3 * wp2,
                      4320) -
             w0);
  g2 = w0 + (GKS_CDIV(4314 * w0 + (4 - 500 * sqrt3) * wn1 - wn2 + 70 * sqrt3 * wn2 + 4 * wp1 + 500 * sqrt3 * wp1 - wp2 -
                          70 * sqrt3 * wp2,
                      4320) -
             w0);
  const double s1 = 48 * sqrt3 * w0 - 72 * wn1 - 26 * sqrt3 * wn1 + 9 * wn2 + 2 * sqrt3 * wn2 + 72 * wp1 -
                    26 * sqrt3 * wp1 - 9 * wp2 + 2 * sqrt3 * wp2;
  const double s2 = 48 * sqrt3 * w0 + 72 * wn1 - 26 * sqrt3 * wn1 - 9 * wn2 + 2 * sqrt3 * wn2 - 72 * wp1 -
                    26 * sqrt3 * wp1 + 9 * wp2 + 2 * sqrt3 * wp2;
#if GKS_FAST
  g1x = s1 * (ih * (1.0 / 108.0));
  g2x = -(s2 * (ih * (1.0 / 108.0)));
#else
  g1x = s1 / (108 * h);
  g2x = -(s2 / (108 * h));
#endif
}

// training label of one cell's normal reconstruction: 1 when the DF sensor
// is nonzero on the stencil, or when for some characteristic (or
// conservative) component and face the WENO5-AO-DF value or derivative
// differs from the linear one by more than tol times that component's range
// over the stencil
GKS_DI float normal_label(const double (*w)[4], const double* alpha, double h, double ih, int recon_var,
                          const Gas& g, double rel_tol) {
  double amax = 0.0;
#pragma unroll
  for (int s = 0; s < 5; s++) amax = fmax(amax, alpha[s]);
  if (amax > 0.0) return 1.0f;
  const double df[4] = {1.0, 1.0, 1.0, 1.0};
  double pn1[4], p0[4], pp1[4];
  convar_to_primvar(pn1, w[1], g);
  convar_to_primvar(p0, w[2], g);
  convar_to_primvar(pp1, w[3], g);
  bool bad = false;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double base[4];
#pragma unroll
    for (int i = 0; i < 4; i++) base[i] = side == 0 ? 0.5 * (pn1[i] + p0[i]) : 0.5 * (pp1[i] + p0[i]);
    double c[5][4];
    if (recon_var == 0) {
#pragma unroll
      for (int s = 0; s < 5; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) c[s][i] = w[s][i];
    } else {
      CharBase cb;
      char_base(cb, base, g);
#pragma unroll
      for (int s = 0; s < 5; s++) to_char(c[s], cb, w[s]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double v, d1, vl, dl;
      if (side == 0) weno5_left(v, d1, c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h, ih);
      else weno5_right(v, d1, c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h, ih);
      weno5_linear_face(side, vl, dl, c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], h, ih);
      double mx = c[0][i], mn = c[0][i];
#pragma unroll
      for (int s = 1; s < 5; s++) {
        mx = fmax(mx, c[s][i]);
        mn = fmin(mn, c[s][i]);
      }
      const double tol = rel_tol * (mx - mn) + 1e-13 * fabs(c[2][i]);
      if (fabs(v - vl) > tol || fabs(d1 - dl) * h > tol) bad = true;
    }
  }
  return bad ? 1.0f : 0.0f;
}

// ------------------------------------------------------------- inference
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// two floats to BF16 (round to nearest even, one v_cvt_pk_bf16_f32), low
// half first
GKS_DI unsigned pack_bf16x2(float a, float b) {
  const f32x2_t v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2_hw));
}
GKS_DI bf16x4_t relu_bf16x4(const f32x4_t& acc, const float* bias) {
  const uint2 u = make_uint2(pack_bf16x2(fmaxf(acc[0] + bias[0], 0.0f), fmaxf(acc[1] + bias[1], 0.0f)),
                             pack_bf16x2(fmaxf(acc[2] + bias[2], 0.0f), fmaxf(acc[3] + bias[3], 0.0f)));
  return __builtin_bit_cast(bf16x4_t, u);
}

// orders this wave's LDS writes before its later LDS reads of other lanes'
// data (the LDS processes one wave's operations in order, so no workgroup
// barrier is needed for data that stays within the wave)
GKS_DI void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Where one wave reads the MLP's operands: straight from global memory
// (P: the float parameters, biases read from it; W: the weights packed in
// the operand layout of mlp_pack.hpp), or from a copy a workgroup staged in
// LDS at its start (mlp_stage), so that their global latency overlaps the
// kernel's own loads instead of following them.
struct MlpGlobal {
  const float* __restrict__ P;
  const unsigned short* __restrict__ W;
  GKS_DI bf16x4_t w(int v, int lane) const {
    return __builtin_bit_cast(bf16x4_t, reinterpret_cast<const uint2*>(W)[v * 64 + lane]);
  }
  GKS_DI float b1(int i) const { return P[kMlpB1 + i]; }
  GKS_DI float b2(int i) const { return P[kMlpB2 + i]; }
  GKS_DI float b3() const { return P[kMlpB3]; }
};
struct MlpLds {
  uint2 w[8 * 64];
  float b[65];  // b1[32], b2[32], b3
};
// every thread of the workgroup calls it; a __syncthreads() must separate
// it from the first read
GKS_DI void mlp_stage(MlpLds& s, const float* __restrict__ P, const unsigned short* __restrict__ W) {
  const uint2* Wv = reinterpret_cast<const uint2*>(W);
  for (int t = threadIdx.x; t < 8 * 64; t += blockDim.x) s.w[t] = Wv[t];
  for (int t = threadIdx.x; t < 65; t += blockDim.x)
    s.b[t] = t < 32 ? P[kMlpB1 + t] : (t < 64 ? P[kMlpB2 + t - 32] : P[kMlpB3]);
}
struct MlpShared {
  const MlpLds* s;
  GKS_DI bf16x4_t w(int v, int lane) const { return __builtin_bit_cast(bf16x4_t, s->w[v * 64 + lane]); }
  GKS_DI float b1(int i) const { return s->b[i]; }
  GKS_DI float b2(int i) const { return s->b[32 + i]; }
  GKS_DI float b3() const { return s->b[64]; }
};

// One wave evaluates the MLP for its 64 lanes' samples (x: this lane's 16
// features) and returns this lane's logit.  Every lane of the wave must
// call it (an inactive lane passes zeros); other waves of the workgroup are
// not involved.  The layers run transposed, H^T = W^T X^T, on 16x16x16 BF16
// MFMAs (FP32 accumulation): the output tile of one layer is then already
// laid out as the B operand of the next, so only the input tile passes
// through LDS.  src: the operands (MlpGlobal or MlpShared); xs: this wave's
// 64 x 16 staging area (BF16, 16-byte aligned), ls: its 64 logits.
template <class S>
GKS_DI float mlp_wave_logit(const float* x, const S& src, short* xs, float* ls) {
  const int lane = threadIdx.x & 63, col = lane & 15, kb = lane >> 4;
  {
    uint4* dst = reinterpret_cast<uint4*>(xs + lane * 16);
    dst[0] = make_uint4(pack_bf16x2(x[0], x[1]), pack_bf16x2(x[2], x[3]), pack_bf16x2(x[4], x[5]),
                        pack_bf16x2(x[6], x[7]));
    dst[1] = make_uint4(pack_bf16x2(x[8], x[9]), pack_bf16x2(x[10], x[11]), pack_bf16x2(x[12], x[13]),
                        pack_bf16x2(x[14], x[15]));
  }
  // A operands (weights, transposed), one 8-byte load each
  bf16x4_t a[8];
#pragma unroll
  for (int v = 0; v < 8; v++) a[v] = src.w(v, lane);
  float b1v[2][4], b2v[2][4];
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      b1v[m][r] = src.b1(m * 16 + 4 * kb + r);
      b2v[m][r] = src.b2(m * 16 + 4 * kb + r);
    }
  const float b3 = src.b3();
  wave_lds_sync();
#pragma unroll
  for (int t = 0; t < 4; t++) {
    // B operand of layer 1: X^T[k = 4 kb + r][n = col] of sample 16 t + col
    const bf16x4_t bx = *reinterpret_cast<const bf16x4_t*>(xs + (16 * t + col) * 16 + 4 * kb);
    bf16x4_t h1[2], h2[2];
#pragma unroll
    for (int m = 0; m < 2; m++) {
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], bx, acc, 0, 0, 0);
      h1[m] = relu_bf16x4(acc, b1v[m]);
    }
#pragma unroll
    for (int m = 0; m < 2; m++) {
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2 + 2 * m], h1[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[3 + 2 * m], h1[1], acc, 0, 0, 0);
      h2[m] = relu_bf16x4(acc, b2v[m]);
    }
    f32x4_t z = {0.f, 0.f, 0.f, 0.f};
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[6], h2[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[7], h2[1], z, 0, 0, 0);
    // row 0 of z^T: lanes with kb == 0 hold the logits of samples 16 t + col
    if (kb == 0) ls[16 * t + col] = z[0] + b3;
  }
  wave_lds_sync();
  return ls[lane];
}
GKS_DI float mlp_wave_logit(const float* x, const float* __restrict__ P, const unsigned short* __restrict__ W,
                            short* xs, float* ls) {
  return mlp_wave_logit(x, MlpGlobal{P, W}, xs, ls);
}

}  // namespace GKS_VARIANT
}  // namespace gksd
