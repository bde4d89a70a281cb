// gks3d_device.hpp — device numerics of the 3D GKS path (gfx950, FP64).
//
// Each function restates one leaf of the reference's 3D scheme (file:line in
// the reference repository) in the reference's floating-point operation
// order, so that the IEEE build (GKS_FAST 0, -ffp-contract=off) differs from
// a CPU run of the same algorithm only through the device erfc / exp / sqrt
// (and the order in which a face's Gauss points are summed).  The fast build
// (GKS_FAST 1) contracts FMAs and multiplies by reciprocals (G3_* below).
//
// Departures from the shipped reference, all deliberate (its 3D path does
// not run as shipped):
//  * the Maxwellian's lambda comes from the conservative state (Lambda,
//    basic_function.cpp:1179-1182); GKS3D (flux_function.cpp:813-1163)
//    passes the pressure there;
//  * the line / point order-reduction fallback is decided afresh at every
//    reconstruction (Recon3d::order_reduce is never cleared in the
//    reference).
//
// The Maxwellian moment tables of MMDF3d (4 x 7x6x6x3 doubles) are never
// materialised: entry (i,j,k,x) is the product u[i] * v[j] * w[k] * xi_x
// formed where it is used, with the same rounding as the table entry.
#pragma once
#include <hip/hip_runtime.h>

#include "gks3d_device_types.hpp"

#ifndef GKS_VARIANT
#error "GKS_VARIANT must name the kernel build"
#endif
#ifndef GKS_FAST
#error "GKS_FAST must be 0 or 1"
#endif

namespace gksd3 {
inline namespace GKS_VARIANT {

#ifndef GKS3_DI
#define GKS3_DI __device__ __forceinline__
#endif

// The fast build (GKS_FAST 1) turns divisions by constants, by the cell
// size and by a denominator shared by several quotients into
// multiplications by a reciprocal formed once (each quotient moves by an ulp
// or two); the IEEE build keeps every division of the reference.
#if GKS_FAST
#define G3_CDIV(a, b) ((a) * (1.0 / (b)))
#else
#define G3_CDIV(a, b) ((a) / (b))
#endif
// a / b and 1 / b for a runtime b that is finite, nonzero and normal: in the
// fast build the hardware reciprocal through two Newton steps (and one
// residual correction for a quotient), 5 and 8 instructions against about
// 11 for IEEE division; NaN for b = 0, so time_coe3's dt / tau_num stays
// IEEE
#if GKS_FAST
GKS3_DI double frcp3(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double t = fma(-b, r, 1.0);
  r = fma(t, r, r);
  t = fma(-b, r, 1.0);
  return fma(t, r, r);
}
GKS3_DI double fdiv3(double a, double b) {
  const double r = frcp3(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
#define G3_DIV(a, b) fdiv3((a), (b))
#define G3_RCP(b) frcp3(b)
#else
#define G3_DIV(a, b) ((a) / (b))
#define G3_RCP(b) (1.0 / (b))
#endif
// the cell size of a reconstruction and, for the fast build, its reciprocals
struct HInv {
  double h, r1, r6, r12, r108;
};
GKS3_DI HInv hinv(double h) {
  HInv r;
  r.h = h;
  r.r1 = 1.0 / h;
  r.r6 = 1.0 / (6 * h);
  r.r12 = 1.0 / (12.0 * h);
  r.r108 = 1.0 / (108 * h);
  return r;
}
#if GKS_FAST
#define G3_DH(a, H) ((a) * (H).r1)
#define G3_D6H(a, H) ((a) * (H).r6)
#define G3_D12H(a, H) ((a) * (H).r12)
#define G3_D108H(a, H) ((a) * (H).r108)
#else
#define G3_DH(a, H) ((a) / (H).h)
#define G3_D6H(a, H) ((a) / (6 * (H).h))
#define G3_D12H(a, H) ((a) / (12.0 * (H).h))
#define G3_D108H(a, H) ((a) / (108 * (H).h))
#endif

// ---------------------------------------------------------------- gas algebra
// Convar_to_primvar_3D: basic_function.cpp:819-826 (Pressure :1162-1167)
GKS3_DI void prim3(double* p, const double* c, const Gas3& g) {
  p[0] = c[0];
#if GKS_FAST
  const double ir = G3_RCP(c[0]);
  p[1] = c[1] * ir;
  p[2] = c[2] * ir;
  p[3] = c[3] * ir;
  p[4] = (g.gamma - 1.0) * (c[4] - 0.5 * c[1] * c[1] * ir - 0.5 * c[2] * c[2] * ir - 0.5 * c[3] * c[3] * ir);
#else
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = c[3] / c[0];
  p[4] = (g.gamma - 1.0) * (c[4] - 0.5 * c[1] * c[1] / c[0] - 0.5 * c[2] * c[2] / c[0] - 0.5 * c[3] * c[3] / c[0]);
#endif
}
// Lambda (3D): basic_function.cpp:1179-1182
GKS3_DI double lambda3(double rho, double u, double v, double w, double rhoE, const Gas3& g) {
  return (g.K + 3.0) * 0.25 * G3_DIV(rho, rhoE - 0.5 * rho * (u * u + v * v + w * w));
}
// (rho, u, v, w, lambda): the state the Maxwellian moments need
GKS3_DI void ulambda3(double* p, const double* c, const Gas3& g) {
  p[0] = c[0];
#if GKS_FAST
  const double ir = G3_RCP(c[0]);
  p[1] = c[1] * ir;
  p[2] = c[2] * ir;
  p[3] = c[3] * ir;
#else
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = c[3] / c[0];
#endif
  p[4] = lambda3(c[0], p[1], p[2], p[3], c[4], g);
}
// Check_Order_Reduce_by_Lambda_3D: reconstruction.cpp:3755-3771
GKS3_DI bool reduce3(const double* c, const Gas3& g) {
#if GKS_FAST
  const double ir = G3_RCP(c[0]);
  const double l = lambda3(c[0], c[1] * ir, c[2] * ir, c[3] * ir, c[4], g);
#else
  const double l = lambda3(c[0], c[1] / c[0], c[2] / c[0], c[3] / c[0], c[4], g);
#endif
  return l <= 0.0 || isnan(l);
}
// Ydirection / Zdirection: basic_function.cpp:1208-1224 (global frame ->
// the frame of a y / z face)
GKS3_DI void to_frame(double* o, const double* f, int dir) {
  o[0] = f[0];
  o[4] = f[4];
  if (dir == 0) {
    o[1] = f[1];
    o[2] = f[2];
    o[3] = f[3];
  } else if (dir == 1) {
    o[1] = f[2];
    o[2] = -f[1];
    o[3] = f[3];
  } else {
    o[1] = f[3];
    o[2] = f[2];
    o[3] = -f[1];
  }
}

// Char_base_3D: basic_function.cpp:994-1035 (the left eigenvectors, scaled)
GKS3_DI void char_base3(double* s, const double* prim, const Gas3& g) {
  const double Gamma = g.gamma;
  const double c = sqrt(G3_DIV(Gamma * prim[4], prim[0]));
  const double overc = G3_RCP(c);
  const double alfa = (Gamma - 1.) * 0.5 * overc * overc;
  const double u = prim[1], v = prim[2], w = prim[3];
  const double r_over = G3_DIV(1.0, Gamma - 1);
  const double H = 0.5 * (u * u + v * v + w * w) + G3_DIV(0.5, alfa);
  s[0] = H + c * r_over * (u - c);
  s[1] = -(u + c * r_over);
  s[2] = -v;
  s[3] = -w;
  s[4] = 1;
  s[5] = -2 * H + 4 * r_over * c * c;
  s[6] = 2 * u;
  s[7] = 2 * v;
  s[8] = 2 * w;
  s[9] = -2;
  s[10] = -2 * v * c * c * r_over;
  s[11] = 0;
  s[12] = 2 * c * c * r_over;
  s[13] = 0;
  s[14] = 0;
  s[15] = -2 * w * c * c * r_over;
  s[16] = 0;
  s[17] = 0;
  s[18] = 2 * c * c * r_over;
  s[19] = 0;
  s[20] = H - c * r_over * (u + c);
  s[21] = -u + c * r_over;
  s[22] = -v;
  s[23] = -w;
  s[24] = 1;
#pragma unroll
  for (int i = 0; i < 25; i++) s[i] = s[i] * alfa;
}
// Convar_to_char_3D: basic_function.cpp:1037-1050
GKS3_DI void to_char3(double* ch, const double* s, const double* c) {
#if GKS_FAST
  // rows 2 and 3 hold three zero entries each: dropped
#pragma unroll
  for (int i = 0; i < 5; i++) {
    if (i == 2) {
      ch[2] = s[10] * c[0] + s[12] * c[2];
      continue;
    }
    if (i == 3) {
      ch[3] = s[15] * c[0] + s[18] * c[3];
      continue;
    }
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 5; j++) a = a + s[i * 5 + j] * c[j];
    ch[i] = a;
  }
  return;
#endif
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 5; j++) a = a + s[i * 5 + j] * c[j];
    ch[i] = a;
  }
}
// Char_to_convar_3D: basic_function.cpp:1052-1093 (the right eigenvectors,
// zero and unit entries kept so the sums run in the reference's order)
struct RightEig {
  double cs, u, v, w, H;
};
GKS3_DI RightEig right_eig(const double* prim, const Gas3& g) {
  RightEig r;
  r.cs = sqrt(G3_DIV(g.gamma * prim[4], prim[0]));
  const double alfa = G3_DIV(g.gamma - 1., 2.0 * r.cs * r.cs);
  r.u = prim[1];
  r.v = prim[2];
  r.w = prim[3];
  r.H = 0.5 * (r.u * r.u + r.v * r.v + r.w * r.w) + G3_DIV(0.5, alfa);
  return r;
}
GKS3_DI void from_char3(double* c, const RightEig& e, const double* ch) {
  const double cs = e.cs, u = e.u, v = e.v, w = e.w, H = e.H;
#if GKS_FAST
  // the zero and unit entries dropped
  c[0] = ch[0] + ch[1] + ch[4];
  c[1] = (u - cs) * ch[0] + u * ch[1] + (u + cs) * ch[4];
  c[2] = v * ch[0] + v * ch[1] + ch[2] + v * ch[4];
  c[3] = w * ch[0] + w * ch[1] + ch[3] + w * ch[4];
  c[4] = (H - u * cs) * ch[0] + 0.5 * (u * u + v * v + w * w) * ch[1] + v * ch[2] + w * ch[3] + (H + u * cs) * ch[4];
  return;
#endif
  const double s[5][5] = {{1., 1.0, 0.0, 0.0, 1.0},
                          {u - cs, u, 0, 0, u + cs},
                          {v, v, 1, 0, v},
                          {w, w, 0, 1, w},
                          {H - u * cs, 0.5 * (u * u + v * v + w * w), v, w, H + u * cs}};
#pragma unroll
  for (int i = 0; i < 5; i++) {
    double a = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) a = a + s[i][j] * ch[j];
    c[i] = a;
  }
}

// ------------------------------------------------------------ WENO5-AO
// AO weights of weno_5th_ao_left / _right / _2gauss without DF (wenotype
// wenoz), reconstruction.cpp:161-358, 2393-2551: fw = (w_k - w3 d_k / d3,
// w3 / d3); eps 1e-10 at the faces, 1e-6 at the Gauss points
GKS3_DI void weno5_weights3(double* fw, double wn2, double wn1, double w0, double wp1, double wp2, double eps) {
  const double dhi = 0.85, dlo = 0.85;
  double d[4], beta[4], alpha[4];
  d[0] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[1] = (1.0 - dhi) * dlo;
  d[2] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[3] = dhi;
  double q = wn2 - 2.0 * wn1 + w0, r = wn2 - 4.0 * wn1 + 3.0 * w0;
  beta[0] = 13.0 / 12.0 * (q * q) + 0.25 * (r * r);
  q = wn1 - 2.0 * w0 + wp1;
  r = wn1 - wp1;
  beta[1] = 13.0 / 12.0 * (q * q) + 0.25 * (r * r);
  q = w0 - 2.0 * wp1 + wp2;
  r = 3.0 * w0 - 4.0 * wp1 + wp2;
  beta[2] = 13.0 / 12.0 * (q * q) + 0.25 * (r * r);
  beta[3] = (1.0 / 5040.0) * (231153.0 * w0 * w0 + 104963.0 * wn1 * wn1 + 6908.0 * wn2 * wn2 -
                              38947.0 * wn2 * wp1 + 104963.0 * wp1 * wp1 +
                              wn1 * (-51001.0 * wn2 + 179098.0 * wp1 - 38947.0 * wp2) -
                              3.0 * w0 * (99692.0 * wn1 - 22641.0 * wn2 + 99692.0 * wp1 - 22641.0 * wp2) +
                              8209.0 * wn2 * wp2 - 51001.0 * wp1 * wp2 + 6908.0 * wp2 * wp2);
  const double tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  double sum_alpha = 0.0;
#if GKS_FAST
  // the four quotients tau5 / (beta_k + eps) from one reciprocal of the
  // product of the denominators (each >= eps, the product stays in range)
  double gq[4];
  {
    const double b0 = beta[0] + eps, b1 = beta[1] + eps, b2 = beta[2] + eps, b3 = beta[3] + eps;
    const double p01 = b0 * b1, p23 = b2 * b3;
    const double r = fdiv3(tau5, p01 * p23);
    gq[0] = r * b1 * p23;
    gq[1] = r * b0 * p23;
    gq[2] = r * p01 * b3;
    gq[3] = r * p01 * b2;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double gd = gq[i];
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum_alpha += alpha[i];
  }
#else
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double gd = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum_alpha += alpha[i];
  }
#endif
  double ww[4];
#if GKS_FAST
  const double rs = frcp3(sum_alpha);
#pragma unroll
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] * rs;
  fw[3] = ww[3] * (1.0 / dhi);
#pragma unroll
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] * (1.0 / dhi) * d[k];
#else
#pragma unroll
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum_alpha;
  fw[3] = ww[3] / d[3];
#pragma unroll
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] / d[3] * d[k];
#endif
}
// weno_5th_ao_left: reconstruction.cpp:161-258 (value, first derivative)
GKS3_DI void weno5_left3(double& var, double& der1, double wn2, double wn1, double w0, double wp1, double wp2,
                         const HInv& H) {
  double fw[4], p[4], px[4];
  weno5_weights3(fw, wn2, wn1, w0, wp1, wp2, 1e-10);
  p[0] = -1.0 / 6.0 * wn2 + 5.0 / 6.0 * wn1 + 1.0 / 3.0 * w0;
  p[1] = 1.0 / 3.0 * wn1 + 5.0 / 6.0 * w0 - 1.0 / 6.0 * wp1;
  p[2] = 11.0 / 6.0 * w0 - 7.0 / 6.0 * wp1 + 1.0 / 3.0 * wp2;
  p[3] = (1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2);
  px[0] = G3_DH(w0 - wn1, H);
  px[1] = G3_DH(w0 - wn1, H);
  px[2] = -(G3_DH(2.0 * w0 - 3.0 * wp1 + wp2, H));
  px[3] = G3_D12H(15.0 * w0 - 15.0 * wn1 + wn2 - wp1, H);
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
// weno_5th_ao_right: reconstruction.cpp:260-358
GKS3_DI void weno5_right3(double& var, double& der1, double wn2, double wn1, double w0, double wp1, double wp2,
                          const HInv& H) {
  double fw[4], p[4], px[4];
  weno5_weights3(fw, wn2, wn1, w0, wp1, wp2, 1e-10);
  p[0] = 1.0 / 3.0 * wn2 - 7.0 / 6.0 * wn1 + 11.0 / 6.0 * w0;
  p[1] = -1.0 / 6.0 * wn1 + 5.0 / 6.0 * w0 + 1.0 / 3.0 * wp1;
  p[2] = 1.0 / 3.0 * w0 + 5.0 / 6.0 * wp1 - 1.0 / 6.0 * wp2;
  p[3] = (1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2);
  px[0] = G3_DH(2.0 * w0 - 3.0 * wn1 + wn2, H);
  px[1] = G3_DH(-w0 + wp1, H);
  px[2] = G3_DH(-w0 + wp1, H);
  px[3] = G3_D12H(-15.0 * w0 + wn1 + 15.0 * wp1 - wp2, H);
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
// weno_5th_ao_2gauss: reconstruction.cpp:2393-2551: values at the Gauss
// points -+ h / (2 sqrt 3) and, when DER, their first derivatives
template <bool DER>
GKS3_DI void weno5_2gauss3(double& g1, double& g1x, double& g2, double& g2x, double wn2, double wn1, double w0,
                           double wp1, double wp2, const HInv& H) {
  double fw[4], p[4];
  weno5_weights3(fw, wn2, wn1, w0, wp1, wp2, 1e-6);
  const double sqrt3 = 1.732050807568877294;
#if GKS_FAST
  // the same candidate polynomials with the coefficients of each stencil
  // value combined at compile time
  {
    constexpr double s3 = 1.732050807568877294;
    constexpr double c43 = 1.0 / (4 * s3), q = 1.0 / 4320;
    constexpr double a0 = 1 - s3 / 4, a1 = 1 / s3, a2 = -c43;
    constexpr double b0 = (4 + s3) / 4, b1 = -s3 / 3, b2 = s3 / 12;
    constexpr double e0 = 4314 * q, e1 = (4 + 500 * s3) * q, e2 = (-1 - 70 * s3) * q, e3 = (4 - 500 * s3) * q,
                     e4 = (-1 + 70 * s3) * q;
    constexpr double k0 = -9 + s3, k1 = -2 * (-6 + s3), k2 = -3 + s3;
    constexpr double m0 = -2 * s3, m1 = 3 + s3, m2 = -3 + s3;
    constexpr double n0 = 9 + s3, n1 = -2 * (6 + s3), n2 = 3 + s3;
    constexpr double t0 = 48 * s3, t1 = -72 - 26 * s3, t2 = 9 + 2 * s3, t3 = 72 - 26 * s3, t4 = -9 + 2 * s3;
    (void)sqrt3;
    p[0] = a0 * w0 + a1 * wn1 + a2 * wn2;
    p[1] = w0 + c43 * wn1 - c43 * wp1;
    p[2] = b0 * w0 + b1 * wp1 + b2 * wp2;
    p[3] = e0 * w0 + e1 * wn1 + e2 * wn2 + e3 * wp1 + e4 * wp2;
    g1 = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) g1 += fw[k] * p[k];
    g1x = 0.0;
    if (DER) {
      double px[4];
      px[0] = -(k0 * w0 + k1 * wn1 + k2 * wn2) * H.r6;
      px[1] = -(m0 * w0 + m1 * wn1 + m2 * wp1) * H.r6;
      px[2] = -(n0 * w0 + n1 * wp1 + n2 * wp2) * H.r6;
      px[3] = (t0 * w0 + t1 * wn1 + t2 * wn2 + t3 * wp1 + t4 * wp2) * H.r108;
#pragma unroll
      for (int k = 0; k < 4; k++) g1x += fw[k] * px[k];
    }
    p[0] = b0 * w0 + b1 * wn1 + b2 * wn2;
    p[1] = w0 - c43 * wn1 + c43 * wp1;
    p[2] = a0 * w0 + a1 * wp1 + a2 * wp2;
    p[3] = e0 * w0 + e3 * wn1 + e4 * wn2 + e1 * wp1 + e2 * wp2;
    g2 = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) g2 += fw[k] * p[k];
    g2x = 0.0;
    if (DER) {
      double px[4];
      px[0] = (n0 * w0 + n1 * wn1 + n2 * wn2) * H.r6;
      px[1] = (m0 * w0 + m2 * wn1 + m1 * wp1) * H.r6;
      px[2] = (k0 * w0 + k1 * wp1 + k2 * wp2) * H.r6;
      px[3] = -((t0 * w0 + t3 * wn1 + t4 * wn2 + t1 * wp1 + t2 * wp2) * H.r108);
#pragma unroll
      for (int k = 0; k < 4; k++) g2x += fw[k] * px[k];
    }
    return;
  }
#endif
  p[0] = w0 - G3_CDIV(sqrt3 * w0, 4) + G3_CDIV(4 * wn1 - wn2, 4 * sqrt3);
  p[1] = w0 + G3_CDIV(wn1 - wp1, 4 * sqrt3);
  p[2] = (1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wp1 + wp2));
  p[3] = G3_CDIV(4314 * w0 + (4 + 500 * sqrt3) * wn1 - wn2 - 70 * sqrt3 * wn2 + 4 * wp1 - 500 * sqrt3 * wp1 - wp2 +
                     70 * sqrt3 * wp2,
                 4320);
  g1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) g1 += fw[k] * p[k];
  g1x = 0.0;
  if (DER) {
    double px[4];
    px[0] = G3_D6H(-((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wn1 + (-3 + sqrt3) * wn2), H);
    px[1] = G3_D6H(-(-2 * sqrt3 * w0 + (3 + sqrt3) * wn1 + (-3 + sqrt3) * wp1), H);
    px[2] = G3_D6H(-((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wp1 + (3 + sqrt3) * wp2), H);
    px[3] = G3_D108H(48 * sqrt3 * w0 - 72 * wn1 - 26 * sqrt3 * wn1 + 9 * wn2 + 2 * sqrt3 * wn2 + 72 * wp1 -
                         26 * sqrt3 * wp1 - 9 * wp2 + 2 * sqrt3 * wp2,
                     H);
#pragma unroll
    for (int k = 0; k < 4; k++) g1x += fw[k] * px[k];
  }
  p[0] = (1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wn1 + wn2));
  p[1] = w0 + G3_CDIV(-wn1 + wp1, 4 * sqrt3);
  p[2] = w0 - G3_CDIV(sqrt3 * w0, 4) + G3_CDIV(4 * wp1 - wp2, 4 * sqrt3);
  p[3] = G3_CDIV(4314 * w0 + (4 - 500 * sqrt3) * wn1 - wn2 + 70 * sqrt3 * wn2 + 4 * wp1 + 500 * sqrt3 * wp1 - wp2 -
                     70 * sqrt3 * wp2,
                 4320);
  g2 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) g2 += fw[k] * p[k];
  g2x = 0.0;
  if (DER) {
    double px[4];
    px[0] = G3_D6H((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wn1 + (3 + sqrt3) * wn2, H);
    px[1] = G3_D6H(-2 * sqrt3 * w0 + (-3 + sqrt3) * wn1 + (3 + sqrt3) * wp1, H);
    px[2] = G3_D6H((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wp1 + (-3 + sqrt3) * wp2, H);
    px[3] = -(G3_D108H(48 * sqrt3 * w0 + 72 * wn1 - 26 * sqrt3 * wn1 - 9 * wn2 + 2 * sqrt3 * wn2 - 72 * wp1 -
                           26 * sqrt3 * wp1 + 9 * wp2 + 2 * sqrt3 * wp2,
                       H));
#pragma unroll
    for (int k = 0; k < 4; k++) g2x += fw[k] * px[k];
  }
}

// ------------------------------------------------------------ GKS3D
// MMDF3d_left / _right / MMDF3d: basic_function.cpp:189-393.  uh: the
// half-space u moments (u > 0 for the left state, u < 0 for the right), uw
// the whole-space ones (mm3_side_from01 and half01 below fill a side).
struct MM3 {
  double uh[7], uw[7], vw[6], ww[6], xi2, xi4;
};
// the whole-space moments of MMDF3d (it divides by lambda where the
// half-space classes multiply by 1/lambda)
GKS3_DI void mm3_whole(MM3& m, double u, double v, double w, double lambda, const Gas3& g) {
#if GKS_FAST
  const double ol = frcp3(lambda);
#define G3_DL(a) ((a) * ol)
#define G3_DLL(a) ((a) * ol * ol)
#else
#define G3_DL(a) ((a) / lambda)
#define G3_DLL(a) ((a) / (lambda * lambda))
#endif
  m.uw[0] = 1;
  m.uw[1] = u;
#pragma unroll
  for (int i = 2; i <= 6; i++) m.uw[i] = u * m.uw[i - 1] + G3_DL(0.5 * (i - 1)) * m.uw[i - 2];
  m.vw[0] = 1;
  m.vw[1] = v;
  m.ww[0] = 1;
  m.ww[1] = w;
#pragma unroll
  for (int i = 2; i <= 5; i++) {
    m.vw[i] = v * m.vw[i - 1] + G3_DL(0.5 * (i - 1)) * m.vw[i - 2];
    m.ww[i] = w * m.ww[i - 1] + G3_DL(0.5 * (i - 1)) * m.ww[i - 2];
  }
  m.xi2 = G3_DL(0.5 * g.K);
  m.xi4 = G3_DLL(0.25 * (g.K * g.K + 2 * g.K));
#undef G3_DL
#undef G3_DLL
}
// table entries upvwxi / unvwxi (half, H) and uvwxi (whole, W)
template <bool HALF>
struct Tab3 {
  const MM3& m;
  GKS3_DI double operator()(int i, int j, int k, int x) const {
    const double a = HALF ? m.uh[i] : m.uw[i];
    const double b = a * m.vw[j] * m.ww[k];
    return x == 0 ? b : (x == 1 ? b * m.xi2 : b * m.xi4);
  }
};
// GL_address / GR_address / G_address (3D): basic_function.h:225-567
template <class T>
GKS3_DI void contract3(int u, int v, int w, int x, double* psi, const double* a, const T& m) {
  psi[0] = a[0] * m(u, v, w, x) + a[1] * m(u + 1, v, w, x) + a[2] * m(u, v + 1, w, x) + a[3] * m(u, v, w + 1, x) +
           a[4] * 0.5 * (m(u + 2, v, w, x) + m(u, v + 2, w, x) + m(u, v, w + 2, x) + m(u, v, w, x + 1));
  psi[1] = a[0] * m(u + 1, v, w, x) + a[1] * m(u + 2, v, w, x) + a[2] * m(u + 1, v + 1, w, x) +
           a[3] * m(u + 1, v, w + 1, x) +
           a[4] * 0.5 * (m(u + 3, v, w, x) + m(u + 1, v + 2, w, x) + m(u + 1, v, w + 2, x) + m(u + 1, v, w, x + 1));
  psi[2] = a[0] * m(u, v + 1, w, x) + a[1] * m(u + 1, v + 1, w, x) + a[2] * m(u, v + 2, w, x) +
           a[3] * m(u, v + 1, w + 1, x) +
           a[4] * 0.5 * (m(u + 2, v + 1, w, x) + m(u, v + 3, w, x) + m(u, v + 1, w + 2, x) + m(u, v + 1, w, x + 1));
  psi[3] = a[0] * m(u, v, w + 1, x) + a[1] * m(u + 1, v, w + 1, x) + a[2] * m(u, v + 1, w + 1, x) +
           a[3] * m(u, v, w + 2, x) +
           a[4] * 0.5 * (m(u + 2, v, w + 1, x) + m(u, v + 2, w + 1, x) + m(u, v, w + 3, x) + m(u, v, w + 1, x + 1));
  psi[4] = 0.5 * (a[0] * (m(u + 2, v, w, x) + m(u, v + 2, w, x) + m(u, v, w + 2, x) + m(u, v, w, x + 1)) +
                  a[1] * (m(u + 3, v, w, x) + m(u + 1, v + 2, w, x) + m(u + 1, v, w + 2, x) + m(u + 1, v, w, x + 1)) +
                  a[2] * (m(u + 2, v + 1, w, x) + m(u, v + 3, w, x) + m(u, v + 1, w + 2, x) + m(u, v + 1, w, x + 1)) +
                  a[3] * (m(u + 2, v, w + 1, x) + m(u, v + 2, w + 1, x) + m(u, v, w + 3, x) + m(u, v, w + 1, x + 1)) +
                  a[4] * 0.5 *
                      (m(u + 4, v, w, x) + m(u, v + 4, w, x) + m(u, v, w + 4, x) + m(u, v, w, x + 2) +
                       2 * m(u + 2, v + 2, w, x) + 2 * m(u + 2, v, w + 2, x) + 2 * m(u + 2, v, w, x + 1) +
                       2 * m(u, v + 2, w + 2, x) + 2 * m(u, v + 2, w, x + 1) + 2 * m(u, v, w + 2, x + 1)));
}
// A (3D): basic_function.cpp:585-601; prim = (rho, u, v, w, lambda)
GKS3_DI void micro_slope3(double* a, const double* der, const double* prim, const Gas3& g) {
  const double density = prim[0], u = prim[1], v = prim[2], w = prim[3], lambda = prim[4];
  const double overden = G3_RCP(density);
  const double overlambda = G3_RCP(lambda);
  const double A = 2 * der[4] * overden - (u * u + v * v + w * w + 0.5 * (g.K + 3) * overlambda) * der[0] * overden;
  const double B = (der[1] - u * der[0]) * overden;
  const double C = (der[2] - v * der[0]) * overden;
  const double D = (der[3] - w * der[0]) * overden;
  a[4] = G3_DIV(4 * lambda * lambda, g.K + 3) * (A - 2 * u * B - 2 * v * C - 2 * w * D);
  a[3] = 2 * lambda * D - w * a[4];
  a[2] = 2 * lambda * C - v * a[4];
  a[1] = 2 * lambda * B - u * a[4];
  a[0] = der[0] * overden - u * a[1] - v * a[2] - w * a[3] -
         0.5 * a[4] * (u * u + v * v + w * w + 0.5 * (g.K + 3) * overlambda);
}
// Get_Tau_NS / Get_Tau: basic_function.cpp:638-708
GKS3_DI double tau_ns3(double rho0, double l0, const Gas3& g) {
  if (g.tau_type == 0) return 0.0;
  if (g.mu > 0.0) return G3_DIV(2.0 * g.mu * l0, rho0);
  if (g.nu > 0.0) return 2 * g.nu * l0;
  return 0.0;
}
GKS3_DI double tau_num3(double rl, double rr, double r0, double ll, double lr, double l0, double dt, const Gas3& g) {
  if (g.tau_type == 0) {
    if (g.c1 <= 0 && g.c2 <= 0) return 0.0;
    const double C = G3_DIV(g.c2 * fabs(G3_DIV(rl, ll) - G3_DIV(rr, lr)), fabs(G3_DIV(rl, ll) + G3_DIV(rr, lr)));
    return g.c1 * dt + dt * C;
  }
  double tau_n = G3_DIV(g.c2 * fabs(G3_DIV(rl, ll) - G3_DIV(rr, lr)), fabs(G3_DIV(rl, ll) + G3_DIV(rr, lr))) * dt;
  if (tau_n != tau_n) tau_n = 0.0;
  if ((g.mu > 0.0 && g.nu > 0.0) || (g.mu < 0.0 && g.nu < 0.0)) return 0.0;
  if (g.mu > 0.0) return tau_n + G3_DIV(2.0 * g.mu * l0, r0);
  if (g.nu > 0.0) return tau_n + 2 * g.nu * l0;
  return 0.0;
}
// the gks2nd time coefficients, flux_function.cpp:899-909
GKS3_DI void time_coe3(double* t, double tau, double tn, double dt) {
  const double eta = exp(-dt / tn);
  t[0] = tn * (1 - eta);
  t[1] = tn * (eta * (dt + tn) - tn) + tau * tn * (eta - 1);
  t[2] = tau * tn * (eta - 1);
  t[3] = tn * eta + dt - tn;
  t[4] = tn * (tn - eta * (dt + tn) - tau * (eta - 1)) - dt * tau;
  t[5] = 0.5 * dt * dt - tau * tn * (eta - 1) - tau * dt;
}

// A Gauss point's one-sided state in the face frame: convar, der1x, der1y,
// der1z (5 each), read through a strided view.
struct Side3 {
  const double* p;
  long long s;
  GKS3_DI double operator[](int k) const { return p[k * s]; }
};

// the half-space moments of one side from its order-0 and order-1 entries
// (the rest of mm3_side, without the erfc / exp)
GKS3_DI void mm3_side_from01(MM3& m, double uh0, double uh1, double u, double v, double w, double lambda,
                             const Gas3& g) {
  const double ol = G3_RCP(lambda);
  m.uw[0] = 1;
  m.uw[1] = u;
  m.uh[0] = uh0;
  m.uh[1] = uh1;
#pragma unroll
  for (int i = 2; i <= 6; i++) m.uw[i] = u * m.uw[i - 1] + 0.5 * (i - 1) * ol * m.uw[i - 2];
#pragma unroll
  for (int i = 2; i <= 6; i++) m.uh[i] = u * m.uh[i - 1] + 0.5 * (i - 1) * ol * m.uh[i - 2];
  m.vw[0] = 1;
  m.vw[1] = v;
  m.ww[0] = 1;
  m.ww[1] = w;
#pragma unroll
  for (int i = 2; i <= 5; i++) {
    m.vw[i] = v * m.vw[i - 1] + 0.5 * (i - 1) * ol * m.vw[i - 2];
    m.ww[i] = w * m.ww[i - 1] + 0.5 * (i - 1) * ol * m.ww[i - 2];
  }
  m.xi2 = 0.5 * g.K * ol;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) * ol * ol;
}
// the order-0 and order-1 half-space u moments (MMDF3d_left / _right,
// basic_function.cpp:189-290): side 0 integrates u > 0, side 1 u < 0
GKS3_DI void half01(double& uh0, double& uh1, int side, double u, double lambda) {
  const double pi = 3.14159265358979323846;
  const double beta = G3_DIV(exp(-lambda * u * u), sqrt(pi * lambda));
  if (side == 0) {
    uh0 = 0.5 * erfc(sqrt(lambda) * (-u));
    uh1 = u * uh0 + 0.5 * beta;
  } else {
    uh0 = 0.5 * erfc(sqrt(lambda) * u);
    uh1 = u * uh0 - 0.5 * beta;
  }
}
// one side's share of the collision state (Collision, basic_function.h:203-212)
GKS3_DI void collision_share(double* c, double rho, const MM3& m) {
  c[0] = rho * m.uh[0];
  c[1] = rho * m.uh[1];
  c[2] = rho * m.vw[1] * m.uh[0];
  c[3] = rho * m.ww[1] * m.uh[0];
  c[4] = 0.5 * rho * (m.uh[2] + m.uh[0] * m.vw[2] + m.uh[0] * m.ww[2] + m.uh[0] * m.xi2);
}

// GKS3D with gks2nd_3d and the S2O4 (dt, dt/2) combination for one Gauss
// point (flux_function.cpp:813-1163), lambda from the conservative state.
// Evaluated in three phases so that one Maxwellian's moments are live at a
// time: (0) both sides' order-0/1 moments, the collision state, tau and
// both time-coefficient sets, g0u; (A) the left side's moment vectors;
// (B) the right side's, folded with the left ones into the dt and dt/2
// fluxes F0, F1 in the reference's summation order (F0: t0, t3, t1, t2,
// t4, t5 terms; F1: t0 ... t5); (C) the collision state's slopes.
// center (optional): the collision state and its three derivatives.
template <class S>
GKS3_DI void gks3d_point_dev(double* f, double* derf, double* center, const S& L, const S& R, double dt,
                             const Gas3& g) {
  double pl[5], pr[5];
  {
    double c[5];
#pragma unroll
    for (int v = 0; v < 5; v++) c[v] = L[v];
    ulambda3(pl, c, g);
#pragma unroll
    for (int v = 0; v < 5; v++) c[v] = R[v];
    ulambda3(pr, c, g);
  }
  const double unit[5] = {1, 0.0, 0.0, 0.0, 0.0};
  // ---- phase 0
  double uhl0, uhl1, uhr0, uhr1;
  half01(uhl0, uhl1, 0, pl[1], pl[4]);
  half01(uhr0, uhr1, 1, pr[1], pr[4]);
  double c0[5];
  {
    MM3 m;
    double cl[5], cr[5];
    mm3_side_from01(m, uhl0, uhl1, pl[1], pl[2], pl[3], pl[4], g);
    collision_share(cl, pl[0], m);
    mm3_side_from01(m, uhr0, uhr1, pr[1], pr[2], pr[3], pr[4], g);
    collision_share(cr, pr[0], m);
#pragma unroll
    for (int v = 0; v < 5; v++) c0[v] = cl[v] + cr[v];
  }
  double p0[5];
  ulambda3(p0, c0, g);
  const double tau = tau_ns3(p0[0], p0[4], g);
  double t0[6], t1[6];
  time_coe3(t0, tau, tau_num3(pl[0], pr[0], p0[0], pl[4], pr[4], p0[4], dt, g), dt);
  const double dt2 = 0.5 * dt;
  time_coe3(t1, tau, tau_num3(pl[0], pr[0], p0[0], pl[4], pr[4], p0[4], dt2, g), dt2);
  double g0u[5];
  {
    MM3 m0;
    mm3_whole(m0, p0[1], p0[2], p0[3], p0[4], g);
    contract3(1, 0, 0, 0, g0u, unit, Tab3<false>{m0});
  }
  // ---- phase A: the left side
  double glu[5], T1L[5], Alul[5], c0d[3][5];
  {
    MM3 ml;
    mm3_side_from01(ml, uhl0, uhl1, pl[1], pl[2], pl[3], pl[4], g);
    const Tab3<true> tl{ml};
    const Tab3<false> wl{ml};
    contract3(1, 0, 0, 0, glu, unit, tl);
    double aul[5];
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      double der[5], al[5], tmp[5];
#pragma unroll
      for (int v = 0; v < 5; v++) der[v] = L[5 * (dd + 1) + v];
      micro_slope3(al, der, pl, g);
      contract3(0, 0, 0, 0, tmp, al, tl);
#pragma unroll
      for (int v = 0; v < 5; v++) c0d[dd][v] = pl[0] * tmp[v];
      contract3(1 + (dd == 0), dd == 1, dd == 2, 0, tmp, al, tl);
#pragma unroll
      for (int v = 0; v < 5; v++) T1L[v] = dd == 0 ? tmp[v] : T1L[v] + tmp[v];
      contract3(dd == 0, dd == 1, dd == 2, 0, tmp, al, wl);
#pragma unroll
      for (int v = 0; v < 5; v++) aul[v] = dd == 0 ? tmp[v] : aul[v] + tmp[v];
    }
    double der_AL[5], Al[5];
#pragma unroll
    for (int v = 0; v < 5; v++) der_AL[v] = -pl[0] * aul[v];
    micro_slope3(Al, der_AL, pl, g);
    contract3(1, 0, 0, 0, Alul, Al, tl);
  }
  // ---- phase B: the right side, folded with the left
  double F0[5], F1[5];
  {
    MM3 mr;
    mm3_side_from01(mr, uhr0, uhr1, pr[1], pr[2], pr[3], pr[4], g);
    const Tab3<true> tr{mr};
    const Tab3<false> wr{mr};
    {
      double gru[5];
      contract3(1, 0, 0, 0, gru, unit, tr);
#pragma unroll
      for (int i = 0; i < 5; i++) {
        F0[i] = pl[0] * t0[0] * glu[i] + pr[0] * t0[0] * gru[i];
        F0[i] = F0[i] + p0[0] * t0[3] * g0u[i];
        F1[i] = t1[0] * (pl[0] * glu[i] + pr[0] * gru[i]);
      }
    }
    double T1R[5], aur[5];
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      double der[5], ar[5], tmp[5];
#pragma unroll
      for (int v = 0; v < 5; v++) der[v] = R[5 * (dd + 1) + v];
      micro_slope3(ar, der, pr, g);
      contract3(0, 0, 0, 0, tmp, ar, tr);
#pragma unroll
      for (int v = 0; v < 5; v++) c0d[dd][v] = c0d[dd][v] + pr[0] * tmp[v];
      contract3(1 + (dd == 0), dd == 1, dd == 2, 0, tmp, ar, tr);
#pragma unroll
      for (int v = 0; v < 5; v++) T1R[v] = dd == 0 ? tmp[v] : T1R[v] + tmp[v];
      contract3(dd == 0, dd == 1, dd == 2, 0, tmp, ar, wr);
#pragma unroll
      for (int v = 0; v < 5; v++) aur[v] = dd == 0 ? tmp[v] : aur[v] + tmp[v];
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const double s1 = pl[0] * T1L[i] + pr[0] * T1R[i];
      F0[i] = F0[i] + t0[1] * s1;
      F1[i] = F1[i] + t1[1] * s1;
    }
    double der_AR[5], Ar[5], Arur[5];
#pragma unroll
    for (int v = 0; v < 5; v++) der_AR[v] = -pr[0] * aur[v];
    micro_slope3(Ar, der_AR, pr, g);
    contract3(1, 0, 0, 0, Arur, Ar, tr);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      const double s2 = pl[0] * Alul[i] + pr[0] * Arur[i];
      F0[i] = F0[i] + t0[2] * s2;
      F1[i] = F1[i] + t1[2] * s2;
      F1[i] = F1[i] + p0[0] * t1[3] * g0u[i];
    }
  }
  if (center) {
#pragma unroll
    for (int v = 0; v < 5; v++) {
      center[v] = c0[v];
      center[5 + v] = c0d[0][v];
      center[10 + v] = c0d[1][v];
      center[15 + v] = c0d[2][v];
    }
  }
  // ---- phase C: the collision state's slopes (t4) and time derivative (t5)
  {
    MM3 m0;
    mm3_whole(m0, p0[1], p0[2], p0[3], p0[4], g);
    const Tab3<false> w0{m0};
    double t4s[5], derA0[5];
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
      double a0[5], au[5], auu[5];
      micro_slope3(a0, c0d[dd], p0, g);
      contract3(dd == 0, dd == 1, dd == 2, 0, au, a0, w0);
      contract3(1 + (dd == 0), dd == 1, dd == 2, 0, auu, a0, w0);
#pragma unroll
      for (int i = 0; i < 5; i++) {
        t4s[i] = dd == 0 ? auu[i] : t4s[i] + auu[i];
        derA0[i] = dd == 0 ? au[i] : derA0[i] + au[i];
      }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
      F0[i] = F0[i] + p0[0] * t0[4] * t4s[i];
      F1[i] = F1[i] + p0[0] * t1[4] * t4s[i];
      derA0[i] = -p0[0] * derA0[i];
    }
    double A0[5], A0u[5];
    micro_slope3(A0, derA0, p0, g);
    contract3(1, 0, 0, 0, A0u, A0, w0);
#pragma unroll
    for (int i = 0; i < 5; i++) {
      F0[i] = F0[i] + p0[0] * t0[5] * (A0u[i]);
      F1[i] = F1[i] + p0[0] * t1[5] * (A0u[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < 5; i++) {
    f[i] = (4.0 * F1[i] - F0[i]);
    derf[i] = 4.0 * (F0[i] - 2.0 * F1[i]);
  }
}

}  // namespace GKS_VARIANT
}  // namespace gksd3
