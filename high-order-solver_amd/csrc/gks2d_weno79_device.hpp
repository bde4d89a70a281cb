// gks2d_weno79_device.hpp — device WENO7-AO / WENO9-AO reconstructions with
// the DF factor (gfx950, FP64), included by both kernel builds after
// gks2d_device.hpp.
//
// Candidate polynomials and smoothness indicators are tables of
// (coefficient, stencil offset) terms summed in the reference's term order
// (reconstruction.cpp: WENO7 :750-974, :1894-2028, :2867-3174; WENO9
// :1072-1364, :2063-2211, :3209-3609).  The tables are local constexpr
// arrays walked by fully unrolled loops, so every coefficient becomes an
// immediate operand.  In the fast build the divisions by constants and by
// the cell size become multiplications by reciprocals.
#pragma once
#include "gks2d_device.hpp"

namespace gksd {
inline namespace GKS_VARIANT {

struct WT {
  double c;
  int o;  // offset from the centre cell
};

// sum of the terms, left to right; w points at the centre of the stencil
template <int N> GKS_DI double tsum(const WT (&t)[N], const double* w) {
  double s = t[0].c == 1.0 ? w[t[0].o] : (t[0].c == -1.0 ? -w[t[0].o] : t[0].c * w[t[0].o]);
#pragma unroll
  for (int k = 1; k < N; k++) {
    if (t[k].c == 1.0) s = s + w[t[k].o];
    else if (t[k].c == -1.0) s = s - w[t[k].o];
    else s = s + t[k].c * w[t[k].o];
  }
  return s;
}

// x / h and x / (k h): divisions in the IEEE build, multiplications by the
// reciprocal cell size in the fast build (the functions below take both h
// and ih = 1/h)
#if GKS_FAST
#define GKS_DH(x) ((x) * ih)
#define GKS_DKH(x, k) ((x) * (ih * (1.0 / (k))))
#else
#define GKS_DH(x) ((x) / h)
#define GKS_DKH(x, k) ((x) / ((k) * h))
#endif

// 7th-order stencil indicator from four 4-point sub-indicators (13/3 and
// 781/20 are the integer quotients 4 and 39 in the reference)
GKS_DI double beta7_global(const double* w) {
  constexpr WT A1_0[4] = {{-19, -3}, {87, -2}, {-177, -1}, {109, 0}};
  constexpr WT A1_1[4] = {{11, -2}, {-63, -1}, {33, 0}, {19, 1}};
  constexpr WT A1_2[4] = {{-19, -1}, {-33, 0}, {63, 1}, {-11, 2}};
  constexpr WT A1_3[4] = {{-109, 0}, {177, 1}, {-87, 2}, {19, 3}};
  constexpr WT A2_0[4] = {{-1, -3}, {4, -2}, {-5, -1}, {2, 0}};
  constexpr WT A2_12[3] = {{1, -1}, {-2, 0}, {1, 1}};
  constexpr WT A2_3[4] = {{2, 0}, {-5, 1}, {4, 2}, {-1, 3}};
  constexpr WT A3_0[4] = {{-1, -3}, {3, -2}, {-3, -1}, {1, 0}};
  constexpr WT A3_1[4] = {{-1, -2}, {3, -1}, {-3, 0}, {1, 1}};
  constexpr WT A3_2[4] = {{-1, -1}, {3, 0}, {-3, 1}, {1, 2}};
  constexpr WT A3_3[4] = {{-1, 0}, {3, 1}, {-3, 2}, {1, 3}};
  double a1[4] = {tsum(A1_0, w), tsum(A1_1, w), tsum(A1_2, w), tsum(A1_3, w)};
  double a2[4] = {tsum(A2_0, w), tsum(A2_12, w), tsum(A2_12, w), tsum(A2_3, w)};
  double a3[4] = {tsum(A3_0, w), tsum(A3_1, w), tsum(A3_2, w), tsum(A3_3, w)};
  double b[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double x1 = GKS_CDIV(a1[k], 60.0), x2 = GKS_CDIV(a2[k], 2.0), x3 = GKS_CDIV(a3[k], 6.0);
    const double t = x1 + GKS_CDIV(x3, 10.0);
    b[k] = t * t + 4.0 * x2 * x2 + 39.0 * x3 * x3;
  }
  return GKS_CDIV(1.0, 20.0) * (b[0] + 9.0 * b[1] + 9.0 * b[2] + b[3]) + fabs(b[0] - b[1] - b[2] + b[3]);
}

// five 5-point sub-indicators of the 9th-order stencil
GKS_DI void beta9_sub(double* bs, const double* w) {
  double a1[5], a2[5], a3[5], a4[5];
  {
    constexpr WT t0[5] = {{-462, -1}, {336, -2}, {-146, -3}, {27, -4}, {245, 0}};
    constexpr WT t1[5] = {{-192, -1}, {66, -2}, {-11, -3}, {110, 0}, {27, 1}};
    constexpr WT t2[4] = {{-82, -1}, {11, -2}, {82, 1}, {-11, 2}};
    constexpr WT t3[5] = {{-27, -1}, {-110, 0}, {192, 1}, {-66, 2}, {11, 3}};
    constexpr WT t4[5] = {{-245, 0}, {462, 1}, {-336, 2}, {146, 3}, {-27, 4}};
    a1[0] = tsum(t0, w), a1[1] = tsum(t1, w), a1[2] = tsum(t2, w), a1[3] = tsum(t3, w), a1[4] = tsum(t4, w);
  }
  {
    constexpr WT t0[5] = {{-240, -1}, {262, -2}, {-128, -3}, {25, -4}, {81, 0}};
    constexpr WT t1[5] = {{10, -1}, {12, -2}, {-3, -3}, {-44, 0}, {25, 1}};
    constexpr WT t2[5] = {{40, -1}, {-3, -2}, {-74, 0}, {40, 1}, {-3, 2}};
    constexpr WT t3[5] = {{25, -1}, {-44, 0}, {10, 1}, {12, 2}, {-3, 3}};
    constexpr WT t4[5] = {{81, 0}, {-240, 1}, {262, 2}, {-128, 3}, {25, 4}};
    a2[0] = tsum(t0, w), a2[1] = tsum(t1, w), a2[2] = tsum(t2, w), a2[3] = tsum(t3, w), a2[4] = tsum(t4, w);
  }
  {
    constexpr WT t0[5] = {{-18, -1}, {24, -2}, {-14, -3}, {3, -4}, {5, 0}};
    constexpr WT t1[5] = {{12, -1}, {-6, -2}, {1, -3}, {-10, 0}, {3, 1}};
    constexpr WT t2[4] = {{2, -1}, {-1, -2}, {-2, 1}, {1, 2}};
    constexpr WT t3[5] = {{-3, -1}, {10, 0}, {-12, 1}, {6, 2}, {-1, 3}};
    constexpr WT t4[5] = {{-5, 0}, {18, 1}, {-24, 2}, {14, 3}, {-3, 4}};
    a3[0] = tsum(t0, w), a3[1] = tsum(t1, w), a3[2] = tsum(t2, w), a3[3] = tsum(t3, w), a3[4] = tsum(t4, w);
  }
  {
    constexpr WT t0[5] = {{-4, -1}, {6, -2}, {-4, -3}, {1, -4}, {1, 0}};
    constexpr WT t1[5] = {{6, -1}, {-4, -2}, {1, -3}, {-4, 0}, {1, 1}};
    constexpr WT t2[5] = {{-4, -1}, {1, -2}, {6, 0}, {-4, 1}, {1, 2}};
    constexpr WT t3[5] = {{1, -1}, {-4, 0}, {6, 1}, {-4, 2}, {1, 3}};
    constexpr WT t4[5] = {{1, 0}, {-4, 1}, {6, 2}, {-4, 3}, {1, 4}};
    a4[0] = tsum(t0, w), a4[1] = tsum(t1, w), a4[2] = tsum(t2, w), a4[3] = tsum(t3, w), a4[4] = tsum(t4, w);
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double x1 = GKS_CDIV(a1[k], 120.0), x2 = GKS_CDIV(a2[k], 56.0), x3 = GKS_CDIV(a3[k], 12.0),
                 x4 = GKS_CDIV(a4[k], 24.0);
    const double t1 = x1 + GKS_CDIV(x3, 10.0);
    const double t2 = x2 + 123.0 / 455.0 * x4;
    bs[k] = t1 * t1 + 13.0 / 3.0 * t2 * t2 + 781.0 / 20.0 * x3 * x3 + 1421461.0 / 2275.0 * x4 * x4;
  }
}

// final AO weights of P(7,5,3,3,3) (NS 5) or P(9,7,5,3,3,3) (NS 6)
template <int ORDER> GKS_DI void ao_weights79(double* fw, const double* w) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  constexpr double h85 = 0.85;
  double d[NS], beta[NS];
  {
    double q, q2;
    q = w[-2] - 2.0 * w[-1] + w[0];
    q2 = w[-2] - 4.0 * w[-1] + 3.0 * w[0];
    beta[0] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
    q = w[-1] - 2.0 * w[0] + w[1];
    q2 = w[-1] - w[1];
    beta[1] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
    q = w[0] - 2.0 * w[1] + w[2];
    q2 = 3.0 * w[0] - 4.0 * w[1] + w[2];
    beta[2] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  }
  if constexpr (ORDER == 7) {
    d[0] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) / 2.0;
    d[1] = (1.0 - h85) * (1.0 - h85) * h85;
    d[2] = d[0];
    d[3] = (1.0 - h85) * h85;
    d[4] = h85;
    beta[3] = GKS_CDIV(1.0, 6.0) * (beta[0] + 4.0 * beta[1] + beta[2]) + fabs(beta[0] - beta[2]);
    beta[4] = beta7_global(w);
  } else {
    d[0] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) * (1.0 - h85) / 2.0;
    d[1] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) * h85;
    d[2] = d[0];
    d[3] = (1.0 - h85) * (1.0 - h85) * h85;
    d[4] = (1.0 - h85) * h85;
    d[5] = h85;
    double bs[5];
    beta9_sub(bs, w);
    beta[3] = bs[2];
    beta[4] = beta7_global(w);
    beta[5] = GKS_CDIV(1.0, 70.0) * (bs[0] + 16.0 * bs[1] + 36.0 * bs[2] + 16.0 * bs[3] + bs[4]) +
              fabs(bs[0] - bs[1] - bs[3] + bs[4]);
  }
  const double top = beta[NS - 1];
  double ts = fabs(top - beta[0]);
#pragma unroll
  for (int k = 1; k < NS - 1; k++) ts = ts + fabs(top - beta[k]);
  const double tau = (ORDER == 7 ? 0.25 : 0.2) * ts;
  double alpha[NS], sum = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const double gd = tau / (beta[k] + 1e-6);
    alpha[k] = d[k] * (1.0 + gd * gd);
    sum += alpha[k];
  }
  double ww[NS];
#if GKS_FAST
  const double isum = 1.0 / sum;
#pragma unroll
  for (int k = 0; k < NS; k++) ww[k] = alpha[k] * isum;
#else
#pragma unroll
  for (int k = 0; k < NS; k++) ww[k] = alpha[k] / sum;
#endif
  const double wt = GKS_CDIV(ww[NS - 1], d[NS - 1]);
  fw[NS - 1] = wt;
#pragma unroll
  for (int k = 0; k < NS - 1; k++) fw[k] = ww[k] - wt * d[k];
}

// DF factors of the NS sub-stencils (a centred on the cell)
template <int ORDER> GKS_DI void df79(double* df, const double* a) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  double s[NS];
  s[0] = a[-2] + a[0];
  s[1] = a[-1] + a[1];
  s[2] = a[0] + a[2];
  s[3] = a[-2] + a[0] + a[2];
  s[4] = a[-3] + a[-1] + a[1] + a[3];
  if constexpr (ORDER == 9) s[5] = a[-4] + a[-2] + a[0] + a[2] + a[4];
#pragma unroll
  for (int i = 0; i < NS; i++) df[i] = s[i] < 2.0 ? 1.0 : 2.0 / s[i];
}

// value and derivative at the cell's left (side 0) / right (side 1) face
template <int ORDER, int SIDE>
GKS_DI void weno79_face(double& var, double& der1, const double* w, const double* df, double h, double ih) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  double fw[NS];
  ao_weights79<ORDER>(fw, w);
  const double w0 = w[0];
  double p[NS], px[NS];
  (void)h;
  (void)ih;
  if constexpr (SIDE == 0) {
    constexpr WT P0[4] = {{-1.0 / 6.0, -2}, {5.0 / 6.0, -1}, {1.0 / 3.0, 0}, {-1, 0}};
    constexpr WT P1[4] = {{1.0 / 3.0, -1}, {5.0 / 6.0, 0}, {-1.0 / 6.0, 1}, {-1, 0}};
    constexpr WT P2[4] = {{11.0 / 6.0, 0}, {-7.0 / 6.0, 1}, {1.0 / 3.0, 2}, {-1, 0}};
    constexpr WT P3[5] = {{47, 0}, {27, -1}, {-3, -2}, {-13, 1}, {2, 2}};
    constexpr WT P4[7] = {{319, 0}, {214, -1}, {-38, -2}, {4, -3}, {-101, 1}, {25, 2}, {-3, 3}};
    constexpr WT X0[2] = {{1, 0}, {-1, -1}};
    constexpr WT X2[3] = {{2, 0}, {-3, 1}, {1, 2}};
    constexpr WT X3[4] = {{15, 0}, {-15, -1}, {1, -2}, {-1, 1}};
    constexpr WT X4[6] = {{245, 0}, {-245, -1}, {25, -2}, {-2, -3}, {-25, 1}, {2, 2}};
    p[0] = w0 + df[0] * tsum(P0, w);
    p[1] = w0 + df[1] * tsum(P1, w);
    p[2] = w0 + df[2] * tsum(P2, w);
    p[3] = w0 + df[3] * ((1.0 / 60.0) * tsum(P3, w) - w0);
    p[4] = w0 + df[4] * ((1.0 / 420.0) * tsum(P4, w) - w0);
    px[0] = GKS_DH(df[0] * tsum(X0, w));
    px[1] = GKS_DH(df[1] * tsum(X0, w));
    px[2] = -df[2] * GKS_DH(tsum(X2, w));
    px[3] = GKS_DKH(df[3] * tsum(X3, w), 12.0);
    px[4] = GKS_DKH(df[4] * tsum(X4, w), 180.0);
    if constexpr (ORDER == 9) {
      constexpr WT P5[9] = {{1879, 0}, {1375, -1}, {-305, -2}, {55, -3}, {-5, -4}, {-641, 1}, {199, 2}, {-41, 3}, {4, 4}};
      constexpr WT X5[8] = {{7175, 0}, {-7175, -1}, {889, -2}, {-119, -3}, {9, -4}, {-889, 1}, {119, 2}, {-9, 3}};
      p[5] = w0 + df[5] * ((1.0 / 2520.0) * tsum(P5, w) - w0);
      px[5] = GKS_DKH(df[5] * tsum(X5, w), 5040.0);
    }
  } else {
    constexpr WT P0[4] = {{1.0 / 3.0, -2}, {-7.0 / 6.0, -1}, {11.0 / 6.0, 0}, {-1, 0}};
    constexpr WT P1[4] = {{-1.0 / 6.0, -1}, {5.0 / 6.0, 0}, {1.0 / 3.0, 1}, {-1, 0}};
    constexpr WT P2[4] = {{1.0 / 3.0, 0}, {5.0 / 6.0, 1}, {-1.0 / 6.0, 2}, {-1, 0}};
    constexpr WT P3[5] = {{47, 0}, {-13, -1}, {2, -2}, {27, 1}, {-3, 2}};
    constexpr WT P4[7] = {{319, 0}, {-101, -1}, {25, -2}, {-3, -3}, {214, 1}, {-38, 2}, {4, 3}};
    constexpr WT X0[3] = {{2, 0}, {-3, -1}, {1, -2}};
    constexpr WT X1[2] = {{-1, 0}, {1, 1}};
    constexpr WT X3[4] = {{-15, 0}, {1, -1}, {15, 1}, {-1, 2}};
    constexpr WT X4[6] = {{-245, 0}, {25, -1}, {-2, -2}, {245, 1}, {-25, 2}, {2, 3}};
    p[0] = w0 + df[0] * tsum(P0, w);
    p[1] = w0 + df[1] * tsum(P1, w);
    p[2] = w0 + df[2] * tsum(P2, w);
    p[3] = w0 + df[3] * ((1.0 / 60.0) * tsum(P3, w) - w0);
    p[4] = w0 + df[4] * ((1.0 / 420.0) * tsum(P4, w) - w0);
    px[0] = GKS_DH(df[0] * tsum(X0, w));
    px[1] = GKS_DH(df[1] * tsum(X1, w));
    px[2] = GKS_DH(df[2] * tsum(X1, w));
    px[3] = GKS_DKH(df[3] * tsum(X3, w), 12.0);
    px[4] = GKS_DKH(df[4] * tsum(X4, w), 180.0);
    if constexpr (ORDER == 9) {
      constexpr WT P5[9] = {{1879, 0}, {-641, -1}, {199, -2}, {-41, -3}, {4, -4}, {1375, 1}, {-305, 2}, {55, 3}, {-5, 4}};
      constexpr WT X5[8] = {{-7175, 0}, {889, -1}, {-119, -2}, {9, -3}, {7175, 1}, {-889, 2}, {119, 3}, {-9, 4}};
      p[5] = w0 + df[5] * ((1.0 / 2520.0) * tsum(P5, w) - w0);
      px[5] = GKS_DKH(df[5] * tsum(X5, w), 5040.0);
    }
  }
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}

// WENO7: 3 Gauss points (face centre, then -/+ sqrt(3/5)/2 of the face)
GKS_DI void weno7_gauss(double* g, double* gx, const double* w, const double* df, double h, double ih) {
  (void)h;
  (void)ih;
  double fw[5];
  ao_weights79<7>(fw, w);
  const double w0 = w[0];
  constexpr double s = 3.872983346207417;  // sqrt(15) as the reference writes it
  double p[5], px[5];
  {
    constexpr WT P0[3] = {{23, 0}, {2, -1}, {-1, -2}};
    constexpr WT P1[3] = {{26, 0}, {-1, -1}, {-1, 1}};
    constexpr WT P2[3] = {{23, 0}, {2, 1}, {-1, 2}};
    constexpr WT P3[5] = {{2134, 0}, {-116, -1}, {9, -2}, {-116, 1}, {9, 2}};
    constexpr WT P4[7] = {{121004, 0}, {-7621, -1}, {954, -2}, {-75, -3}, {-7621, 1}, {954, 2}, {-75, 3}};
    constexpr WT X0[3] = {{3, 0}, {-4, -1}, {1, -2}};
    constexpr WT X1[2] = {{-1, -1}, {1, 1}};
    constexpr WT X2[3] = {{-3, 0}, {4, 1}, {-1, 2}};
    constexpr WT X3[4] = {{-34, -1}, {5, -2}, {34, 1}, {-5, 2}};
    constexpr WT X4[6] = {{-9455, -1}, {2236, -2}, {-259, -3}, {9455, 1}, {-2236, 2}, {259, 3}};
    p[0] = w0 + df[0] * (GKS_CDIV(tsum(P0, w), 24.0) - w0);
    p[1] = w0 + df[1] * (GKS_CDIV(tsum(P1, w), 24.0) - w0);
    p[2] = w0 + df[2] * (GKS_CDIV(tsum(P2, w), 24.0) - w0);
    p[3] = w0 + df[3] * (GKS_CDIV(tsum(P3, w), 1920.0) - w0);
    p[4] = w0 + df[4] * (GKS_CDIV(tsum(P4, w), 107520.0) - w0);
    px[0] = GKS_DKH(df[0] * tsum(X0, w), 2.0);
    px[1] = GKS_DKH(df[1] * tsum(X1, w), 2.0);
    px[2] = GKS_DKH(df[2] * tsum(X2, w), 2.0);
    px[3] = GKS_DKH(df[3] * tsum(X3, w), 48.0);
    px[4] = GKS_DKH(df[4] * tsum(X4, w), 11520.0);
    g[0] = 0.0;
    gx[0] = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[0] += fw[k] * p[k];
      gx[0] += fw[k] * px[k];
    }
  }
  {
    constexpr WT A0[3] = {{62 - 9 * s, 0}, {4 * (-1 + 3 * s), -1}, {2 - 3 * s, -2}};
    constexpr WT A1[3] = {{56, 0}, {2 + 3 * s, -1}, {2 - 3 * s, 1}};
    constexpr WT A2[3] = {{62 + 9 * s, 0}, {-(4 * (1 + 3 * s)), 1}, {2 + 3 * s, 2}};
    constexpr WT A3[8] = {{2186, 0}, {4 * (29 + 41 * s), -1}, {-9, -2}, {-(22 * s), -2},
                          {116, 1}, {-(164 * s), 1}, {-9, 2}, {22 * s, 2}};
    constexpr WT A4[12] = {{4534440, 0}, {5 * (57144 + 78421 * s), -1}, {-35748, -2}, {-(84364 * s), -2},
                           {2808, -3}, {9541 * s, -3}, {285720, 1}, {-(392105 * s), 1},
                           {-35748, 2}, {84364 * s, 2}, {2808, 3}, {-(9541 * s), 3}};
    constexpr WT B0[3] = {{-(-15 + s), 0}, {2 * (-10 + s), -1}, {-(-5 + s), -2}};
    constexpr WT B1[3] = {{2 * s, 0}, {-(5 + s), -1}, {-(-5 + s), 1}};
    constexpr WT B2[3] = {{-(15 + s), 0}, {2 * (10 + s), 1}, {-(5 + s), 2}};
    constexpr WT B3[8] = {{78 * s, 0}, {-(2 * (95 + 21 * s)), -1}, {20, -2}, {3 * s, -2},
                          {190, 1}, {-(42 * s), 1}, {-20, 2}, {3 * s, 2}};
    constexpr WT B4[12] = {{103860 * s, 0}, {-(5 * (49925 + 11619 * s)), -1}, {41300, -2}, {6678 * s, -2},
                           {-4325, -3}, {-(513 * s), -3}, {249625, 1}, {-(58095 * s), 1},
                           {-41300, 2}, {6678 * s, 2}, {4325, 3}, {-(513 * s), 3}};
    p[0] = w0 + df[0] * (GKS_CDIV(tsum(A0, w), 60.0) - w0);
    p[1] = w0 + df[1] * (GKS_CDIV(tsum(A1, w), 60.0) - w0);
    p[2] = w0 + df[2] * (GKS_CDIV(tsum(A2, w), 60.0) - w0);
    p[3] = w0 + df[3] * (GKS_CDIV(tsum(A3, w), 2400.0) - w0);
    p[4] = w0 + df[4] * (GKS_CDIV(tsum(A4, w), 5040000.0) - w0);
    px[0] = GKS_DKH(df[0] * tsum(B0, w), 10.0);
    px[1] = GKS_DKH(df[1] * tsum(B1, w), 10.0);
    px[2] = GKS_DKH(df[2] * tsum(B2, w), 10.0);
    px[3] = GKS_DKH(df[3] * tsum(B3, w), 300.0);
    px[4] = GKS_DKH(df[4] * tsum(B4, w), 360000.0);
    g[1] = 0.0;
    gx[1] = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[1] += fw[k] * p[k];
      gx[1] += fw[k] * px[k];
    }
  }
  {
    constexpr WT A0[3] = {{62 + 9 * s, 0}, {-(4 * (1 + 3 * s)), -1}, {2 + 3 * s, -2}};
    constexpr WT A1[3] = {{56, 0}, {2 - 3 * s, -1}, {2 + 3 * s, 1}};
    constexpr WT A2[3] = {{62 - 9 * s, 0}, {4 * (-1 + 3 * s), 1}, {2 - 3 * s, 2}};
    constexpr WT A3[8] = {{2186, 0}, {-(4 * (-29 + 41 * s)), -1}, {-9, -2}, {22 * s, -2},
                          {116, 1}, {164 * s, 1}, {-9, 2}, {-(22 * s), 2}};
    constexpr WT A4[12] = {{4534440, 0}, {-(5 * (-57144 + 78421 * s)), -1}, {-35748, -2}, {84364 * s, -2},
                           {2808, -3}, {-(9541 * s), -3}, {285720, 1}, {392105 * s, 1},
                           {-35748, 2}, {-(84364 * s), 2}, {2808, 3}, {9541 * s, 3}};
    constexpr WT B0[3] = {{15 + s, 0}, {-(2 * (10 + s)), -1}, {5 + s, -2}};
    constexpr WT B1[3] = {{-(2 * s), 0}, {-(5 - s), -1}, {5 + s, 1}};
    constexpr WT B2[3] = {{-15 + s, 0}, {-(2 * (-10 + s)), 1}, {-5 + s, 2}};
    constexpr WT B3[8] = {{-(78 * s), 0}, {2 * (-95 + 21 * s), -1}, {20, -2}, {-(3 * s), -2},
                          {190, 1}, {42 * s, 1}, {-20, 2}, {-(3 * s), 2}};
    constexpr WT B4[12] = {{-(103860 * s), 0}, {5 * (-49925 + 11619 * s), -1}, {41300, -2}, {-(6678 * s), -2},
                           {-4325, -3}, {513 * s, -3}, {249625, 1}, {58095 * s, 1},
                           {-41300, 2}, {-(6678 * s), 2}, {4325, 3}, {513 * s, 3}};
    p[0] = w0 + df[0] * (GKS_CDIV(tsum(A0, w), 60.0) - w0);
    p[1] = w0 + df[1] * (GKS_CDIV(tsum(A1, w), 60.0) - w0);
    p[2] = w0 + df[2] * (GKS_CDIV(tsum(A2, w), 60.0) - w0);
    p[3] = w0 + df[3] * (GKS_CDIV(tsum(A3, w), 2400.0) - w0);
    p[4] = w0 + df[4] * (GKS_CDIV(tsum(A4, w), 5040000.0) - w0);
    px[0] = GKS_DKH(df[0] * tsum(B0, w), 10.0);
    px[1] = GKS_DKH(df[1] * tsum(B1, w), 10.0);
    px[2] = GKS_DKH(df[2] * tsum(B2, w), 10.0);
    px[3] = GKS_DKH(df[3] * tsum(B3, w), 300.0);
    px[4] = GKS_DKH(df[4] * tsum(B4, w), 360000.0);
    g[2] = 0.0;
    gx[2] = 0.0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      g[2] += fw[k] * p[k];
      gx[2] += fw[k] * px[k];
    }
  }
}

// one Polynomial_9th candidate: v = sum c_n x^n, dv = sum n c_n x^(n-1),
// powers formed as ((c*x)*x)... as the reference writes them
template <int NC>
GKS_DI void poly_cand(double& p, double& px, const double* cf, double x, double w0, double df, double h,
                      double ih) {
  (void)h;
  (void)ih;
  double v = cf[0];
#pragma unroll
  for (int n = 1; n < NC; n++) {
    double m = cf[n];
#pragma unroll
    for (int k = 0; k < n; k++) m = m * x;
    v = v + m;
  }
  p = w0 + df * (v - w0);
  double dv = cf[1];
#pragma unroll
  for (int n = 2; n < NC; n++) {
    double m = double(n) * cf[n];
#pragma unroll
    for (int k = 0; k < n - 1; k++) m = m * x;
    dv = dv + m;
  }
  px = GKS_DH(df * dv);
}

// WENO9: 4 Lobatto points at x = -1, 0, -1/2 -/+ sqrt(5)/10 of the face
// (Polynomial_9th, reconstruction.cpp:3551-3609)
GKS_DI void weno9_gauss(double* g, double* gx, const double* w, const double* df, double h, double ih) {
  double fw[6];
  ao_weights79<9>(fw, w);
  const double w0 = w[0];
  // x-independent polynomial coefficients of the six candidates
  double c3[3][3], c5[5], c7[7], c9[9];
  {
    constexpr WT A0[3] = {{11, 0}, {-7, -1}, {2, -2}};
    constexpr WT A1[3] = {{5, 0}, {-1, -1}, {2, 1}};
    constexpr WT A2[3] = {{2, 0}, {5, 1}, {-1, 2}};
    constexpr WT B0[3] = {{2, 0}, {-3, -1}, {1, -2}};
    constexpr WT B12[2] = {{-1, 0}, {1, 1}};
    constexpr WT C0[3] = {{1, 0}, {-2, -1}, {1, -2}};
    constexpr WT C1[3] = {{-2, 0}, {1, -1}, {1, 1}};
    constexpr WT C2[3] = {{1, 0}, {-2, 1}, {1, 2}};
    c3[0][0] = GKS_CDIV(tsum(A0, w), 6.0), c3[0][1] = tsum(B0, w), c3[0][2] = GKS_CDIV(tsum(C0, w), 2.0);
    c3[1][0] = GKS_CDIV(tsum(A1, w), 6.0), c3[1][1] = tsum(B12, w), c3[1][2] = GKS_CDIV(tsum(C1, w), 2.0);
    c3[2][0] = GKS_CDIV(tsum(A2, w), 6.0), c3[2][1] = tsum(B12, w), c3[2][2] = GKS_CDIV(tsum(C2, w), 2.0);
  }
  {
    constexpr WT T0[5] = {{47, 0}, {-13, -1}, {2, -2}, {27, 1}, {-3, 2}};
    constexpr WT T1[4] = {{-15, 0}, {1, -1}, {15, 1}, {-1, 2}};
    constexpr WT T2[5] = {{-8, 0}, {6, -1}, {-1, -2}, {2, 1}, {1, 2}};
    constexpr WT T3[4] = {{3, 0}, {-1, -1}, {-3, 1}, {1, 2}};
    constexpr WT T4[5] = {{6, 0}, {-4, -1}, {1, -2}, {-4, 1}, {1, 2}};
    c5[0] = GKS_CDIV(tsum(T0, w), 60.0), c5[1] = GKS_CDIV(tsum(T1, w), 12.0), c5[2] = GKS_CDIV(tsum(T2, w), 8.0);
    c5[3] = GKS_CDIV(tsum(T3, w), 6.0), c5[4] = GKS_CDIV(tsum(T4, w), 24.0);
  }
  {
    constexpr WT T0[7] = {{319, 0}, {-101, -1}, {25, -2}, {-3, -3}, {214, 1}, {-38, 2}, {4, 3}};
    constexpr WT T1[6] = {{-245, 0}, {25, -1}, {-2, -2}, {245, 1}, {-25, 2}, {2, 3}};
    constexpr WT T2[7] = {{-230, 0}, {210, -1}, {-57, -2}, {7, -3}, {15, 1}, {63, 2}, {-8, 3}};
    constexpr WT T3[6] = {{28, 0}, {-11, -1}, {1, -2}, {-28, 1}, {11, 2}, {-1, 3}};
    constexpr WT T4[7] = {{46, 0}, {-39, -1}, {15, -2}, {-2, -3}, {-24, 1}, {3, 2}, {1, 3}};
    constexpr WT T5[6] = {{-10, 0}, {5, -1}, {-1, -2}, {10, 1}, {-5, 2}, {1, 3}};
    constexpr WT T6[7] = {{-20, 0}, {15, -1}, {-6, -2}, {1, -3}, {15, 1}, {-6, 2}, {1, 3}};
    c7[0] = GKS_CDIV(tsum(T0, w), 420.0), c7[1] = GKS_CDIV(tsum(T1, w), 180.0);
    c7[2] = GKS_CDIV(tsum(T2, w), 240.0), c7[3] = GKS_CDIV(tsum(T3, w), 36.0);
    c7[4] = GKS_CDIV(tsum(T4, w), 144.0), c7[5] = GKS_CDIV(tsum(T5, w), 120.0);
    c7[6] = GKS_CDIV(tsum(T6, w), 720.0);
  }
  {
    constexpr WT T0[9] = {{1879, 0}, {-641, -1}, {199, -2}, {-41, -3}, {4, -4}, {1375, 1}, {-305, 2}, {55, 3}, {-5, 4}};
    constexpr WT T1[8] = {{-7175, 0}, {889, -1}, {-119, -2}, {9, -3}, {7175, 1}, {-889, 2}, {119, 3}, {-9, 4}};
    constexpr WT T2[9] = {{-27895, 0}, {28679, -1}, {-9835, -2}, {2081, -3}, {-205, -4},
                          {-2065, 1}, {11459, 2}, {-2455, 3}, {236, 4}};
    constexpr WT T3[8] = {{1365, 0}, {-587, -1}, {89, -2}, {-7, -3}, {-1365, 1}, {587, 2}, {-89, 3}, {7, 4}};
    constexpr WT T4[9] = {{1174, 0}, {-1160, -1}, {556, -2}, {-128, -3}, {13, -4}, {-464, 1}, {-68, 2}, {88, 3}, {-11, 4}};
    constexpr WT T5[8] = {{-75, 0}, {41, -1}, {-11, -2}, {1, -3}, {75, 1}, {-41, 2}, {11, 3}, {-1, 4}};
    constexpr WT T6[9] = {{-380, 0}, {334, -1}, {-170, -2}, {46, -3}, {-5, -4}, {250, 1}, {-86, 2}, {10, 3}, {1, 4}};
    constexpr WT T7[8] = {{35, 0}, {-21, -1}, {7, -2}, {-1, -3}, {-35, 1}, {21, 2}, {-7, 3}, {1, 4}};
    constexpr WT T8[9] = {{70, 0}, {-56, -1}, {28, -2}, {-8, -3}, {1, -4}, {-56, 1}, {28, 2}, {-8, 3}, {1, 4}};
    c9[0] = GKS_CDIV(tsum(T0, w), 2520.0), c9[1] = GKS_CDIV(tsum(T1, w), 5040.0);
    c9[2] = GKS_CDIV(tsum(T2, w), 30240.0), c9[3] = GKS_CDIV(tsum(T3, w), 1440.0);
    c9[4] = GKS_CDIV(tsum(T4, w), 3456.0), c9[5] = GKS_CDIV(tsum(T5, w), 480.0);
    c9[6] = GKS_CDIV(tsum(T6, w), 8640.0), c9[7] = GKS_CDIV(tsum(T7, w), 5040.0);
    c9[8] = GKS_CDIV(tsum(T8, w), 40320.0);
  }
  constexpr double sqrt5 = 2.236067977499790;
  const double xs[4] = {-1.0, 0.0, -0.5 - sqrt5 / 10.0, -0.5 + sqrt5 / 10.0};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double p[6], px[6];
#pragma unroll
    for (int k = 0; k < 3; k++) poly_cand<3>(p[k], px[k], c3[k], xs[q], w0, df[k], h, ih);
    poly_cand<5>(p[3], px[3], c5, xs[q], w0, df[3], h, ih);
    poly_cand<7>(p[4], px[4], c7, xs[q], w0, df[4], h, ih);
    poly_cand<9>(p[5], px[5], c9, xs[q], w0, df[5], h, ih);
    g[q] = 0.0;
    gx[q] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      g[q] += fw[k] * p[k];
      gx[q] += fw[k] * px[k];
    }
  }
}
#undef GKS_DH
#undef GKS_DKH

template <int ORDER>
GKS_DI void weno79_gauss(double* g, double* gx, const double* w, const double* df, double h, double ih) {
  if constexpr (ORDER == 7) weno7_gauss(g, gx, w, df, h, ih);
  else weno9_gauss(g, gx, w, df, h, ih);
}

}  // namespace GKS_VARIANT
}  // namespace gksd
