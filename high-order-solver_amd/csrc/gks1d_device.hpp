// gks1d_device.hpp — device numerics of the 1D GKS path (gfx950, FP64).
// Restates the reference's 1D leaves with their operation order; included
// after gks2d_device.hpp by both kernel builds (it reuses the WENO5-AO-DF
// leaves, the collision-time and time-coefficient helpers from there).
#pragma once
#include "gks2d_device.hpp"

namespace gksd {
inline namespace GKS_VARIANT {

// Lambda (1D), Convar_to_primvar_1D, Convar_to_ULambda_1d:
// basic_function.cpp:804-809, 828-833, 1169-1172
GKS_DI double lambda1(double rho, double u, double rhoE, const Gas& g) {
  return (g.K + 1.0) * 0.25 * (rho / (rhoE - 0.5 * rho * (u * u)));
}
GKS_DI void prim1(double* p, const double* c, const Gas& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = (g.gamma - 1.0) * (c[2] - 0.5 * c[1] * c[1] / c[0]);
}
GKS_DI void ulambda1(double* p, const double* c, const Gas& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = lambda1(c[0], p[1], c[2], g);
}
// Convar_to_char1D / Char_to_convar1D: basic_function.cpp:868-921
GKS_DI void to_char1(double* ch, const double* prim, const double* c, const Gas& g) {
  const double cs = sqrt(g.gamma * prim[2] / prim[0]);
  const double alfa = (g.gamma - 1.0) / (2.0 * cs * cs);
  const double u = prim[1];
  const double s[3][3] = {{alfa * (0.5 * u * u + u * cs / (g.gamma - 1.0)), alfa * (-u - cs / (g.gamma - 1.0)), alfa},
                          {alfa * (-u * u + 2.0 * cs * cs / (g.gamma - 1.0)), alfa * 2.0 * u, -2.0 * alfa},
                          {alfa * (0.5 * u * u - u * cs / (g.gamma - 1.0)), alfa * (-u + cs / (g.gamma - 1.0)), alfa}};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) a = a + s[i][j] * c[j];
    ch[i] = a;
  }
}
GKS_DI void from_char1(double* c, const double* prim, const double* ch, const Gas& g) {
  const double cs = sqrt(g.gamma * prim[2] / prim[0]);
  const double u = prim[1];
  const double h = 0.5 * u * u + cs * cs / (g.gamma - 1.0);
  const double s[3][3] = {{1.0, 1.0, 1.0}, {u - cs, u, u + cs}, {h - u * cs, u * u / 2.0, h + u * cs}};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) a = a + s[i][j] * ch[j];
    c[i] = a;
  }
}

// MMDF1d (basic_function.cpp:16-99): half and whole moments up to order 9,
// xi moments up to 6; table entries formed where used
struct MM1 {
  double up[10], un[10], uw[10], xi2, xi4, xi6;
};
GKS_DI void mmdf1(MM1& m, double u, double lambda, const Gas& g) {
  const double pi = 3.14159265358979323846;
  m.uw[0] = 1.0;
  m.uw[1] = u;
  const double sl = sqrt(lambda);
  m.up[0] = 0.5 * erfc(sl * (-u));
  m.un[0] = 0.5 * erfc(sl * u);
  const double beta = exp(-lambda * u * u) / sqrt(pi * lambda);
  m.up[1] = u * m.up[0] + 0.5 * beta;
  m.un[1] = u * m.un[0] - 0.5 * beta;
#pragma unroll
  for (int i = 2; i <= 9; i++) m.uw[i] = u * m.uw[i - 1] + 0.5 * (i - 1) / lambda * m.uw[i - 2];
#pragma unroll
  for (int i = 2; i <= 9; i++) {
    m.up[i] = u * m.up[i - 1] + 0.5 * (i - 1) / lambda * m.up[i - 2];
    m.un[i] = u * m.un[i - 1] + 0.5 * (i - 1) / lambda * m.un[i - 2];
  }
  m.xi2 = 0.5 * g.K / lambda;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) / (lambda * lambda);
  m.xi6 = 0.5 * (g.K + 4) / lambda * m.xi4;
}
struct Tab1 {
  const double* mom;
  const MM1& m;
  GKS_DI double operator()(int i, int k) const {
    return k == 0 ? mom[i] : mom[i] * (k == 1 ? m.xi2 : (k == 2 ? m.xi4 : m.xi6));
  }
};
// G / GL / GR (1D): basic_function.cpp:522-549
GKS_DI void contract1(int u, int x, double* psi, const double* a, const Tab1& M) {
  psi[0] = a[0] * M(u, x) + a[1] * M(u + 1, x) + a[2] * 0.5 * (M(u + 2, x) + M(u, x + 1));
  psi[1] = a[0] * M(u + 1, x) + a[1] * M(u + 2, x) + a[2] * 0.5 * (M(u + 3, x) + M(u + 1, x + 1));
  psi[2] = 0.5 * (a[0] * (M(u + 2, x) + M(u, x + 1)) + a[1] * (M(u + 3, x) + M(u + 1, x + 1)) +
                  a[2] * 0.5 * (M(u + 4, x) + M(u, x + 2) + 2 * M(u + 2, x + 1)));
}
// Microslope (1D): basic_function.cpp:551-559
GKS_DI void slope1(double* a, const double* der, const double* prim, const Gas& g) {
  const double R4 = der[2] / prim[0] - 0.5 * (prim[1] * prim[1] + 0.5 * (g.K + 1) / prim[2]) * der[0] / prim[0];
  const double R2 = (der[1] - prim[1] * der[0]) / prim[0];
  a[2] = 4 * prim[2] * prim[2] / (g.K + 1) * (2 * R4 - 2 * prim[1] * R2);
  a[1] = 2 * prim[2] * R2 - prim[1] * a[2];
  a[0] = der[0] / prim[0] - prim[1] * a[1] - 0.5 * a[2] * (prim[1] * prim[1] + 0.5 * (g.K + 1) / prim[2]);
}

// Center_collision (reconstruction.cpp:1375-1407) + GKS (flux_function.cpp:
// 16-276, gks2nd, two-stage).  L/R/C: convar[0..2], der1[3..5].  The dt/2
// evaluation keeps the reference's tau_num(dt) (:236).
GKS_DI void gks1_point(double* f, double* derf, double* center, const double* L, const double* Rr, double dt,
                       const Gas& g) {
  double pl[3], pr[3];
  ulambda1(pl, L, g);
  ulambda1(pr, Rr, g);
  MM1 ml, mr;
  mmdf1(ml, pl[1], pl[2], g);
  mmdf1(mr, pr[1], pr[2], g);
  const Tab1 tl{ml.up, ml}, tr{mr.un, mr}, wl{ml.uw, ml}, wr{mr.uw, mr};
  const double unit[3] = {1.0, 0.0, 0.0};
  double c0[6];
  {
    double gl[3], gr[3], axl[3], axr[3], a0l[3], a0r[3];
    contract1(0, 0, gl, unit, tl);
    contract1(0, 0, gr, unit, tr);
    slope1(axl, L + 3, pl, g);
    slope1(axr, Rr + 3, pr, g);
    contract1(0, 0, a0l, axl, tl);
    contract1(0, 0, a0r, axr, tr);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      c0[i] = L[0] * gl[i] + Rr[0] * gr[i];
      c0[3 + i] = L[0] * a0l[i] + Rr[0] * a0r[i];
    }
  }
  if (center)
#pragma unroll
    for (int i = 0; i < 6; i++) center[i] = c0[i];
  double p0[3];
  ulambda1(p0, c0, g);
  const double tau = tau_ns(p0[0], p0[2], g);
  double t[6];
  time_coe(t, tau, tau_num(pl[0], pr[0], p0[0], pl[2], pr[2], p0[2], dt, g), dt);
  double glu[3], gru[3], alxuul[3], arxuur[3], Alul[3], Arur[3];
  contract1(1, 0, glu, unit, tl);
  contract1(1, 0, gru, unit, tr);
  double alx[3], arx[3];
  slope1(alx, L + 3, pl, g);
  contract1(2, 0, alxuul, alx, tl);
  slope1(arx, Rr + 3, pr, g);
  contract1(2, 0, arxuur, arx, tr);
  {
    double xu[3], der[3], A[3];
    contract1(1, 0, xu, alx, wl);
#pragma unroll
    for (int i = 0; i < 3; i++) der[i] = -pl[0] * (xu[i]);
    slope1(A, der, pl, g);
    contract1(1, 0, Alul, A, tl);
    contract1(1, 0, xu, arx, wr);
#pragma unroll
    for (int i = 0; i < 3; i++) der[i] = -pr[0] * (xu[i]);
    slope1(A, der, pr, g);
    contract1(1, 0, Arur, A, tr);
  }
  double g0u[3], a0xuu[3], A0u[3];
  {
    MM1 m0;
    mmdf1(m0, p0[1], p0[2], g);
    const Tab1 w0{m0.uw, m0};
    contract1(1, 0, g0u, unit, w0);
    double a0x[3], a0xu[3], derA0[3], A0[3];
    slope1(a0x, c0 + 3, p0, g);
    contract1(1, 0, a0xu, a0x, w0);
    contract1(2, 0, a0xuu, a0x, w0);
#pragma unroll
    for (int i = 0; i < 3; i++) derA0[i] = -p0[0] * (a0xu[i]);
    slope1(A0, derA0, p0, g);
    contract1(1, 0, A0u, A0, w0);
  }
  double F0[3], F1[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    F0[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
    F0[i] = F0[i] + p0[0] * t[3] * g0u[i];
    F0[i] = F0[i] + pl[0] * t[1] * (alxuul[i]) + pr[0] * t[1] * (arxuur[i]);
    F0[i] = F0[i] + pl[0] * t[2] * (Alul[i]) + pr[0] * t[2] * (Arur[i]);
    F0[i] = F0[i] + p0[0] * t[4] * (a0xuu[i]);
    F0[i] = F0[i] + p0[0] * t[5] * (A0u[i]);
  }
  const double dt2 = 0.5 * dt;
  const double tn = tau_num(pl[0], pr[0], p0[0], pl[2], pr[2], p0[2], dt, g);
  const double eta = exp(-dt2 / tn);
  t[0] = tn * (1 - eta);
  t[1] = tn * (eta * (dt2 + tn) - tn) + tau * tn * (eta - 1);
  t[2] = tau * tn * (eta - 1);
  t[3] = tn * eta + dt2 - tn;
  t[4] = tn * (tn - eta * (dt2 + tn) - tau * (eta - 1)) - dt2 * tau;
  t[5] = 0.5 * dt2 * dt2 - tau * tn * (eta - 1) - tau * dt2;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    F1[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
    F1[i] = F1[i] + pl[0] * t[1] * (alxuul[i]) + pr[0] * t[1] * (arxuur[i]);
    F1[i] = F1[i] + pl[0] * t[2] * (Alul[i]) + pr[0] * t[2] * (Arur[i]);
    F1[i] = F1[i] + p0[0] * t[3] * g0u[i];
    F1[i] = F1[i] + p0[0] * t[4] * (a0xuu[i]);
    F1[i] = F1[i] + p0[0] * t[5] * (A0u[i]);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    f[i] = (4.0 * F1[i] - F0[i]);
    derf[i] = 4.0 * (F0[i] - 2.0 * F1[i]);
  }
}

// weno_5th_ao_left / _right without DF (reconstruction.cpp:161-360): full
// beta3, eps 1e-10, WENO-Z; values and first derivatives
GKS_DI void weno5n_weights(double* fw, double wn2, double wn1, double w0, double wp1, double wp2) {
  const double dhi = 0.85, dlo = 0.85, eps = 1e-10;
  const double d[4] = {(1.0 - dhi) * (1.0 - dlo) / 2.0, (1.0 - dhi) * dlo, (1.0 - dhi) * (1.0 - dlo) / 2.0, dhi};
  double beta[4], q, q2;
  q = wn2 - 2.0 * wn1 + w0;
  q2 = wn2 - 4.0 * wn1 + 3.0 * w0;
  beta[0] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = wn1 - 2.0 * w0 + wp1;
  q2 = wn1 - wp1;
  beta[1] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = w0 - 2.0 * wp1 + wp2;
  q2 = 3.0 * w0 - 4.0 * wp1 + wp2;
  beta[2] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  beta[3] = (1.0 / 5040.0) * (231153.0 * w0 * w0 + 104963.0 * wn1 * wn1 + 6908.0 * wn2 * wn2 -
                              38947.0 * wn2 * wp1 + 104963.0 * wp1 * wp1 +
                              wn1 * (-51001.0 * wn2 + 179098.0 * wp1 - 38947.0 * wp2) -
                              3.0 * w0 * (99692.0 * wn1 - 22641.0 * wn2 + 99692.0 * wp1 - 22641.0 * wp2) +
                              8209.0 * wn2 * wp2 - 51001.0 * wp1 * wp2 + 6908.0 * wp2 * wp2);
  const double tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  double alpha[4], sum = 0.0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double gd = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum += alpha[i];
  }
  double ww[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum;
  fw[3] = ww[3] / d[3];
#pragma unroll
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] / d[3] * d[k];
}
GKS_DI void weno5n(bool left, double& var, double& der1, double wn2, double wn1, double w0, double wp1, double wp2,
                   double h) {
  double fw[4], p[4], px[4];
  weno5n_weights(fw, wn2, wn1, w0, wp1, wp2);
  if (left) {
    p[0] = -1.0 / 6.0 * wn2 + 5.0 / 6.0 * wn1 + 1.0 / 3.0 * w0;
    p[1] = 1.0 / 3.0 * wn1 + 5.0 / 6.0 * w0 - 1.0 / 6.0 * wp1;
    p[2] = 11.0 / 6.0 * w0 - 7.0 / 6.0 * wp1 + 1.0 / 3.0 * wp2;
    p[3] = (1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2);
    px[0] = (w0 - wn1) / h;
    px[1] = (w0 - wn1) / h;
    px[2] = -((2.0 * w0 - 3.0 * wp1 + wp2) / h);
    px[3] = (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) / (12.0 * h);
  } else {
    p[0] = 1.0 / 3.0 * wn2 - 7.0 / 6.0 * wn1 + 11.0 / 6.0 * w0;
    p[1] = -1.0 / 6.0 * wn1 + 5.0 / 6.0 * w0 + 1.0 / 3.0 * wp1;
    p[2] = 1.0 / 3.0 * w0 + 5.0 / 6.0 * wp1 - 1.0 / 6.0 * wp2;
    p[3] = (1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2);
    px[0] = (2.0 * w0 - 3.0 * wn1 + wn2) / h;
    px[1] = (-w0 + wp1) / h;
    px[2] = (-w0 + wp1) / h;
    px[3] = (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) / (12.0 * h);
  }
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}

// WENO5_AO (reconstruction.cpp:85-159) / WENO5_AO_with_DF (:394-478) for one
// cell; lo = state at the cell's left face, ro = at its right face
// (convar[0..2], der1[3..5]); Check_Order_Reduce (:13-44)
GKS_DI void weno5_cell1(double* lo, double* ro, const double (*w)[3], const double* alpha, bool with_df, double h,
                        double ih, int recon_var, const Gas& g) {
  double df[4] = {1.0, 1.0, 1.0, 1.0};
  if (with_df) df5(df, alpha);
  auto recon = [&](bool left, double& var, double& d1, double a, double b, double c, double d, double e) {
    if (with_df) {
      if (left) weno5_left(var, d1, a, b, c, d, e, df, h, ih);
      else weno5_right(var, d1, a, b, c, d, e, df, h, ih);
    } else {
      weno5n(left, var, d1, a, b, c, d, e, h);
    }
  };
  if (recon_var == 0) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      recon(false, ro[i], ro[3 + i], w[0][i], w[1][i], w[2][i], w[3][i], w[4][i]);
      recon(true, lo[i], lo[3 + i], w[0][i], w[1][i], w[2][i], w[3][i], w[4][i]);
    }
  } else {
    double pn1[3], p0[3], pp1[3];
    prim1(pn1, w[1], g);
    prim1(p0, w[2], g);
    prim1(pp1, w[3], g);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      double base[3];
#pragma unroll
      for (int i = 0; i < 3; i++) base[i] = side == 0 ? 0.5 * (pn1[i] + p0[i]) : 0.5 * (pp1[i] + p0[i]);
      double c[5][3], var[3], d1[3];
#pragma unroll
      for (int s = 0; s < 5; s++) to_char1(c[s], base, w[s], g);
#pragma unroll
      for (int i = 0; i < 3; i++) recon(side == 0, var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i]);
      double* out = side == 0 ? lo : ro;
      from_char1(out, base, var, g);
      from_char1(out + 3, base, d1, g);
    }
  }
  auto bad = [&](const double* c) {
    const double lam = lambda1(c[0], c[1] / c[0], c[2], g);
    return lam <= 0.0 || !(lam == lam);
  };
  const bool r0 = bad(ro), r1 = bad(lo);
  if (r0 || r1) {
#pragma unroll
    for (int m = 0; m < 3; m++) {
      lo[m] = w[2][m];
      ro[m] = w[2][m];
      lo[3 + m] = 0.0;
      ro[3 + m] = 0.0;
    }
  }
}

// Calculate_alpha_k_1d: reconstruction.cpp:362-372
GKS_DI double alpha_k1(const double* pl, const double* pr, const Gas& g) {
  const double ml = pl[1] / sqrt(g.gamma * pl[2] / pl[0]);
  const double mr = pr[1] / sqrt(g.gamma * pr[2] / pr[0]);
  double A = fabs(pl[2] - pr[2]) / pl[2] + fabs(pl[2] - pr[2]) / pr[2];
  const double d = ml - mr;
  A = A + d * d;
  return A;
}

}  // namespace GKS_VARIANT
}  // namespace gksd
