// solver1d.hpp — CPU restatement of the reference's 1D GKS path (test
// infrastructure only; see gks_core.hpp for the conventions).  Covers the
// SodTubeProblem configuration: WENO5-AO (first step) / WENO5-AO-DF (later
// steps) in characteristic or conservative variables, Center_collision, GKS
// with gks2nd, S2O4, Update_alpha, Get_CFL, free / periodic / reflective
// boundaries.
#pragma once
#include <vector>

#include "gks_core.hpp"

namespace orc {

// Lambda (1D): basic_function.cpp:1169-1172
template <class R> inline R lambda_1d(R rho, R u, R rhoE, const Gas& g) {
  return (g.K + 1.0) * 0.25 * (rho / (rhoE - 0.5 * rho * (u * u)));
}
// Convar_to_primvar_1D / Convar_to_ULambda_1d: basic_function.cpp:804-809, 828-833
template <class R> inline void convar_to_primvar_1d(R* p, const R* c, const Gas& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = (g.gamma - 1.0) * (c[2] - 0.5 * c[1] * c[1] / c[0]);
}
template <class R> inline void convar_to_ulambda_1d(R* p, const R* c, const Gas& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = lambda_1d(c[0], p[1], c[2], g);
}
// Primvar_to_convar_1D with DensityE: basic_function.cpp:853-858, 1102-1105
template <class R> inline void primvar_to_convar_1d(R* c, const R* p, const Gas& g) {
  c[0] = p[0];
  c[1] = p[0] * p[1];
  c[2] = p[0] * (p[2] / p[0] / (g.gamma - 1) + 0.5 * (p[1] * p[1]));
}
// Convar_to_char1D / Char_to_convar1D: basic_function.cpp:868-921
template <class R> inline void convar_to_char_1d(R* ch, const R* prim, const R* c, const Gas& g) {
  R cs = sqrt(g.gamma * prim[2] / prim[0]);
  R alfa = (g.gamma - 1.0) / (2.0 * cs * cs);
  R u = prim[1];
  R s[3][3];
  s[0][0] = alfa * (0.5 * u * u + u * cs / (g.gamma - 1.0));
  s[0][1] = alfa * (-u - cs / (g.gamma - 1.0));
  s[0][2] = alfa;
  s[1][0] = alfa * (-u * u + 2.0 * cs * cs / (g.gamma - 1.0));
  s[1][1] = alfa * 2.0 * u;
  s[1][2] = -2.0 * alfa;
  s[2][0] = alfa * (0.5 * u * u - u * cs / (g.gamma - 1.0));
  s[2][1] = alfa * (-u + cs / (g.gamma - 1.0));
  s[2][2] = alfa;
  for (int i = 0; i < 3; i++) {
    R acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + s[i][j] * c[j];
    ch[i] = acc;
  }
}
template <class R> inline void char_to_convar_1d(R* c, const R* prim, const R* ch, const Gas& g) {
  R cs = sqrt(g.gamma * prim[2] / prim[0]);
  R u = prim[1];
  R h = 0.5 * u * u + cs * cs / (g.gamma - 1.0);
  R s[3][3];
  s[0][0] = 1.0;
  s[0][1] = 1.0;
  s[0][2] = 1.0;
  s[1][0] = u - cs;
  s[1][1] = u;
  s[1][2] = u + cs;
  s[2][0] = h - u * cs;
  s[2][1] = u * u / 2.0;
  s[2][2] = h + u * cs;
  for (int i = 0; i < 3; i++) {
    R acc = 0.0;
    for (int j = 0; j < 3; j++) acc = acc + s[i][j] * ch[j];
    c[i] = acc;
  }
}

// MMDF1d: basic_function.cpp:16-99 (moments up to order 9 in u, xi up to 6)
template <class R> struct MM1 {
  R uwhole[10], uplus[10], uminus[10];
  R xi2, xi4, xi6;
  // upxi / unxi / uxi [i][k] = u-moment i times xi^(2k)
  R up(int i, int k) const { return k == 0 ? uplus[i] : uplus[i] * xi(k); }
  R un(int i, int k) const { return k == 0 ? uminus[i] : uminus[i] * xi(k); }
  R uw(int i, int k) const { return k == 0 ? uwhole[i] : uwhole[i] * xi(k); }
  R xi(int k) const { return k == 1 ? xi2 : (k == 2 ? xi4 : xi6); }
};
template <class R> inline void mmdf1(MM1<R>& m, R u, R lambda, const Gas& g) {
  m.uwhole[0] = 1;
  m.uwhole[1] = u;
  m.uplus[0] = 0.5 * alpha_fn(lambda, -u);
  m.uminus[0] = 0.5 * alpha_fn(lambda, u);
  m.uplus[1] = u * m.uplus[0] + 0.5 * beta_fn(lambda, u);
  m.uminus[1] = u * m.uminus[0] - 0.5 * beta_fn(lambda, u);
  for (int i = 2; i <= 9; i++) m.uwhole[i] = u * m.uwhole[i - 1] + 0.5 * (i - 1) / lambda * m.uwhole[i - 2];
  for (int i = 2; i <= 9; i++) {
    m.uplus[i] = u * m.uplus[i - 1] + 0.5 * (i - 1) / lambda * m.uplus[i - 2];
    m.uminus[i] = u * m.uminus[i - 1] + 0.5 * (i - 1) / lambda * m.uminus[i - 2];
  }
  m.xi2 = 0.5 * g.K / lambda;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) / (lambda * lambda);
  m.xi6 = 0.5 * (g.K + 4) / lambda * m.xi4;
}
// G / GL / GR (1D): basic_function.cpp:522-549; which = 0 whole, 1 u>0, 2 u<0
template <class R>
inline void g1_contract(int nu, int nx, R* psi, const R* a, const MM1<R>& m, int which) {
  auto M = [&](int i, int k) { return which == 1 ? m.up(i, k) : (which == 2 ? m.un(i, k) : m.uw(i, k)); };
  psi[0] = a[0] * M(nu, nx) + a[1] * M(nu + 1, nx) + a[2] * 0.5 * (M(nu + 2, nx) + M(nu, nx + 1));
  psi[1] = a[0] * M(nu + 1, nx) + a[1] * M(nu + 2, nx) + a[2] * 0.5 * (M(nu + 3, nx) + M(nu + 1, nx + 1));
  psi[2] = 0.5 * (a[0] * (M(nu + 2, nx) + M(nu, nx + 1)) + a[1] * (M(nu + 3, nx) + M(nu + 1, nx + 1)) +
                  a[2] * 0.5 * (M(nu + 4, nx) + M(nu, nx + 2) + 2 * M(nu + 2, nx + 1)));
}
// Microslope (1D): basic_function.cpp:551-559
template <class R> inline void micro_slope_1d(R* a, const R* der, const R* prim, const Gas& g) {
  R R4 = der[2] / prim[0] - 0.5 * (prim[1] * prim[1] + 0.5 * (g.K + 1) / prim[2]) * der[0] / prim[0];
  R R2 = (der[1] - prim[1] * der[0]) / prim[0];
  a[2] = 4 * prim[2] * prim[2] / (g.K + 1) * (2 * R4 - 2 * prim[1] * R2);
  a[1] = 2 * prim[2] * R2 - prim[1] * a[2];
  a[0] = der[0] / prim[0] - prim[1] * a[1] - 0.5 * a[2] * (prim[1] * prim[1] + 0.5 * (g.K + 1) / prim[2]);
}

template <class R> struct Pt1 {
  R convar[3];
  R der1[3];
};

// Center_collision (1D): reconstruction.cpp:1375-1407
template <class R> inline void center_collision_1d(Pt1<R>& c, const Pt1<R>& l, const Pt1<R>& r, const Gas& g) {
  R pl[3], pr[3];
  convar_to_ulambda_1d(pl, l.convar, g);
  convar_to_ulambda_1d(pr, r.convar, g);
  MM1<R> ml, mr;
  mmdf1(ml, pl[1], pl[2], g);
  mmdf1(mr, pr[1], pr[2], g);
  const R unit[3] = {1.0, 0.0, 0.0};
  R gl[3], gr[3], axl[3], axr[3], ax0l[3], ax0r[3];
  g1_contract(0, 0, gl, unit, ml, 1);
  g1_contract(0, 0, gr, unit, mr, 2);
  micro_slope_1d(axl, l.der1, pl, g);
  micro_slope_1d(axr, r.der1, pr, g);
  g1_contract(0, 0, ax0l, axl, ml, 1);
  g1_contract(0, 0, ax0r, axr, mr, 2);
  for (int i = 0; i < 3; i++) {
    c.convar[i] = l.convar[0] * gl[i] + r.convar[0] * gr[i];
    c.der1[i] = l.convar[0] * ax0l[i] + r.convar[0] * ax0r[i];
  }
}

// GKS (1D, gks2nd, two-stage scheme): flux_function.cpp:16-276.  The dt/2
// evaluation keeps the reference's tau_num(dt) (:236).
template <class R>
inline void gks1d_flux(R* f, R* derf, const Pt1<R>& l, const Pt1<R>& r, const Pt1<R>& c, double dt, const Gas& g) {
  R F0[3], F1[3];
  R pl[3], pr[3], p0[3];
  convar_to_ulambda_1d(pl, l.convar, g);
  convar_to_ulambda_1d(pr, r.convar, g);
  convar_to_ulambda_1d(p0, c.convar, g);
  R tau = tau_ns(p0[0], p0[2], g);
  R tn = tau_num(pl[0], pr[0], p0[0], pl[2], pr[2], p0[2], dt, g);
  R t[6];
  time_coe(t, tau, tn, dt);
  MM1<R> ml, mr, m0;
  mmdf1(ml, pl[1], pl[2], g);
  mmdf1(mr, pr[1], pr[2], g);
  const R unit[3] = {1.0, 0.0, 0.0};
  R glu[3], gru[3];
  g1_contract(1, 0, glu, unit, ml, 1);
  g1_contract(1, 0, gru, unit, mr, 2);
  for (int i = 0; i < 3; i++) F0[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
  mmdf1(m0, p0[1], p0[2], g);
  R g0u[3];
  g1_contract(1, 0, g0u, unit, m0, 0);
  for (int i = 0; i < 3; i++) F0[i] = F0[i] + p0[0] * t[3] * g0u[i];
  R alx[3], alxuul[3], arx[3], arxuur[3];
  micro_slope_1d(alx, l.der1, pl, g);
  g1_contract(2, 0, alxuul, alx, ml, 1);
  micro_slope_1d(arx, r.der1, pr, g);
  g1_contract(2, 0, arxuur, arx, mr, 2);
  for (int i = 0; i < 3; i++) F0[i] = F0[i] + pl[0] * t[1] * (alxuul[i]) + pr[0] * t[1] * (arxuur[i]);
  R alxu[3], arxu[3], derl[3], derr[3], Al[3], Ar[3], Alul[3], Arur[3];
  g1_contract(1, 0, alxu, alx, ml, 0);
  g1_contract(1, 0, arxu, arx, mr, 0);
  for (int i = 0; i < 3; i++) {
    derl[i] = -pl[0] * (alxu[i]);
    derr[i] = -pr[0] * (arxu[i]);
  }
  micro_slope_1d(Al, derl, pl, g);
  micro_slope_1d(Ar, derr, pr, g);
  g1_contract(1, 0, Alul, Al, ml, 1);
  g1_contract(1, 0, Arur, Ar, mr, 2);
  for (int i = 0; i < 3; i++) F0[i] = F0[i] + pl[0] * t[2] * (Alul[i]) + pr[0] * t[2] * (Arur[i]);
  R a0x[3], a0xu[3], a0xuu[3];
  micro_slope_1d(a0x, c.der1, p0, g);
  g1_contract(1, 0, a0xu, a0x, m0, 0);
  g1_contract(2, 0, a0xuu, a0x, m0, 0);
  for (int i = 0; i < 3; i++) F0[i] = F0[i] + p0[0] * t[4] * (a0xuu[i]);
  R derA0[3], A0[3], A0u[3];
  for (int i = 0; i < 3; i++) derA0[i] = -p0[0] * (a0xu[i]);
  micro_slope_1d(A0, derA0, p0, g);
  g1_contract(1, 0, A0u, A0, m0, 0);
  for (int i = 0; i < 3; i++) F0[i] = F0[i] + p0[0] * t[5] * (A0u[i]);
  // dt/2, with tau_num evaluated with dt (flux_function.cpp:236)
  double dt2 = 0.5 * dt;
  tn = tau_num(pl[0], pr[0], p0[0], pl[2], pr[2], p0[2], dt, g);
  R eta = exp(-dt2 / tn);
  t[0] = tn * (1 - eta);
  t[1] = tn * (eta * (dt2 + tn) - tn) + tau * tn * (eta - 1);
  t[2] = tau * tn * (eta - 1);
  t[3] = tn * eta + dt2 - tn;
  t[4] = tn * (tn - eta * (dt2 + tn) - tau * (eta - 1)) - dt2 * tau;
  t[5] = 0.5 * dt2 * dt2 - tau * tn * (eta - 1) - tau * dt2;
  for (int i = 0; i < 3; i++) {
    F1[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
    F1[i] = F1[i] + pl[0] * t[1] * (alxuul[i]) + pr[0] * t[1] * (arxuur[i]);
    F1[i] = F1[i] + pl[0] * t[2] * (Alul[i]) + pr[0] * t[2] * (Arur[i]);
    F1[i] = F1[i] + p0[0] * t[3] * g0u[i];
    F1[i] = F1[i] + p0[0] * t[4] * (a0xuu[i]);
    F1[i] = F1[i] + p0[0] * t[5] * (A0u[i]);
  }
  for (int i = 0; i < 3; i++) {
    f[i] = (4.0 * F1[i] - F0[i]);
    derf[i] = 4.0 * (F0[i] - 2.0 * F1[i]);
  }
}

// weno_5th_ao_left / _right (no DF; full beta3; eps 1e-10; WENO-Z):
// reconstruction.cpp:161-360.  The second derivative they also form is not
// used by any caller and is not computed.
template <class R> inline void weno5_nodf_weights(R* fw, R wn2, R wn1, R w0, R wp1, R wp2) {
  const double dhi = 0.85, dlo = 0.85, eps = 1e-10;
  double d[4];
  d[0] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[1] = (1.0 - dhi) * dlo;
  d[2] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[3] = dhi;
  R beta[4], q, q2;
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
  R tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  R alpha[4], sum = 0.0;
  for (int i = 0; i < 4; i++) {
    R gd = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum += alpha[i];
  }
  R ww[4];
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum;
  fw[3] = ww[3] / d[3];
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] / d[3] * d[k];
}
template <class R> inline void weno5_nodf_left(R& var, R& der1, R wn2, R wn1, R w0, R wp1, R wp2, double h) {
  R fw[4], p[4], px[4];
  weno5_nodf_weights(fw, wn2, wn1, w0, wp1, wp2);
  p[0] = -1.0 / 6.0 * wn2 + 5.0 / 6.0 * wn1 + 1.0 / 3.0 * w0;
  p[1] = 1.0 / 3.0 * wn1 + 5.0 / 6.0 * w0 - 1.0 / 6.0 * wp1;
  p[2] = 11.0 / 6.0 * w0 - 7.0 / 6.0 * wp1 + 1.0 / 3.0 * wp2;
  p[3] = (1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2);
  px[0] = (w0 - wn1) / h;
  px[1] = (w0 - wn1) / h;
  px[2] = -((2.0 * w0 - 3.0 * wp1 + wp2) / h);
  px[3] = (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) / (12.0 * h);
  var = 0.0;
  der1 = 0.0;
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
template <class R> inline void weno5_nodf_right(R& var, R& der1, R wn2, R wn1, R w0, R wp1, R wp2, double h) {
  R fw[4], p[4], px[4];
  weno5_nodf_weights(fw, wn2, wn1, w0, wp1, wp2);
  p[0] = 1.0 / 3.0 * wn2 - 7.0 / 6.0 * wn1 + 11.0 / 6.0 * w0;
  p[1] = -1.0 / 6.0 * wn1 + 5.0 / 6.0 * w0 + 1.0 / 3.0 * wp1;
  p[2] = 1.0 / 3.0 * w0 + 5.0 / 6.0 * wp1 - 1.0 / 6.0 * wp2;
  p[3] = (1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2);
  px[0] = (2.0 * w0 - 3.0 * wn1 + wn2) / h;
  px[1] = (-w0 + wp1) / h;
  px[2] = (-w0 + wp1) / h;
  px[3] = (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) / (12.0 * h);
  var = 0.0;
  der1 = 0.0;
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}

// WENO5_AO (reconstruction.cpp:85-159) and WENO5_AO_with_DF (:394-478) for one
// cell: left / right face states (the reference's `left` argument is the
// state at the cell's left face, stored as interfaces[i].right).
template <class R>
inline void weno5_1d(Pt1<R>& left, Pt1<R>& right, const R (*w)[3], const R* alpha, bool with_df, double h,
                     int recon_var, const Gas& g) {
  R df[4] = {1.0, 1.0, 1.0, 1.0};
  if (with_df) df_factors5(df, alpha);
  auto recon = [&](bool is_left, R& var, R& d1, R a, R b, R c0, R d, R e) {
    if (with_df) {
      if (is_left) weno5_df_left(var, d1, a, b, c0, d, e, df, h);
      else weno5_df_right(var, d1, a, b, c0, d, e, df, h);
    } else {
      if (is_left) weno5_nodf_left(var, d1, a, b, c0, d, e, h);
      else weno5_nodf_right(var, d1, a, b, c0, d, e, h);
    }
  };
  if (recon_var == 0) {
    for (int i = 0; i < 3; i++) {
      recon(false, right.convar[i], right.der1[i], w[0][i], w[1][i], w[2][i], w[3][i], w[4][i]);
      recon(true, left.convar[i], left.der1[i], w[0][i], w[1][i], w[2][i], w[3][i], w[4][i]);
    }
  } else {
    R pn1[3], p0[3], pp1[3], bl[3], br[3];
    convar_to_primvar_1d(pn1, w[1], g);
    convar_to_primvar_1d(p0, w[2], g);
    convar_to_primvar_1d(pp1, w[3], g);
    for (int i = 0; i < 3; i++) {
      bl[i] = 0.5 * (pn1[i] + p0[i]);
      br[i] = 0.5 * (pp1[i] + p0[i]);
    }
    R c[5][3], var[3], d1[3];
    for (int s = 0; s < 5; s++) convar_to_char_1d(c[s], bl, w[s], g);
    for (int i = 0; i < 3; i++) recon(true, var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i]);
    char_to_convar_1d(left.convar, bl, var, g);
    char_to_convar_1d(left.der1, bl, d1, g);
    for (int s = 0; s < 5; s++) convar_to_char_1d(c[s], br, w[s], g);
    for (int i = 0; i < 3; i++) recon(false, var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i]);
    char_to_convar_1d(right.convar, br, var, g);
    char_to_convar_1d(right.der1, br, d1, g);
  }
  // Check_Order_Reduce: reconstruction.cpp:13-44
  auto bad = [&](const R* c) {
    R lam = lambda_1d(c[0], c[1] / c[0], c[2], g);
    return lam <= 0.0 || !(lam == lam);
  };
  bool r0 = bad(right.convar), r1 = bad(left.convar);
  if (r0 || r1)
    for (int m = 0; m < 3; m++) {
      left.convar[m] = w[2][m];
      right.convar[m] = w[2][m];
      left.der1[m] = 0.0;
      right.der1[m] = 0.0;
    }
}

// Calculate_alpha_k_1d: reconstruction.cpp:362-372
template <class R> inline R alpha_k_1d(const R* pl, const R* pr, const Gas& g) {
  R ml = pl[1] / sqrt(g.gamma * pl[2] / pl[0]);
  R mr = pr[1] / sqrt(g.gamma * pr[2] / pr[0]);
  R A = fabs(pl[2] - pr[2]) / pl[2] + fabs(pl[2] - pr[2]) / pr[2];
  R d = ml - mr;
  A = A + d * d;
  return A;
}

template <class R> struct Solver1d {
  int N, G, nx;
  double dx, left_x, cfl;
  int recon_var, use_df, df_from_start;
  Gas g;
  std::vector<R> U[3], Uold[3], alpha;
  std::vector<Pt1<R>> fL, fR, fC;  // per interface: left / right / centre state
  std::vector<R> fl[2];            // per stage: [face*6 + (0..2 f, 3..5 derf)]
  double t = 0.0;
  int step = 0;

  Solver1d(int nodex, int ghost, double left, double right, double cfl_, int recon, int df, int df0, const Gas& gas)
      : N(nodex), G(ghost), cfl(cfl_), recon_var(recon), use_df(df), df_from_start(df0), g(gas) {
    nx = N + 2 * G;
    dx = (right - left) / N;
    left_x = left;
    for (int v = 0; v < 3; v++) {
      U[v].assign(nx, R(0.0));
      Uold[v].assign(nx, R(0.0));
    }
    alpha.assign(nx, R(0.0));
    Pt1<R> z;
    for (int k = 0; k < 3; k++) z.convar[k] = z.der1[k] = R(0.0);
    fL.assign(nx + 1, z);
    fR.assign(nx + 1, z);
    fC.assign(nx + 1, z);
    for (int s = 0; s < 2; s++) fl[s].assign((size_t)(nx + 1) * 6, R(0.0));
  }
  // boundary_condition.cpp:4-64
  void boundary(int side, int kind) {
    for (int k = 0; k < G; k++) {
      int i = side == 0 ? G - 1 - k : N + G + k;
      int src;
      if (kind == 0) src = side == 0 ? i + 1 : i - 1;
      else if (kind == 1) src = side == 0 ? i + N : i - N;
      else src = side == 0 ? 2 * G - 1 - i : 2 * (nx - G) - 1 - i;
      for (int v = 0; v < 3; v++) {
        R val = U[v][src];
        if (kind == 2 && v == 1) val = -val;
        U[v][i] = val;
      }
    }
  }
  // Get_CFL / Dtx (1D): time_advance.cpp:75-111 (current convar)
  R cfl_dt(double tstop) const {
    R d = dx;
    for (int i = 0; i < N; i++) {
      R c[3] = {U[0][i + G], U[1][i + G], U[2][i + G]}, p[3];
      convar_to_primvar_1d(p, c, g);
      R tmp = fabs(p[1]) + sqrt(g.gamma * p[2] / p[0]);
      if (tmp > cfl * dx / d) d = cfl * dx / tmp;
      if (d > 0.25 * cfl * dx * dx / g.mu && g.tau_type == 1 && g.mu > 0) d = 0.25 * cfl * dx * dx / g.mu;
    }
    if (t + d > tstop) d = tstop - t + 1e-15;
    return d;
  }
  void stage_run(int stage, double dt, const int* bc, bool with_df) {
    boundary(0, bc[0]);
    boundary(1, bc[1]);
    // Reconstruction_within_cell: cells G-1 .. nx-G (reconstruction.cpp:46-53)
    for (int i = G - 1; i < nx - G + 1; i++) {
      R w[5][3], a[5];
      for (int s = 0; s < 5; s++) {
        for (int v = 0; v < 3; v++) w[s][v] = U[v][i - 2 + s];
        a[s] = alpha[i - 2 + s];
      }
      weno5_1d(fR[i], fL[i + 1], w, a, with_df, dx, recon_var, g);
    }
    // Reconstruction_forg0 + Calculate_flux: faces G .. nx-G
    for (int i = G; i < nx - G + 1; i++) center_collision_1d(fC[i], fL[i], fR[i], g);
    for (int i = G; i < N + G + 1; i++) gks1d_flux(&fl[stage][i * 6], &fl[stage][i * 6 + 3], fL[i], fR[i], fC[i], dt, g);
    // Update (S2O4 branch): time_advance.cpp:167-192
    static const double c[2][2][2] = {{{0.5, 1.0 / 8.0}, {0.0, 0.0}}, {{1.0, 1.0 / 6.0}, {0.0, 1.0 / 3.0}}};
    std::vector<R> F((size_t)(nx + 1) * 3);
    for (int i = G; i < N + G + 1; i++)
      for (int j = 0; j < 3; j++) {
        R acc = 0.0;
        for (int k = 0; k < stage + 1; k++) acc = acc + c[stage][k][0] * fl[k][i * 6 + j] + c[stage][k][1] * fl[k][i * 6 + 3 + j];
        F[i * 3 + j] = acc;
      }
    for (int i = G; i < N + G; i++)
      for (int j = 0; j < 3; j++) U[j][i] = Uold[j][i] + 1.0 / dx * (F[i * 3 + j] - F[(i + 1) * 3 + j]);
    // Update_alpha (1D): reconstruction.cpp:374-392
    if (use_df)
      for (int i = G; i < N + G; i++) {
        R pll[3], plr[3], prl[3], prr[3];
        convar_to_primvar_1d(pll, fL[i].convar, g);
        convar_to_primvar_1d(plr, fR[i].convar, g);
        convar_to_primvar_1d(prl, fL[i + 1].convar, g);
        convar_to_primvar_1d(prr, fR[i + 1].convar, g);
        R al = alpha_k_1d(pll, plr, g), ar = alpha_k_1d(prl, prr, g);
        al = al < 2.0 ? R(0.0) : al;
        ar = ar < 2.0 ? R(0.0) : ar;
        alpha[i] = al + ar;
      }
  }
  // SodTubeProblem's loop (1Dproblem.cpp:79-148): the DF reconstruction from
  // the start when df_from_start, from step 1 on when use_df
  int run(const int* bc, double tstop, int max_steps, double* dt_hist) {
    int n = 0;
    while (t < tstop && n < max_steps) {
      for (int i = G; i < G + N; i++)
        for (int v = 0; v < 3; v++) Uold[v][i] = U[v][i];
      R dt = cfl_dt(tstop);
      double d = as_double(dt);
      if (dt_hist) dt_hist[n] = d;
      bool with_df = df_from_start || (use_df && step > 0);
      for (int s = 0; s < 2; s++) stage_run(s, d, bc, with_df);
      step++;
      t = t + d;
      n++;
    }
    return 0;
  }
};

}  // namespace orc
