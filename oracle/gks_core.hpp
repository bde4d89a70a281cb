// gks_core.hpp — CPU restatement of the reference GKS numerics (2D leaves).
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it, never the product path.
//
// Every function restates one reference function (cited file:line, relative
// to the reference repository) with the same floating-point operation order,
// so that compiled with g++ -O2 -ffp-contract=off it reproduces the
// reference bit for bit (checked in tests/test_oracle.py against the compiled
// reference, oracle/_ref).  Templated on the scalar type so that a counting
// scalar (flopcount.cpp) can fix the algorithmic FLOP count per cell update.
#pragma once
#include <cmath>

namespace orc {

inline double as_double(double x) { return x; }

using std::erfc;
using std::exp;
using std::fabs;
using std::sqrt;

struct Gas {
  int K;
  double gamma, mu, nu, pr, c1, c2;
  int tau_type;     // 0 Euler, 1 NS
  int prandtl_fix;
};

// Point2d without geometry (fluid_mesh.h:125-137)
template <class R> struct Pt {
  R convar[4];
  R der1x[4];
  R der1y[4];
};

// ---------------------------------------------------------------- gas algebra
// basic_function.cpp:811-817
template <class R> inline void convar_to_primvar_2d(R* prim, const R* c, const Gas& g) {
  prim[0] = c[0];
  prim[1] = c[1] / c[0];
  prim[2] = c[2] / c[0];
  prim[3] = (g.gamma - 1.0) * (c[3] - 0.5 * c[1] * c[1] / c[0] - 0.5 * c[2] * c[2] / c[0]);
}
// basic_function.cpp:1174-1177 (Lambda, 2D overload)
template <class R> inline R lambda_2d(R rho, R u, R v, R rhoE, const Gas& g) {
  return (g.K + 2.0) * 0.25 * (rho / (rhoE - 0.5 * rho * (u * u + v * v)));
}
// basic_function.cpp:835-842
template <class R> inline void convar_to_ulambda_2d(R* prim, const R* c, const Gas& g) {
  prim[0] = c[0];
  prim[1] = c[1] / c[0];
  prim[2] = c[2] / c[0];
  prim[3] = lambda_2d(c[0], prim[1], prim[2], c[3], g);
}
// basic_function.cpp:860-866 with Q_densityE (:1127-1131)
template <class R> inline void primvar_to_convar_2d(R* c, const R* p, const Gas& g) {
  c[0] = p[0];
  c[1] = p[0] * p[1];
  c[2] = p[0] * p[2];
  c[3] = p[0] * (p[3] / p[0] / (g.gamma - 1) + 0.5 * (p[1] * p[1] + p[2] * p[2]));
}
// basic_function.cpp:1192-1198: rotate a y-face state into the x-like frame
template <class R> inline void y_to_x(R* o, const R* f) {
  o[0] = f[0];
  o[1] = f[2];
  o[2] = -f[1];
  o[3] = f[3];
}
// basic_function.cpp:1200-1206
template <class R> inline void x_to_y(R* o, const R* f) {
  o[0] = f[0];
  o[1] = -f[2];
  o[2] = f[1];
  o[3] = f[3];
}
// basic_function.cpp:1226-1233
template <class R> inline void local_to_global(R* c, const double* normal) {
  R t0 = c[1], t1 = c[2];
  c[1] = t0 * normal[0] - t1 * normal[1];
  c[2] = t0 * normal[1] + t1 * normal[0];
}

// basic_function.cpp:923-958 (Convar_to_char, 2D)
template <class R> inline void convar_to_char(R* ch, const R* prim, const R* c, const Gas& g) {
  R cs = sqrt(g.gamma * prim[3] / prim[0]);
  R overc = 1.0 / cs;
  R alfa = (g.gamma - 1.0) * 0.5 * overc * overc;
  R u = prim[1], v = prim[2];
  R s[4][4];
  s[0][0] = 1.0 - alfa * (u * u + v * v);
  s[0][1] = 2.0 * alfa * u;
  s[0][2] = 2.0 * alfa * v;
  s[0][3] = -2.0 * alfa;
  s[1][0] = -v;
  s[1][1] = 0.;
  s[1][2] = 1.;
  s[1][3] = 0.;
  s[2][0] = 0.5 * (-u * overc + alfa * (u * u + v * v));
  s[2][1] = 0.5 * overc - alfa * u;
  s[2][2] = -alfa * v;
  s[2][3] = alfa;
  s[3][0] = 0.5 * (u * overc + alfa * (u * u + v * v));
  s[3][1] = -0.5 * overc - alfa * u;
  s[3][2] = -alfa * v;
  s[3][3] = alfa;
  for (int i = 0; i < 4; i++) {
    R acc = 0.0;
    for (int j = 0; j < 4; j++) acc = acc + s[i][j] * c[j];
    ch[i] = acc;
  }
}
// basic_function.cpp:960-992 (Char_to_convar, 2D)
template <class R> inline void char_to_convar(R* c, const R* prim, const R* ch, const Gas& g) {
  R cs = sqrt(g.gamma * prim[3] / prim[0]);
  R u = prim[1], v = prim[2];
  R s[4][4];
  s[0][0] = 1.;
  s[0][1] = 0.0;
  s[0][2] = 1.0;
  s[0][3] = 1.0;
  s[1][0] = u;
  s[1][1] = 0.;
  s[1][2] = u + cs;
  s[1][3] = u - cs;
  s[2][0] = v;
  s[2][1] = 1.;
  s[2][2] = v;
  s[2][3] = v;
  s[3][0] = 0.5 * (u * u + v * v);
  s[3][1] = v;
  s[3][2] = 0.5 * (u * u + v * v) + cs * u + cs * cs / (g.gamma - 1.0);
  s[3][3] = 0.5 * (u * u + v * v) - cs * u + cs * cs / (g.gamma - 1.0);
  for (int i = 0; i < 4; i++) {
    R acc = 0.0;
    for (int j = 0; j < 4; j++) acc = acc + s[i][j] * ch[j];
    c[i] = acc;
  }
}

// --------------------------------------------------------- Maxwellian moments
// Alpha / Beta: basic_function.cpp:710-719
template <class R> inline R alpha_fn(R lambda, R u) { return erfc(sqrt(lambda) * u); }
template <class R> inline R beta_fn(R lambda, R u) {
  const double pi = 3.14159265358979323846;
  return exp(-lambda * u * u) / sqrt(pi * lambda);
}

// MMDF2d (basic_function.h:41-60, basic_function.cpp:103-187)
template <class R> struct MM2 {
  R uwhole[7], uplus[7], uminus[7], vwhole[7];
  R upvxi[7][7][3], unvxi[7][7][3], uvxi[7][7][3];
  R xi2, xi4;
};
template <class R> inline void mmdf2(MM2<R>& m, R u, R v, R lambda, const Gas& g) {
  m.uwhole[0] = 1.0;
  m.uwhole[1] = u;
  m.uplus[0] = 0.5 * alpha_fn(lambda, -u);
  m.uminus[0] = 1.0 - m.uplus[0];
  m.uplus[1] = u * m.uplus[0] + 0.5 * beta_fn(lambda, u);
  m.uminus[1] = u - m.uplus[1];
  R ol = 1.0 / lambda;
  for (int i = 2; i <= 6; i++) m.uwhole[i] = u * m.uwhole[i - 1] + 0.5 * (i - 1) * ol * m.uwhole[i - 2];
  for (int i = 2; i <= 6; i++) {
    m.uplus[i] = u * m.uplus[i - 1] + 0.5 * (i - 1) * ol * m.uplus[i - 2];
    m.uminus[i] = m.uwhole[i] - m.uplus[i];
  }
  m.vwhole[0] = 1.0;
  m.vwhole[1] = v;
  for (int i = 2; i <= 6; i++) m.vwhole[i] = v * m.vwhole[i - 1] + 0.5 * (i - 1) * ol * m.vwhole[i - 2];
  m.xi2 = 0.5 * g.K * ol;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) * ol * ol;
  for (int i = 0; i < 7; i++)
    for (int j = 0; j < 7; j++)
      for (int k = 0; k < 3; k++) {
        if (i + j + 2 * k > 6) continue;
        if (k == 0) {
          m.upvxi[i][j][k] = m.uplus[i] * m.vwhole[j];
          m.unvxi[i][j][k] = m.uminus[i] * m.vwhole[j];
        } else {
          R xi = (k == 1) ? m.xi2 : m.xi4;
          m.upvxi[i][j][k] = m.uplus[i] * m.vwhole[j] * xi;
          m.unvxi[i][j][k] = m.uminus[i] * m.vwhole[j] * xi;
        }
        m.uvxi[i][j][k] = m.upvxi[i][j][k] + m.unvxi[i][j][k];
      }
}

// GL_address / GR_address / G_address (basic_function.cpp:603-636): moment
// contraction of a microscopic slope a[4] against one of the three tables.
template <class R, class T>
inline void g_contract(int u, int v, int x, R* psi, const R* a, const T& m) {
  psi[0] = a[0] * m[u][v][x] + a[1] * m[u + 1][v][x] + a[2] * m[u][v + 1][x] +
           a[3] * 0.5 * (m[u + 2][v][x] + m[u][v + 2][x] + m[u][v][x + 1]);
  psi[1] = a[0] * m[u + 1][v][x] + a[1] * m[u + 2][v][x] + a[2] * m[u + 1][v + 1][x] +
           a[3] * 0.5 * (m[u + 3][v][x] + m[u + 1][v + 2][x] + m[u + 1][v][x + 1]);
  psi[2] = a[0] * m[u][v + 1][x] + a[1] * m[u + 1][v + 1][x] + a[2] * m[u][v + 2][x] +
           a[3] * 0.5 * (m[u + 2][v + 1][x] + m[u][v + 3][x] + m[u][v + 1][x + 1]);
  psi[3] = 0.5 * (a[0] * (m[u + 2][v][x] + m[u][v + 2][x] + m[u][v][x + 1]) +
                  a[1] * (m[u + 3][v][x] + m[u + 1][v + 2][x] + m[u + 1][v][x + 1]) +
                  a[2] * (m[u + 2][v + 1][x] + m[u][v + 3][x] + m[u][v + 1][x + 1]) +
                  a[3] * 0.5 *
                      (m[u + 4][v][x] + m[u][v + 4][x] + m[u][v][x + 2] + 2 * m[u + 2][v + 2][x] +
                       2 * m[u + 2][v][x + 1] + 2 * m[u][v + 2][x + 1]));
}
template <class R> inline void GL(int u, int v, int x, R* psi, const R* a, const MM2<R>& m) {
  g_contract(u, v, x, psi, a, m.upvxi);
}
template <class R> inline void GR(int u, int v, int x, R* psi, const R* a, const MM2<R>& m) {
  g_contract(u, v, x, psi, a, m.unvxi);
}
template <class R> inline void G(int u, int v, int x, R* psi, const R* a, const MM2<R>& m) {
  g_contract(u, v, x, psi, a, m.uvxi);
}

// A (2D overload): basic_function.cpp:571-583 — microscopic slope from the
// macroscopic derivative der[4] and (rho, U, V, lambda).
template <class R> inline void micro_slope(R* a, const R* der, const R* prim, const Gas& g) {
  R overden = 1.0 / prim[0];
  R R4 = der[3] * overden -
         0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + 0.5 * (g.K + 2) / prim[3]) * der[0] * overden;
  R R3 = (der[2] - prim[2] * der[0]) * overden;
  R R2 = (der[1] - prim[1] * der[0]) * overden;
  a[3] = (4.0 / (g.K + 2)) * prim[3] * prim[3] * (2 * R4 - 2 * prim[1] * R2 - 2 * prim[2] * R3);
  a[2] = 2 * prim[3] * R3 - prim[2] * a[3];
  a[1] = 2 * prim[3] * R2 - prim[1] * a[3];
  a[0] = der[0] * overden - prim[1] * a[1] - prim[2] * a[2] -
         0.5 * a[3] * (prim[1] * prim[1] + prim[2] * prim[2] + 0.5 * (g.K + 2) / prim[3]);
}

// Collision: basic_function.cpp:561-569
template <class R>
inline void collision(R* w0, R left, R right, const MM2<R>& m2, const MM2<R>& m3) {
  w0[0] = left * m2.uplus[0] + right * m3.uminus[0];
  w0[1] = left * m2.uplus[1] + right * m3.uminus[1];
  w0[2] = left * m2.vwhole[1] * m2.uplus[0] + right * m3.vwhole[1] * m3.uminus[0];
  w0[3] = 0.5 * left * (m2.uplus[2] + m2.uplus[0] * m2.vwhole[2] + m2.uplus[0] * m2.xi2) +
          0.5 * right * (m3.uminus[2] + m3.uminus[0] * m3.vwhole[2] + m3.uminus[0] * m3.xi2);
}

// Get_Tau_NS / Get_Tau: basic_function.cpp:638-708
template <class R> inline R tau_ns(R rho0, R lambda0, const Gas& g) {
  if (g.tau_type == 0) return R(0.0);
  if (g.mu > 0.0) return 2.0 * g.mu * lambda0 / rho0;
  if (g.nu > 0.0) return 2 * g.nu * lambda0;
  return R(0.0);
}
template <class R>
inline R tau_num(R rl, R rr, R r0, R ll, R lr, R l0, double dt, const Gas& g) {
  if (g.tau_type == 0) {
    if (g.c1 <= 0 && g.c2 <= 0) return R(0.0);
    R C = g.c2 * fabs(rl / ll - rr / lr) / fabs(rl / ll + rr / lr);
    return g.c1 * dt + dt * C;
  }
  R tau_n = g.c2 * fabs(rl / ll - rr / lr) / fabs(rl / ll + rr / lr) * dt;
  if (tau_n != tau_n) tau_n = 0.0;
  if ((g.mu > 0.0 && g.nu > 0.0) || (g.mu < 0.0 && g.nu < 0.0)) return R(0.0);
  if (g.mu > 0.0) return tau_n + 2.0 * g.mu * l0 / r0;
  if (g.nu > 0.0) return tau_n + 2 * g.nu * l0;
  return R(0.0);
}

// Center_all_collision_2d_multi: reconstruction.cpp:3652-3687 — collision
// centre state of one Gauss point (convar, der1x, der1y).
// ml / mr receive the left and right Maxwellian moments, which GKS2D needs
// again (the reference recomputes them from the same inputs).
template <class R>
inline void center_collision(Pt<R>& center, MM2<R>& ml, MM2<R>& mr, const Pt<R>& left, const Pt<R>& right,
                             const Gas& g) {
  R pl[4], pr[4];
  convar_to_ulambda_2d(pl, left.convar, g);
  convar_to_ulambda_2d(pr, right.convar, g);
  mmdf2(ml, pl[1], pl[2], pl[3], g);
  mmdf2(mr, pr[1], pr[2], pr[3], g);
  collision(center.convar, pl[0], pr[0], ml, mr);
  R alx[4], arx[4], al0x[4], ar0x[4];
  micro_slope(alx, left.der1x, pl, g);
  micro_slope(arx, right.der1x, pr, g);
  GL(0, 0, 0, al0x, alx, ml);
  GR(0, 0, 0, ar0x, arx, mr);
  R aly[4], ary[4], al0y[4], ar0y[4];
  micro_slope(aly, left.der1y, pl, g);
  micro_slope(ary, right.der1y, pr, g);
  GL(0, 0, 0, al0y, aly, ml);
  GR(0, 0, 0, ar0y, ary, mr);
  for (int k = 0; k < 4; k++) {
    center.der1x[k] = pl[0] * al0x[k] + pr[0] * ar0x[k];
    center.der1y[k] = pl[0] * al0y[k] + pr[0] * ar0y[k];
  }
}

// time-integration coefficients of gks2nd (flux_function.cpp:410-424)
template <class R> inline void time_coe(R* t, R tau, R tau_n, double dt) {
  R eta = exp(-dt / tau_n);
  t[0] = tau_n * (1 - eta);
  t[1] = tau_n * (eta * (dt + tau_n) - tau_n) + tau * tau_n * (eta - 1);
  t[2] = tau * tau_n * (eta - 1);
  t[3] = tau_n * eta + dt - tau_n;
  t[4] = tau_n * (tau_n - eta * (dt + tau_n) - tau * (eta - 1)) - dt * tau;
  t[5] = 0.5 * dt * dt - tau * tau_n * (eta - 1) - tau * dt;
}

// GKS2D with gks2nd_2d and a two-stage (not S1O1) time scheme:
// flux_function.cpp:345-698.  The centre state and the left/right moments
// are taken from center_collision (the reference recomputes both from the
// same inputs at :363-401).  Outputs f = 4F(dt/2) -
// F(dt) and derf = 4(F(dt) - 2F(dt/2)) in the face-local frame.
template <class R>
inline void gks2d_flux(R* f, R* derf, const Pt<R>& left, const Pt<R>& right, const Pt<R>& center,
                       const MM2<R>& ml, const MM2<R>& mr, double dt, const Gas& g) {
  R Flux[2][4];
  R pl[4], pr[4], p0[4];
  convar_to_ulambda_2d(pl, left.convar, g);
  convar_to_ulambda_2d(pr, right.convar, g);
  convar_to_ulambda_2d(p0, center.convar, g);
  R tau = tau_ns(p0[0], p0[3], g);
  R tn = tau_num(pl[0], pr[0], p0[0], pl[3], pr[3], p0[3], dt, g);
  R t[6];
  time_coe(t, tau, tn, dt);

  const R unit[4] = {1.0, 0.0, 0.0, 0.0};
  R glu[4], gru[4];
  GL(1, 0, 0, glu, unit, ml);
  GR(1, 0, 0, gru, unit, mr);
  for (int i = 0; i < 4; i++) Flux[0][i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];

  MM2<R> m0;
  mmdf2(m0, p0[1], p0[2], p0[3], g);
  R g0u[4];
  G(1, 0, 0, g0u, unit, m0);
  for (int i = 0; i < 4; i++) Flux[0][i] = Flux[0][i] + p0[0] * t[3] * g0u[i];

  R alx[4], alxuul[4], arx[4], arxuur[4], aly[4], alyuvl[4], ary[4], aryuvr[4];
  micro_slope(alx, left.der1x, pl, g);
  GL(2, 0, 0, alxuul, alx, ml);
  micro_slope(arx, right.der1x, pr, g);
  GR(2, 0, 0, arxuur, arx, mr);
  micro_slope(aly, left.der1y, pl, g);
  GL(1, 1, 0, alyuvl, aly, ml);
  micro_slope(ary, right.der1y, pr, g);
  GR(1, 1, 0, aryuvr, ary, mr);
  for (int i = 0; i < 4; i++)
    Flux[0][i] = Flux[0][i] + t[1] * (pl[0] * (alxuul[i] + alyuvl[i]) + pr[0] * (arxuur[i] + aryuvr[i]));

  R alxu[4], alyv[4], arxu[4], aryv[4];
  G(1, 0, 0, alxu, alx, ml);
  G(1, 0, 0, arxu, arx, mr);
  G(0, 1, 0, alyv, aly, ml);
  G(0, 1, 0, aryv, ary, mr);
  R Al[4], Ar[4], der_AL[4], der_AR[4];
  for (int i = 0; i < 4; i++) {
    der_AL[i] = -pl[0] * (alxu[i] + alyv[i]);
    der_AR[i] = -pr[0] * (arxu[i] + aryv[i]);
  }
  micro_slope(Al, der_AL, pl, g);
  micro_slope(Ar, der_AR, pr, g);
  R Alul[4], Arur[4];
  GL(1, 0, 0, Alul, Al, ml);
  GR(1, 0, 0, Arur, Ar, mr);
  for (int i = 0; i < 4; i++) Flux[0][i] = Flux[0][i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);

  R a0x[4], a0xu[4], a0xuu[4], a0y[4], a0yv[4], a0yuv[4];
  micro_slope(a0x, center.der1x, p0, g);
  G(1, 0, 0, a0xu, a0x, m0);
  G(2, 0, 0, a0xuu, a0x, m0);
  micro_slope(a0y, center.der1y, p0, g);
  G(0, 1, 0, a0yv, a0y, m0);
  G(1, 1, 0, a0yuv, a0y, m0);
  for (int i = 0; i < 4; i++) Flux[0][i] = Flux[0][i] + p0[0] * t[4] * (a0xuu[i] + a0yuv[i]);

  R derA0[4], A0[4], A0u[4];
  for (int i = 0; i < 4; i++) derA0[i] = -p0[0] * (a0xu[i] + a0yv[i]);
  micro_slope(A0, derA0, p0, g);
  G(1, 0, 0, A0u, A0, m0);
  for (int i = 0; i < 4; i++) Flux[0][i] = Flux[0][i] + p0[0] * t[5] * (A0u[i]);
  if (g.prandtl_fix) {
    R qs = a0xuu[3] + A0u[3] - p0[1] * (a0xuu[1] + A0u[1]) - p0[2] * (a0xuu[2] + A0u[2]);
    qs *= -p0[0] * tau * dt;
    Flux[0][3] += (1.0 / g.pr - 1.0) * qs;
  }

  // second evaluation over dt/2 (flux_function.cpp:648-690)
  double dt2 = 0.5 * dt;
  tn = tau_num(pl[0], pr[0], p0[0], pl[3], pr[3], p0[3], dt2, g);
  time_coe(t, tau, tn, dt2);
  for (int i = 0; i < 4; i++) {
    Flux[1][i] = t[0] * (pl[0] * glu[i] + pr[0] * gru[i]);
    Flux[1][i] = Flux[1][i] + t[1] * (pl[0] * (alxuul[i] + alyuvl[i]) + pr[0] * (arxuur[i] + aryuvr[i]));
    Flux[1][i] = Flux[1][i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);
    Flux[1][i] = Flux[1][i] + p0[0] * t[3] * g0u[i];
    Flux[1][i] = Flux[1][i] + p0[0] * t[4] * (a0xuu[i] + a0yuv[i]);
    Flux[1][i] = Flux[1][i] + p0[0] * t[5] * (A0u[i]);
  }
  if (g.prandtl_fix) {
    R qs = a0xuu[3] + A0u[3] - p0[1] * (a0xuu[1] + A0u[1]) - p0[2] * (a0xuu[2] + A0u[2]);
    qs *= -p0[0] * tau * 0.5 * dt;
    Flux[1][3] += (1.0 / g.pr - 1.0) * qs;
  }
  for (int i = 0; i < 4; i++) {
    f[i] = (4.0 * Flux[1][i] - Flux[0][i]);
    derf[i] = 4.0 * (Flux[0][i] - 2.0 * Flux[1][i]);
  }
}

// ------------------------------------------------------------- WENO5-AO (+DF)
// Shared smoothness indicators and AO weights of the DF variants
// (reconstruction.cpp:480-511, 569-600, 2742-2776): returns the final
// combination weights fw[4] = (w_k - w3 d_k / d3, w3 / d3).
template <class R>
inline void weno5_ao_df_weights(R* fw, R wn2, R wn1, R w0, R wp1, R wp2) {
  const double dhi = 0.85, dlo = 0.85;
  const double eps = 1e-6;
  double d[4];
  d[0] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[1] = (1.0 - dhi) * dlo;
  d[2] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[3] = dhi;
  R q;
  R beta[4];
  q = wn2 - 2.0 * wn1 + w0;
  R q2 = wn2 - 4.0 * wn1 + 3.0 * w0;
  beta[0] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = wn1 - 2.0 * w0 + wp1;
  q2 = wn1 - wp1;
  beta[1] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = w0 - 2.0 * wp1 + wp2;
  q2 = 3.0 * w0 - 4.0 * wp1 + wp2;
  beta[2] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  beta[3] = 1.0 / 6.0 * (beta[0] + 4.0 * beta[1] + beta[2]) + fabs(beta[0] - beta[2]);
  R tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  R alpha[4];
  R sum_alpha = 0.0;
  for (int i = 0; i < 4; i++) {
    R gd = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum_alpha += alpha[i];
  }
  R ww[4];
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum_alpha;
  fw[3] = ww[3] / d[3];
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] / d[3] * d[k];
}

// weno_5th_ao_with_df_left: reconstruction.cpp:480-567 — value and first
// derivative at the cell's left face.
template <class R>
inline void weno5_df_left(R& var, R& der1, R wn2, R wn1, R w0, R wp1, R wp2, const R* df, double h) {
  R fw[4];
  weno5_ao_df_weights(fw, wn2, wn1, w0, wp1, wp2);
  R p[4], px[4];
  p[0] = w0 + df[0] * (-1.0 / 6.0 * wn2 + 5.0 / 6.0 * wn1 + 1.0 / 3.0 * w0 - w0);
  p[1] = w0 + df[1] * (1.0 / 3.0 * wn1 + 5.0 / 6.0 * w0 - 1.0 / 6.0 * wp1 - w0);
  p[2] = w0 + df[2] * (11.0 / 6.0 * w0 - 7.0 / 6.0 * wp1 + 1.0 / 3.0 * wp2 - w0);
  p[3] = w0 + df[3] * ((1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2) - w0);
  px[0] = df[0] * (w0 - wn1) / h;
  px[1] = df[1] * (w0 - wn1) / h;
  px[2] = -df[2] * ((2.0 * w0 - 3.0 * wp1 + wp2) / h);
  px[3] = df[3] * (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) / (12.0 * h);
  var = 0.0;
  der1 = 0.0;
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
// weno_5th_ao_with_df_right: reconstruction.cpp:569-657
template <class R>
inline void weno5_df_right(R& var, R& der1, R wn2, R wn1, R w0, R wp1, R wp2, const R* df, double h) {
  R fw[4];
  weno5_ao_df_weights(fw, wn2, wn1, w0, wp1, wp2);
  R p[4], px[4];
  p[0] = w0 + df[0] * (1.0 / 3.0 * wn2 - 7.0 / 6.0 * wn1 + 11.0 / 6.0 * w0 - w0);
  p[1] = w0 + df[1] * (-1.0 / 6.0 * wn1 + 5.0 / 6.0 * w0 + 1.0 / 3.0 * wp1 - w0);
  p[2] = w0 + df[2] * (1.0 / 3.0 * w0 + 5.0 / 6.0 * wp1 - 1.0 / 6.0 * wp2 - w0);
  p[3] = w0 + df[3] * ((1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2) - w0);
  px[0] = df[0] * (2.0 * w0 - 3.0 * wn1 + wn2) / h;
  px[1] = df[1] * (-w0 + wp1) / h;
  px[2] = df[2] * (-w0 + wp1) / h;
  px[3] = df[3] * (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) / (12.0 * h);
  var = 0.0;
  der1 = 0.0;
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}

// weno_5th_ao_with_df_2gauss: reconstruction.cpp:2742-2838 — values and
// tangential derivatives at the two Gauss points -/+ h/(2 sqrt 3).
template <class R>
inline void weno5_df_2gauss(R& g1, R& g1x, R& g2, R& g2x, const R* df, R wn2, R wn1, R w0, R wp1,
                            R wp2, double h) {
  R fw[4];
  weno5_ao_df_weights(fw, wn2, wn1, w0, wp1, wp2);
  const double sqrt3 = 1.732050807568877294;
  R p[4], px[4];
  p[0] = w0 + df[0] * (w0 - (sqrt3 * w0) / 4 + (4 * wn1 - wn2) / (4 * sqrt3) - w0);
  p[1] = w0 + df[1] * (w0 + (wn1 - wp1) / (4 * sqrt3) - w0);
  p[2] = w0 + df[2] * ((1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wp1 + wp2)) - w0);
  p[3] = w0 + df[3] * ((4314 * w0 + (4 + 500 * sqrt3) * wn1 - wn2 - 70 * sqrt3 * wn2 + 4 * wp1 -
                        500 * sqrt3 * wp1 - wp2 + 70 * sqrt3 * wp2) / 4320 - w0);
  px[0] = -df[0] * ((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wn1 + (-3 + sqrt3) * wn2) / (6 * h);
  px[1] = -df[1] * (-2 * sqrt3 * w0 + (3 + sqrt3) * wn1 + (-3 + sqrt3) * wp1) / (6 * h);
  px[2] = -df[2] * ((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wp1 + (3 + sqrt3) * wp2) / (6 * h);
  px[3] = df[3] * (48 * sqrt3 * w0 - 72 * wn1 - 26 * sqrt3 * wn1 + 9 * wn2 + 2 * sqrt3 * wn2 +
                   72 * wp1 - 26 * sqrt3 * wp1 - 9 * wp2 + 2 * sqrt3 * wp2) / (108 * h);
  g1 = 0.0;
  g1x = 0.0;
  for (int k = 0; k < 4; k++) {
    g1 += fw[k] * p[k];
    g1x += fw[k] * px[k];
  }
  p[0] = w0 + df[0] * ((1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wn1 + wn2)) - w0);
  p[1] = w0 + df[1] * (w0 + (-wn1 + wp1) / (4 * sqrt3) - w0);
  p[2] = w0 + df[2] * (w0 - (sqrt3 * w0) / 4 + (4 * wp1 - wp2) / (4 * sqrt3) - w0);
  p[3] = w0 + df[3] * ((4314 * w0 + (4 - 500 * sqrt3) * wn1 - wn2 + 70 * sqrt3 * wn2 + 4 * wp1 +
                        500 * sqrt3 * wp1 - wp2 - 70 * sqrt3 * wp2) / 4320 - w0);
  px[0] = df[0] * ((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wn1 + (3 + sqrt3) * wn2) / (6 * h);
  px[1] = df[1] * (-2 * sqrt3 * w0 + (-3 + sqrt3) * wn1 + (3 + sqrt3) * wp1) / (6 * h);
  px[2] = df[2] * ((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wp1 + (-3 + sqrt3) * wp2) / (6 * h);
  px[3] = -df[3] * ((48 * sqrt3 * w0 + 72 * wn1 - 26 * sqrt3 * wn1 - 9 * wn2 + 2 * sqrt3 * wn2 -
                     72 * wp1 - 26 * sqrt3 * wp1 + 9 * wp2 + 2 * sqrt3 * wp2) / (108 * h));
  g2 = 0.0;
  g2x = 0.0;
  for (int k = 0; k < 4; k++) {
    g2 += fw[k] * p[k];
    g2x += fw[k] * px[k];
  }
}

// DF factors from the 5 alphas of a stencil (reconstruction.cpp:1733-1742,
// 2580-2596)
template <class R> inline void df_factors5(R* df, const R* a) {
  R s[4];
  s[0] = a[0] + a[2];
  s[1] = a[1] + a[3];
  s[2] = a[2] + a[4];
  s[3] = a[0] + a[2] + a[4];
  for (int i = 0; i < 4; i++) df[i] = s[i] < 2.0 ? R(1.0) : 2.0 / s[i];
}

// Check_Order_Reduce_by_Lambda_2D: reconstruction.cpp:1439-1449
template <class R> inline bool reduce_order(const R* c, const Gas& g) {
  R lam = lambda_2d(c[0], c[1] / c[0], c[2] / c[0], c[3], g);
  return (lam <= 0.0 || !(lam == lam));
}

// WENO5_AO_with_df: reconstruction.cpp:1731-1859.  One cell, one direction
// (states already rotated to the x-like frame): writes the line state at the
// cell's left face (`left`) and right face (`right`), convar + der1x.
template <class R>
inline void weno5_df_normal(Pt<R>& left, Pt<R>& right, const R* alpha, const R* wn2, const R* wn1,
                            const R* w, const R* wp1, const R* wp2, double h, int recon_var,
                            const Gas& g) {
  R df[4];
  df_factors5(df, alpha);
  R pn1[4], p0[4], pp1[4], bl[4], br[4];
  convar_to_primvar_2d(pn1, wn1, g);
  convar_to_primvar_2d(p0, w, g);
  convar_to_primvar_2d(pp1, wp1, g);
  for (int i = 0; i < 4; i++) {
    bl[i] = 0.5 * (pn1[i] + p0[i]);
    br[i] = 0.5 * (pp1[i] + p0[i]);
  }
  R c[5][4], var[4], d1[4];
  const R* src[5] = {wn2, wn1, w, wp1, wp2};
  for (int side = 0; side < 2; side++) {
    const R* base = side == 0 ? bl : br;
    for (int s = 0; s < 5; s++) {
      if (recon_var == 0)
        for (int i = 0; i < 4; i++) c[s][i] = src[s][i];
      else
        convar_to_char(c[s], base, src[s], g);
    }
    for (int i = 0; i < 4; i++) {
      if (side == 0)
        weno5_df_left(var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h);
      else
        weno5_df_right(var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h);
    }
    Pt<R>& out = side == 0 ? left : right;
    if (recon_var == 0) {
      for (int i = 0; i < 4; i++) {
        out.convar[i] = var[i];
        out.der1x[i] = d1[i];
      }
    } else {
      char_to_convar(out.convar, base, var, g);
      char_to_convar(out.der1x, base, d1, g);
    }
  }
  bool rr = reduce_order(right.convar, g);
  bool rl = reduce_order(left.convar, g);
  if (rl || rr) {
    for (int m = 0; m < 4; m++) {
      right.convar[m] = w[m];
      left.convar[m] = w[m];
      right.der1x[m] = 0.0;
      left.der1x[m] = 0.0;
    }
  }
}

// weno_5th_ao_with_df_tangential with gausspoint == 2:
// reconstruction.cpp:2578-2740.  One side (left or right state) of one face:
// from the 5 line states along the face (wn2..wp2 = line[0..4]) produce the
// two Gauss-point states (convar, der1x, der1y).
template <class R>
inline void weno5_df_tangent_side(Pt<R>* gp, const Pt<R>* line, const R* alpha, double h,
                                  int recon_var, const Gas& g) {
  R df[4];
  df_factors5(df, alpha);
  R base[4];
  convar_to_primvar_2d(base, line[2].convar, g);
  if (recon_var == 0) {
    for (int i = 0; i < 4; i++)
      weno5_df_2gauss(gp[0].convar[i], gp[0].der1y[i], gp[1].convar[i], gp[1].der1y[i], df,
                      line[0].convar[i], line[1].convar[i], line[2].convar[i], line[3].convar[i],
                      line[4].convar[i], h);
  } else {
    R c[5][4];
    for (int s = 0; s < 5; s++) convar_to_char(c[s], base, line[s].convar, g);
    R var[2][4], d1[2][4];
    for (int i = 0; i < 4; i++)
      weno5_df_2gauss(var[0][i], d1[0][i], var[1][i], d1[1][i], df, c[0][i], c[1][i], c[2][i],
                      c[3][i], c[4][i], h);
    for (int q = 0; q < 2; q++) {
      char_to_convar(gp[q].convar, base, var[q], g);
      char_to_convar(gp[q].der1y, base, d1[q], g);
    }
  }
  for (int i = 0; i < 4; i++) {
    R t0, t1;
    weno5_df_2gauss(gp[0].der1x[i], t0, gp[1].der1x[i], t1, df, line[0].der1x[i], line[1].der1x[i],
                    line[2].der1x[i], line[3].der1x[i], line[4].der1x[i], h);
  }
  for (int q = 0; q < 2; q++) {
    if (reduce_order(gp[q].convar, g)) {
      for (int v = 0; v < 4; v++) {
        gp[q].convar[v] = line[2].convar[v];
        gp[q].der1x[v] = line[2].der1x[v];
        gp[q].der1y[v] = 0.0;
      }
    }
  }
}

// Calculate_alpha_k: reconstruction.cpp:1451-1464
template <class R> inline R alpha_k(const R* pl, const R* pr, const Gas& g) {
  R cl = sqrt(g.gamma * pl[3] / pl[0]);
  R cr = sqrt(g.gamma * pr[3] / pr[0]);
  R mln = pl[1] / cl;
  R mlt = pl[2] / cl;
  R mrn = pr[1] / cr;
  R mrt = pr[2] / cr;
  R A = fabs(pl[3] - pr[3]) / pl[3] + fabs(pl[3] - pr[3]) / pr[3];
  R dn = mln - mrn, dtg = mlt - mrt;
  A = A + dn * dn + dtg * dtg;
  return A;
}

}  // namespace orc
