// gks2d_device.hpp — device numerics of the 2D GKS hot path (gfx950, FP64).
//
// Each function restates one reference leaf (cited file:line in the
// reference repository) with the reference's floating-point operation order.
// Two builds include this header:
//  * GKS_FAST == 0 (kernels_ieee.hip, -ffp-contract=off): every division of
//    the reference stays a division, so the result differs from the CPU
//    reference only through the device erfc/exp/sqrt;
//  * GKS_FAST == 1 (kernels_fast.hip, -ffp-contract=fast-honor-pragmas): divisions by
//    constants, by the cell size and by a denominator shared by several
//    quotients become multiplications by a reciprocal formed once (each
//    quotient moves by at most an ulp or two); FMAs are contracted.
// The functions live in a namespace named after the build so the two
// definitions never meet at link time.
//
// The Maxwellian moment tables of the reference (MMDF2d, 3 x 7x7x3 doubles)
// are not materialised: an entry (i,j,k) is the product uplus[i]*vwhole[j]*xi_k
// formed where it is used, which gives the same value the table holds and
// keeps the working set in registers (4 length-7 vectors per Maxwellian).
#pragma once
#include <hip/hip_runtime.h>

#include "gks2d_types.hpp"

#ifndef GKS_VARIANT
#error "GKS_VARIANT must name the kernel build"
#endif
#ifndef GKS_FAST
#error "GKS_FAST must be 0 or 1"
#endif

namespace gksd {
inline namespace GKS_VARIANT {

#ifndef GKS_DI
#define GKS_DI __device__ __forceinline__
#endif
// a / b where b is a compile-time constant
#if GKS_FAST
#define GKS_CDIV(a, b) ((a) * (1.0 / (b)))
#else
#define GKS_CDIV(a, b) ((a) / (b))
#endif
// a / b and 1 / b for a runtime b that is finite, nonzero and normal (a
// density, a pressure, lambda, a sound speed, a sum of positive terms).
// The fast build takes the hardware reciprocal (about 24 bits) through two
// Newton steps and, for a quotient, one residual correction: 5 and 8
// instructions against about 11 for the IEEE division sequence (on 2^20
// random operands over 2^+-60 both agreed with IEEE division bit for bit).
// Unlike IEEE division they give NaN for b = 0 or infinite b, so divisions
// whose denominator can vanish (time_coe's dt / tau_num) stay IEEE.
#if GKS_FAST
GKS_DI double frcp(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double t = fma(-b, r, 1.0);
  r = fma(t, r, r);
  t = fma(-b, r, 1.0);
  return fma(t, r, r);
}
GKS_DI double fdiv(double a, double b) {
  const double r = frcp(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
#define GKS_DIV(a, b) fdiv((a), (b))
#define GKS_RCP(b) frcp(b)
#else
#define GKS_DIV(a, b) ((a) / (b))
#define GKS_RCP(b) (1.0 / (b))
#endif

// ---------------------------------------------------------------- gas algebra
// Convar_to_primvar_2D: basic_function.cpp:811-817 (+ U/V/Pressure :1137-1160)
GKS_DI void convar_to_primvar(double* p, const double* c, const Gas& g) {
  p[0] = c[0];
#if GKS_FAST
  const double ir = GKS_RCP(c[0]);
  p[1] = c[1] * ir;
  p[2] = c[2] * ir;
  p[3] = (g.gamma - 1.0) * (c[3] - 0.5 * c[1] * c[1] * ir - 0.5 * c[2] * c[2] * ir);
#else
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = (g.gamma - 1.0) * (c[3] - 0.5 * c[1] * c[1] / c[0] - 0.5 * c[2] * c[2] / c[0]);
#endif
}
// Lambda (2D): basic_function.cpp:1174-1177
GKS_DI double lambda2(double rho, double u, double v, double rhoE, const Gas& g) {
  return (g.K + 2.0) * 0.25 * GKS_DIV(rho, rhoE - 0.5 * rho * (u * u + v * v));
}
// Convar_to_ULambda_2d: basic_function.cpp:835-842
GKS_DI void convar_to_ulambda(double* p, const double* c, const Gas& g) {
  p[0] = c[0];
#if GKS_FAST
  const double ir = GKS_RCP(c[0]);
  p[1] = c[1] * ir;
  p[2] = c[2] * ir;
#else
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
#endif
  p[3] = lambda2(c[0], p[1], p[2], c[3], g);
}

// Convar_to_char / Char_to_convar (2D): basic_function.cpp:923-992.  The
// 4x4 projections are kept as written (zero and unit entries included) so
// the IEEE build sums in the reference's order.
struct CharBase {
  double u, v, c, overc, alfa, q2, h;
};
GKS_DI void char_base(CharBase& b, const double* prim, const Gas& g) {
  b.c = sqrt(GKS_DIV(g.gamma * prim[3], prim[0]));
  b.overc = GKS_RCP(b.c);
  b.alfa = (g.gamma - 1.0) * 0.5 * b.overc * b.overc;
  b.u = prim[1];
  b.v = prim[2];
  b.q2 = b.u * b.u + b.v * b.v;
}
GKS_DI void to_char(double* ch, const CharBase& b, const double* c) {
  const double u = b.u, v = b.v, alfa = b.alfa, overc = b.overc;
  double s0[4] = {1.0 - alfa * b.q2, 2.0 * alfa * u, 2.0 * alfa * v, -2.0 * alfa};
  double s2[4] = {0.5 * (-u * overc + alfa * b.q2), 0.5 * overc - alfa * u, -alfa * v, alfa};
  double s3[4] = {0.5 * (u * overc + alfa * b.q2), -0.5 * overc - alfa * u, -alfa * v, alfa};
  double a;
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s0[j] * c[j];
  ch[0] = a;
#if GKS_FAST
  ch[1] = c[2] - v * c[0];  // the zero and unit entries dropped
#else
  a = 0.0;
  a = a + (-v) * c[0];
  a = a + 0. * c[1];
  a = a + 1. * c[2];
  a = a + 0. * c[3];
  ch[1] = a;
#endif
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s2[j] * c[j];
  ch[2] = a;
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s3[j] * c[j];
  ch[3] = a;
}
GKS_DI void from_char(double* c, const CharBase& b, const double* ch, const Gas& g) {
  const double u = b.u, v = b.v, cs = b.c;
  const double hq = 0.5 * (u * u + v * v);
  const double s0[4] = {1., 0.0, 1.0, 1.0};
  const double s1[4] = {u, 0., u + cs, u - cs};
  const double s2[4] = {v, 1., v, v};
#if GKS_FAST
  const double h0 = cs * cs * g.inv_gm1;
#else
  const double h0 = cs * cs / (g.gamma - 1.0);
#endif
  const double s3[4] = {hq, v, hq + cs * u + h0, hq - cs * u + h0};
  double a;
#if GKS_FAST
  // the zero entries of the first two rows dropped
  (void)s0;
  (void)s1;
  c[0] = ch[0] + ch[2] + ch[3];
  c[1] = u * ch[0] + (u + cs) * ch[2] + (u - cs) * ch[3];
#else
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s0[j] * ch[j];
  c[0] = a;
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s1[j] * ch[j];
  c[1] = a;
#endif
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s2[j] * ch[j];
  c[2] = a;
  a = 0.0;
#pragma unroll
  for (int j = 0; j < 4; j++) a = a + s3[j] * ch[j];
  c[3] = a;
}

// --------------------------------------------------------- Maxwellian moments
// MMDF2d::calcualte_MMDF: basic_function.cpp:110-187
struct MM {
  double up[7], un[7], vw[7], xi2, xi4;
#if GKS_FAST
  double uw[7];  // whole-space u moments
#endif
};
GKS_DI void mmdf(MM& m, double u, double v, double lambda, const Gas& g) {
  const double pi = 3.14159265358979323846;
  double uw[7];
  uw[0] = 1.0;
  uw[1] = u;
  m.up[0] = 0.5 * erfc(sqrt(lambda) * (-u));
  m.un[0] = 1.0 - m.up[0];
  m.up[1] = u * m.up[0] + 0.5 * GKS_DIV(exp(-lambda * u * u), sqrt(pi * lambda));
  m.un[1] = u - m.up[1];
  const double ol = GKS_RCP(lambda);
#pragma unroll
  for (int i = 2; i <= 6; i++) uw[i] = u * uw[i - 1] + 0.5 * (i - 1) * ol * uw[i - 2];
#pragma unroll
  for (int i = 2; i <= 6; i++) {
    m.up[i] = u * m.up[i - 1] + 0.5 * (i - 1) * ol * m.up[i - 2];
    m.un[i] = uw[i] - m.up[i];
  }
  m.vw[0] = 1.0;
  m.vw[1] = v;
#pragma unroll
  for (int i = 2; i <= 6; i++) m.vw[i] = v * m.vw[i - 1] + 0.5 * (i - 1) * ol * m.vw[i - 2];
  m.xi2 = 0.5 * g.K * ol;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) * ol * ol;
#if GKS_FAST
#pragma unroll
  for (int i = 0; i <= 6; i++) m.uw[i] = uw[i];
#endif
}
// table entries upvxi / unvxi / uvxi [i][j][k]
// Each table is separable, entry (i, j, k) = U[i] V[j] X[k] with X = (1,
// xi2, xi4); U, V, x2, x4 expose the factors (contract_sep).
struct TabP {
  const MM& m;
  GKS_DI double operator()(int i, int j, int k) const {
    return k == 0 ? m.up[i] * m.vw[j] : (k == 1 ? m.up[i] * m.vw[j] * m.xi2 : m.up[i] * m.vw[j] * m.xi4);
  }
  GKS_DI double U(int i) const { return m.up[i]; }
  GKS_DI double V(int j) const { return m.vw[j]; }
  GKS_DI double x2() const { return m.xi2; }
  GKS_DI double x4() const { return m.xi4; }
};
struct TabN {
  const MM& m;
  GKS_DI double operator()(int i, int j, int k) const {
    return k == 0 ? m.un[i] * m.vw[j] : (k == 1 ? m.un[i] * m.vw[j] * m.xi2 : m.un[i] * m.vw[j] * m.xi4);
  }
  GKS_DI double U(int i) const { return m.un[i]; }
  GKS_DI double V(int j) const { return m.vw[j]; }
  GKS_DI double x2() const { return m.xi2; }
  GKS_DI double x4() const { return m.xi4; }
};
// uvxi: the reference forms upvxi + unvxi (basic_function.cpp:161), which is
// uwhole[i] vwhole[j] xi up to rounding; the fast build takes the product of
// the whole-space moments directly (fewer operations and live registers)
struct TabW {
  const MM& m;
  GKS_DI double operator()(int i, int j, int k) const {
#if GKS_FAST
    return k == 0 ? m.uw[i] * m.vw[j] : (k == 1 ? m.uw[i] * m.vw[j] * m.xi2 : m.uw[i] * m.vw[j] * m.xi4);
#else
    return TabP{m}(i, j, k) + TabN{m}(i, j, k);
#endif
  }
#if GKS_FAST
  GKS_DI double U(int i) const { return m.uw[i]; }
  GKS_DI double V(int j) const { return m.vw[j]; }
  GKS_DI double x2() const { return m.xi2; }
  GKS_DI double x4() const { return m.xi4; }
#endif
};

#if GKS_FAST
// contract() at x = 0 through the factors of a separable table: the 13
// products U[u+i] V[v+j] it needs are formed once and the xi terms are
// grouped (S_ij = U[u+i] V[v+j]; T0 = S20 + S02 + xi2 S00 is psi[0]'s a3
// bracket, and so on); the same sums as the term-by-term form, regrouped
template <class T> GKS_DI void contract_sep(int u, int v, double* psi, const double* a, const T& m) {
  const double U0 = m.U(u), U1 = m.U(u + 1), U2 = m.U(u + 2), U3 = m.U(u + 3), U4 = m.U(u + 4);
  const double V0 = m.V(v), V1 = m.V(v + 1), V2 = m.V(v + 2), V3 = m.V(v + 3), V4 = m.V(v + 4);
  const double x2 = m.x2(), x4 = m.x4();
  const double S00 = U0 * V0, S10 = U1 * V0, S01 = U0 * V1, S20 = U2 * V0, S02 = U0 * V2, S11 = U1 * V1;
  const double S30 = U3 * V0, S12 = U1 * V2, S21 = U2 * V1, S03 = U0 * V3;
  const double S40 = U4 * V0, S04 = U0 * V4, S22 = U2 * V2;
  const double E0 = S20 + S02;
  const double T0 = E0 + x2 * S00, T1 = (S30 + S12) + x2 * S10, T2 = (S21 + S03) + x2 * S01;
  const double h3 = 0.5 * a[3];
  psi[0] = a[0] * S00 + a[1] * S10 + a[2] * S01 + h3 * T0;
  psi[1] = a[0] * S10 + a[1] * S20 + a[2] * S11 + h3 * T1;
  psi[2] = a[0] * S01 + a[1] * S11 + a[2] * S02 + h3 * T2;
  psi[3] = 0.5 * (a[0] * T0 + a[1] * T1 + a[2] * T2 + h3 * (S40 + S04 + 2 * S22 + x4 * S00 + 2 * x2 * E0));
}
#endif

// GL_address / GR_address / G_address: basic_function.cpp:603-636
template <class T> GKS_DI void contract(int u, int v, int x, double* psi, const double* a, const T& m) {
#if GKS_FAST
  if (x == 0) {
    contract_sep(u, v, psi, a, m);
    return;
  }
#endif
  psi[0] = a[0] * m(u, v, x) + a[1] * m(u + 1, v, x) + a[2] * m(u, v + 1, x) +
           a[3] * 0.5 * (m(u + 2, v, x) + m(u, v + 2, x) + m(u, v, x + 1));
  psi[1] = a[0] * m(u + 1, v, x) + a[1] * m(u + 2, v, x) + a[2] * m(u + 1, v + 1, x) +
           a[3] * 0.5 * (m(u + 3, v, x) + m(u + 1, v + 2, x) + m(u + 1, v, x + 1));
  psi[2] = a[0] * m(u, v + 1, x) + a[1] * m(u + 1, v + 1, x) + a[2] * m(u, v + 2, x) +
           a[3] * 0.5 * (m(u + 2, v + 1, x) + m(u, v + 3, x) + m(u, v + 1, x + 1));
  psi[3] = 0.5 * (a[0] * (m(u + 2, v, x) + m(u, v + 2, x) + m(u, v, x + 1)) +
                  a[1] * (m(u + 3, v, x) + m(u + 1, v + 2, x) + m(u + 1, v, x + 1)) +
                  a[2] * (m(u + 2, v + 1, x) + m(u, v + 3, x) + m(u, v + 1, x + 1)) +
                  a[3] * 0.5 *
                      (m(u + 4, v, x) + m(u, v + 4, x) + m(u, v, x + 2) + 2 * m(u + 2, v + 2, x) +
                       2 * m(u + 2, v, x + 1) + 2 * m(u, v + 2, x + 1)));
}

// A (2D): basic_function.cpp:571-583
GKS_DI void micro_slope(double* a, const double* der, const double* prim, const Gas& g) {
  double overden = GKS_RCP(prim[0]);
  double R4 = der[3] * overden -
              0.5 * (prim[1] * prim[1] + prim[2] * prim[2] + GKS_DIV(0.5 * (g.K + 2), prim[3])) * der[0] * overden;
  double R3 = (der[2] - prim[2] * der[0]) * overden;
  double R2 = (der[1] - prim[1] * der[0]) * overden;
  a[3] = GKS_DIV(4.0, g.K + 2) * prim[3] * prim[3] * (2 * R4 - 2 * prim[1] * R2 - 2 * prim[2] * R3);
  a[2] = 2 * prim[3] * R3 - prim[2] * a[3];
  a[1] = 2 * prim[3] * R2 - prim[1] * a[3];
  a[0] = der[0] * overden - prim[1] * a[1] - prim[2] * a[2] -
         0.5 * a[3] * (prim[1] * prim[1] + prim[2] * prim[2] + GKS_DIV(0.5 * (g.K + 2), prim[3]));
}

// Get_Tau_NS / Get_Tau: basic_function.cpp:638-708
GKS_DI double tau_ns(double rho0, double lambda0, const Gas& g) {
  if (g.tau_type == 0) return 0.0;
  if (g.mu > 0.0) return GKS_DIV(2.0 * g.mu * lambda0, rho0);
  if (g.nu > 0.0) return 2 * g.nu * lambda0;
  return 0.0;
}
GKS_DI double tau_num(double rl, double rr, double r0, double ll, double lr, double l0, double dt,
                      const Gas& g) {
  if (g.tau_type == 0) {
    if (g.c1 <= 0 && g.c2 <= 0) return 0.0;
    double C = GKS_DIV(g.c2 * fabs(GKS_DIV(rl, ll) - GKS_DIV(rr, lr)), fabs(GKS_DIV(rl, ll) + GKS_DIV(rr, lr)));
    return g.c1 * dt + dt * C;
  }
  double tau_n = GKS_DIV(g.c2 * fabs(GKS_DIV(rl, ll) - GKS_DIV(rr, lr)), fabs(GKS_DIV(rl, ll) + GKS_DIV(rr, lr))) * dt;
  if (tau_n != tau_n) tau_n = 0.0;
  if ((g.mu > 0.0 && g.nu > 0.0) || (g.mu < 0.0 && g.nu < 0.0)) return 0.0;
  if (g.mu > 0.0) return tau_n + GKS_DIV(2.0 * g.mu * l0, r0);
  if (g.nu > 0.0) return tau_n + 2 * g.nu * l0;
  return 0.0;
}
// flux_function.cpp:410-424
GKS_DI void time_coe(double* t, double tau, double tn, double dt) {
  double eta = exp(-dt / tn);
  t[0] = tn * (1 - eta);
  t[1] = tn * (eta * (dt + tn) - tn) + tau * tn * (eta - 1);
  t[2] = tau * tn * (eta - 1);
  t[3] = tn * eta + dt - tn;
  t[4] = tn * (tn - eta * (dt + tn) - tau * (eta - 1)) - dt * tau;
  t[5] = 0.5 * dt * dt - tau * tn * (eta - 1) - tau * dt;
}

// whole-space moments alone (no erfc / exp): the collision state's table in
// the fast build, and the smooth-flow flux
struct MMW {
  double uw[7], vw[7], xi2, xi4;
};
GKS_DI void mmdf_whole(MMW& m, double u, double v, double lambda, const Gas& g) {
  const double ol = GKS_RCP(lambda);
  m.uw[0] = 1.0;
  m.uw[1] = u;
  m.vw[0] = 1.0;
  m.vw[1] = v;
#pragma unroll
  for (int i = 2; i <= 6; i++) {
    m.uw[i] = u * m.uw[i - 1] + 0.5 * (i - 1) * ol * m.uw[i - 2];
    m.vw[i] = v * m.vw[i - 1] + 0.5 * (i - 1) * ol * m.vw[i - 2];
  }
  m.xi2 = 0.5 * g.K * ol;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) * ol * ol;
}
struct TabWhole {
  const MMW& m;
  GKS_DI double operator()(int i, int j, int k) const {
    return k == 0 ? m.uw[i] * m.vw[j] : (k == 1 ? m.uw[i] * m.vw[j] * m.xi2 : m.uw[i] * m.vw[j] * m.xi4);
  }
  GKS_DI double U(int i) const { return m.uw[i]; }
  GKS_DI double V(int j) const { return m.vw[j]; }
  GKS_DI double x2() const { return m.xi2; }
  GKS_DI double x4() const { return m.xi4; }
};
// a strided read-only view (e.g. one face's column of an LDS tile)
struct Strided {
  const double* p;
  int s;
  GKS_DI double operator[](int k) const { return p[k * s]; }
};

// One Gauss point: collision centre state (Center_all_collision_2d_multi,
// reconstruction.cpp:3652-3687) followed by GKS2D with gks2nd_2d
// (flux_function.cpp:345-698).  L/R: convar[0..3], der1x[4..7], der1y[8..11]
// in the face-local frame.  Writes f[4], derf[4]; optionally the centre.
//
// Register plan: every quantity that involves the left/right Maxwellians is
// a sum (left term) + (right term) or a single-side term, so the left side
// is evaluated with only its moments live, then the right side; the centre
// Maxwellian comes after both.  F(dt) and F(dt/2) are summed from the stored
// 4-vectors, each in the reference's own term order
// (flux_function.cpp:443-629 and 650-684).
//
// L and Rr are read through operator[] (a local array, or a strided view of
// LDS): each side's values are read where that side is evaluated, so the
// right side's twelve values are not held through the left side.
template <class V>
GKS_DI void gks_point(double* f, double* derf, double* center, const V& L, const V& Rr, double dt, const Gas& g) {
  double pl[4], pr[4];
  {
    double c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = L[k];
    convar_to_ulambda(pl, c, g);
  }
  const double unit[4] = {1.0, 0.0, 0.0, 0.0};
  double c0[4], c0x[4], c0y[4];
  double glu[4], gru[4], s1[4], Alul[4], Arur[4];
  // ---- left state: u > 0 half moments
  {
    MM ml;
    mmdf(ml, pl[1], pl[2], pl[3], g);
    const TabP tl{ml};
    const TabW wl{ml};
    // left terms of Collision (basic_function.cpp:561-569)
    c0[0] = pl[0] * ml.up[0];
    c0[1] = pl[0] * ml.up[1];
    c0[2] = pl[0] * ml.vw[1] * ml.up[0];
    c0[3] = 0.5 * pl[0] * (ml.up[2] + ml.up[0] * ml.vw[2] + ml.up[0] * ml.xi2);
    double alx[4], aly[4], t4[4];
    {
      double dx[4], dy[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        dx[k] = L[4 + k];
        dy[k] = L[8 + k];
      }
      micro_slope(alx, dx, pl, g);
      micro_slope(aly, dy, pl, g);
    }
    contract(0, 0, 0, t4, alx, tl);
#pragma unroll
    for (int k = 0; k < 4; k++) c0x[k] = pl[0] * t4[k];
    contract(0, 0, 0, t4, aly, tl);
#pragma unroll
    for (int k = 0; k < 4; k++) c0y[k] = pl[0] * t4[k];
    contract(1, 0, 0, glu, unit, tl);
    {
      double a[4], c[4];
      contract(2, 0, 0, a, alx, tl);
      contract(1, 1, 0, c, aly, tl);
#pragma unroll
      for (int i = 0; i < 4; i++) s1[i] = pl[0] * (a[i] + c[i]);
    }
    {
      double xu[4], yv[4], der[4], A[4];
      contract(1, 0, 0, xu, alx, wl);
      contract(0, 1, 0, yv, aly, wl);
#pragma unroll
      for (int i = 0; i < 4; i++) der[i] = -pl[0] * (xu[i] + yv[i]);
      micro_slope(A, der, pl, g);
      contract(1, 0, 0, Alul, A, tl);
    }
  }
  // keep the scheduler from interleaving the two sides (it would raise the
  // number of live registers)
  __builtin_amdgcn_sched_barrier(0);
  // ---- right state: u < 0 half moments
  {
    double c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) c[k] = Rr[k];
    convar_to_ulambda(pr, c, g);
  }
  {
    MM mr;
    mmdf(mr, pr[1], pr[2], pr[3], g);
    const TabN tr{mr};
    const TabW wr{mr};
    c0[0] = c0[0] + pr[0] * mr.un[0];
    c0[1] = c0[1] + pr[0] * mr.un[1];
    c0[2] = c0[2] + pr[0] * mr.vw[1] * mr.un[0];
    c0[3] = c0[3] + 0.5 * pr[0] * (mr.un[2] + mr.un[0] * mr.vw[2] + mr.un[0] * mr.xi2);
    double arx[4], ary[4], t4[4];
    {
      double dx[4], dy[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        dx[k] = Rr[4 + k];
        dy[k] = Rr[8 + k];
      }
      micro_slope(arx, dx, pr, g);
      micro_slope(ary, dy, pr, g);
    }
    contract(0, 0, 0, t4, arx, tr);
#pragma unroll
    for (int k = 0; k < 4; k++) c0x[k] = c0x[k] + pr[0] * t4[k];
    contract(0, 0, 0, t4, ary, tr);
#pragma unroll
    for (int k = 0; k < 4; k++) c0y[k] = c0y[k] + pr[0] * t4[k];
    contract(1, 0, 0, gru, unit, tr);
    {
      double b[4], d[4];
      contract(2, 0, 0, b, arx, tr);
      contract(1, 1, 0, d, ary, tr);
#pragma unroll
      for (int i = 0; i < 4; i++) s1[i] = s1[i] + pr[0] * (b[i] + d[i]);
    }
    {
      double xu[4], yv[4], der[4], A[4];
      contract(1, 0, 0, xu, arx, wr);
      contract(0, 1, 0, yv, ary, wr);
#pragma unroll
      for (int i = 0; i < 4; i++) der[i] = -pr[0] * (xu[i] + yv[i]);
      micro_slope(A, der, pr, g);
      contract(1, 0, 0, Arur, A, tr);
    }
  }
  if (center) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      center[k] = c0[k];
      center[4 + k] = c0x[k];
      center[8 + k] = c0y[k];
    }
  }

  __builtin_amdgcn_sched_barrier(0);
  // equilibrium (centre) part
  double p0[4];
  convar_to_ulambda(p0, c0, g);
  double g0u[4], s4[4], A0u[4];
  double qs_raw = 0.0;  // Prandtl-fix heat-flux term (flux_function.cpp:630-636)
  {
    // the collision state's moments enter only through the whole-space
    // table; the fast build forms it without the half-space integrals
#if GKS_FAST
    MMW m0;
    mmdf_whole(m0, p0[1], p0[2], p0[3], g);
    const TabWhole w0{m0};
#else
    MM m0;
    mmdf(m0, p0[1], p0[2], p0[3], g);
    const TabW w0{m0};
#endif
    const double unit[4] = {1.0, 0.0, 0.0, 0.0};
    contract(1, 0, 0, g0u, unit, w0);
    double a0x[4], a0y[4], a0xu[4], a0xuu[4], a0yv[4], a0yuv[4];
    micro_slope(a0x, c0x, p0, g);
    contract(1, 0, 0, a0xu, a0x, w0);
    contract(2, 0, 0, a0xuu, a0x, w0);
    micro_slope(a0y, c0y, p0, g);
    contract(0, 1, 0, a0yv, a0y, w0);
    contract(1, 1, 0, a0yuv, a0y, w0);
#pragma unroll
    for (int i = 0; i < 4; i++) s4[i] = a0xuu[i] + a0yuv[i];
    double derA0[4], A0[4];
#pragma unroll
    for (int i = 0; i < 4; i++) derA0[i] = -p0[0] * (a0xu[i] + a0yv[i]);
    micro_slope(A0, derA0, p0, g);
    contract(1, 0, 0, A0u, A0, w0);
    if (g.prandtl_fix)
      qs_raw = a0xuu[3] + A0u[3] - p0[1] * (a0xuu[1] + A0u[1]) - p0[2] * (a0xuu[2] + A0u[2]);
  }

  const double tau = tau_ns(p0[0], p0[3], g);
  double t[6];
  // F(dt): flux_function.cpp:406-629
  double F0[4];
  time_coe(t, tau, tau_num(pl[0], pr[0], p0[0], pl[3], pr[3], p0[3], dt, g), dt);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    F0[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
    F0[i] = F0[i] + p0[0] * t[3] * g0u[i];
    F0[i] = F0[i] + t[1] * s1[i];
    F0[i] = F0[i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);
    F0[i] = F0[i] + p0[0] * t[4] * s4[i];
    F0[i] = F0[i] + p0[0] * t[5] * (A0u[i]);
  }
  if (g.prandtl_fix) {
    double qs = qs_raw;
    qs *= -p0[0] * tau * dt;
    F0[3] += (1.0 / g.pr - 1.0) * qs;
  }
  // F(dt/2): flux_function.cpp:648-684
  const double dt2 = 0.5 * dt;
  time_coe(t, tau, tau_num(pl[0], pr[0], p0[0], pl[3], pr[3], p0[3], dt2, g), dt2);
  double F1[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    F1[i] = t[0] * (pl[0] * glu[i] + pr[0] * gru[i]);
    F1[i] = F1[i] + t[1] * s1[i];
    F1[i] = F1[i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);
    F1[i] = F1[i] + p0[0] * t[3] * g0u[i];
    F1[i] = F1[i] + p0[0] * t[4] * s4[i];
    F1[i] = F1[i] + p0[0] * t[5] * (A0u[i]);
  }
  if (g.prandtl_fix) {
    double qs = qs_raw;
    qs *= -p0[0] * tau * 0.5 * dt;
    F1[3] += (1.0 / g.pr - 1.0) * qs;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f[i] = (4.0 * F1[i] - F0[i]);
    derf[i] = 4.0 * (F0[i] - 2.0 * F1[i]);
  }
}

// ------------------------------------------------------------- WENO5-AO (+DF)
// ------------------------------------------------- smooth-flow GKS flux
// No counterpart in the reference: the flux of a Gauss point whose two
// states both come from linear-weight (smooth) reconstructions.  For
// continuous data the BGK solution at the interface is the equilibrium
// g0 [1 - tau (a u + b v + A) + t A] of the interface state (Xu's smooth
// GKS flux); here the interface state and its slopes are the averages of
// the two sides.  Its time integrals F(T) = rho0 [T g0u - tau T s4
// + (T^2/2 - tau T) A0u] are what GKS2D's (dt, dt/2) combination gives when
// the two states coincide (then the numerical collision time tau_num drops
// out: t0 + t3 = T, t1 + t4 = -tau T, t2 + t5 = T^2/2 - tau T), so
// f = 4 F(dt/2) - F(dt) = rho0 dt [g0u - tau (s4 + A0u)] and
// derf = 4 (F(dt) - 2 F(dt/2)) = rho0 dt^2 A0u.  Only whole-space moments:
// no erfc or exp.
template <class V>
GKS_DI void gks_point_smooth(double* f, double* derf, const V& L, const V& Rr, double dt, const Gas& g) {
  double c0[4], c0x[4], c0y[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    c0[k] = 0.5 * (L[k] + Rr[k]);
    c0x[k] = 0.5 * (L[4 + k] + Rr[4 + k]);
    c0y[k] = 0.5 * (L[8 + k] + Rr[8 + k]);
  }
  double p0[4];
  convar_to_ulambda(p0, c0, g);
  MMW m0;
  mmdf_whole(m0, p0[1], p0[2], p0[3], g);
  const TabWhole w0{m0};
  const double unit[4] = {1.0, 0.0, 0.0, 0.0};
  double g0u[4], a0x[4], a0y[4], a0xu[4], a0xuu[4], a0yv[4], a0yuv[4], derA0[4], A0[4], A0u[4];
  const double tau = tau_ns(p0[0], p0[3], g);
  contract(1, 0, 0, g0u, unit, w0);
  micro_slope(a0x, c0x, p0, g);
  contract(1, 0, 0, a0xu, a0x, w0);
  micro_slope(a0y, c0y, p0, g);
  contract(0, 1, 0, a0yv, a0y, w0);
  // the second-order slope terms s4 = a0xuu + a0yuv enter f only through
  // tau * s4 (and the heat-flux correction): with tau = 0 (the Euler
  // collision time) they are not formed
  if (tau != 0.0 || g.prandtl_fix) {
    contract(2, 0, 0, a0xuu, a0x, w0);
    contract(1, 1, 0, a0yuv, a0y, w0);
  } else {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      a0xuu[i] = 0.0;
      a0yuv[i] = 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) derA0[i] = -p0[0] * (a0xu[i] + a0yv[i]);
  micro_slope(A0, derA0, p0, g);
  contract(1, 0, 0, A0u, A0, w0);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    f[i] = p0[0] * dt * (g0u[i] - tau * ((a0xuu[i] + a0yuv[i]) + A0u[i]));
    derf[i] = p0[0] * dt * dt * A0u[i];
  }
  if (g.prandtl_fix) {  // the heat-flux correction of GKS2D, integrated the same way
    const double qs = a0xuu[3] + A0u[3] - p0[1] * (a0xuu[1] + A0u[1]) - p0[2] * (a0xuu[2] + A0u[2]);
    f[3] += (1.0 / g.pr - 1.0) * (-p0[0] * tau * dt * qs);
  }
}

// smoothness indicators and AO weights (reconstruction.cpp:480-511, 2742-2776)
GKS_DI void weno5_weights(double* fw, double wn2, double wn1, double w0, double wp1, double wp2) {
  const double dhi = 0.85, dlo = 0.85;
  const double eps = 1e-6;
  const double d[4] = {(1.0 - dhi) * (1.0 - dlo) / 2.0, (1.0 - dhi) * dlo, (1.0 - dhi) * (1.0 - dlo) / 2.0, dhi};
  double q, q2, beta[4];
  q = wn2 - 2.0 * wn1 + w0;
  q2 = wn2 - 4.0 * wn1 + 3.0 * w0;
  beta[0] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = wn1 - 2.0 * w0 + wp1;
  q2 = wn1 - wp1;
  beta[1] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = w0 - 2.0 * wp1 + wp2;
  q2 = 3.0 * w0 - 4.0 * wp1 + wp2;
  beta[2] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  beta[3] = 1.0 / 6.0 * (beta[0] + 4.0 * beta[1] + beta[2]) + fabs(beta[0] - beta[2]);
  const double tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  double alpha[4], sum = 0.0;
#if GKS_FAST
  // the four quotients tau5 / (beta_k + eps) from one reciprocal of the
  // product of the denominators (each >= eps, the product stays in range)
  double ib[4];
  {
    const double b0 = beta[0] + eps, b1 = beta[1] + eps, b2 = beta[2] + eps, b3 = beta[3] + eps;
    const double p01 = b0 * b1, p23 = b2 * b3;
    const double r = fdiv(tau5, p01 * p23);
    ib[0] = r * b1 * p23;
    ib[1] = r * b0 * p23;
    ib[2] = r * p01 * b3;
    ib[3] = r * p01 * b2;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double gd = ib[i];
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum += alpha[i];
  }
#else
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double gd = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + gd * gd);
    sum += alpha[i];
  }
#endif
  double ww[4];
#if GKS_FAST
  const double isum = frcp(sum);
#pragma unroll
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] * isum;
#else
#pragma unroll
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum;
#endif
  const double w3d = GKS_CDIV(ww[3], d[3]);
  fw[3] = w3d;
#pragma unroll
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - w3d * d[k];
}
// weno_5th_ao_with_df_left: reconstruction.cpp:480-567
GKS_DI void weno5_left(double& var, double& der1, double wn2, double wn1, double w0, double wp1, double wp2,
                       const double* df, double h, double ih) {
  double fw[4];
  weno5_weights(fw, wn2, wn1, w0, wp1, wp2);
  double p[4], px[4];
  p[0] = w0 + df[0] * (-1.0 / 6.0 * wn2 + 5.0 / 6.0 * wn1 + 1.0 / 3.0 * w0 - w0);
  p[1] = w0 + df[1] * (1.0 / 3.0 * wn1 + 5.0 / 6.0 * w0 - 1.0 / 6.0 * wp1 - w0);
  p[2] = w0 + df[2] * (11.0 / 6.0 * w0 - 7.0 / 6.0 * wp1 + 1.0 / 3.0 * wp2 - w0);
  p[3] = w0 + df[3] * ((1.0 / 60.0) * (47.0 * w0 + 27.0 * wn1 - 3.0 * wn2 - 13.0 * wp1 + 2.0 * wp2) - w0);
#if GKS_FAST
  px[0] = df[0] * (w0 - wn1) * ih;
  px[1] = df[1] * (w0 - wn1) * ih;
  px[2] = -df[2] * ((2.0 * w0 - 3.0 * wp1 + wp2) * ih);
  px[3] = df[3] * (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) * (ih * (1.0 / 12.0));
#else
  (void)ih;
  px[0] = df[0] * (w0 - wn1) / h;
  px[1] = df[1] * (w0 - wn1) / h;
  px[2] = -df[2] * ((2.0 * w0 - 3.0 * wp1 + wp2) / h);
  px[3] = df[3] * (15.0 * w0 - 15.0 * wn1 + wn2 - wp1) / (12.0 * h);
#endif
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
// weno_5th_ao_with_df_right: reconstruction.cpp:569-657
GKS_DI void weno5_right(double& var, double& der1, double wn2, double wn1, double w0, double wp1, double wp2,
                        const double* df, double h, double ih) {
  double fw[4];
  weno5_weights(fw, wn2, wn1, w0, wp1, wp2);
  double p[4], px[4];
  p[0] = w0 + df[0] * (1.0 / 3.0 * wn2 - 7.0 / 6.0 * wn1 + 11.0 / 6.0 * w0 - w0);
  p[1] = w0 + df[1] * (-1.0 / 6.0 * wn1 + 5.0 / 6.0 * w0 + 1.0 / 3.0 * wp1 - w0);
  p[2] = w0 + df[2] * (1.0 / 3.0 * w0 + 5.0 / 6.0 * wp1 - 1.0 / 6.0 * wp2 - w0);
  p[3] = w0 + df[3] * ((1.0 / 60.0) * (47.0 * w0 - 13.0 * wn1 + 2.0 * wn2 + 27.0 * wp1 - 3.0 * wp2) - w0);
#if GKS_FAST
  px[0] = df[0] * (2.0 * w0 - 3.0 * wn1 + wn2) * ih;
  px[1] = df[1] * (-w0 + wp1) * ih;
  px[2] = df[2] * (-w0 + wp1) * ih;
  px[3] = df[3] * (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) * (ih * (1.0 / 12.0));
#else
  (void)ih;
  px[0] = df[0] * (2.0 * w0 - 3.0 * wn1 + wn2) / h;
  px[1] = df[1] * (-w0 + wp1) / h;
  px[2] = df[2] * (-w0 + wp1) / h;
  px[3] = df[3] * (-15.0 * w0 + wn1 + 15.0 * wp1 - wp2) / (12.0 * h);
#endif
  var = 0.0;
  der1 = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}
// weno_5th_ao_with_df_2gauss: reconstruction.cpp:2742-2838
GKS_DI void weno5_2gauss(double& g1, double& g1x, double& g2, double& g2x, const double* df, double wn2,
                         double wn1, double w0, double wp1, double wp2, double h, double ih) {
  double fw[4];
  weno5_weights(fw, wn2, wn1, w0, wp1, wp2);
  const double sqrt3 = 1.732050807568877294;
  // the reference's divisors: divisions in the IEEE build, multiplications
  // by reciprocals formed once in the fast build
#if GKS_FAST
  const double r6h = ih * (1.0 / 6.0), r108h = ih * (1.0 / 108.0);
#define GKS_D4S3 *(1.0 / (4 * sqrt3))
#define GKS_D4320 *(1.0 / 4320)
#define GKS_D6H *r6h
#define GKS_D108H *r108h
#else
  (void)ih;
#define GKS_D4S3 / (4 * sqrt3)
#define GKS_D4320 / 4320
#define GKS_D6H / (6 * h)
#define GKS_D108H / (108 * h)
#endif
  double p[4], px[4];
#if GKS_FAST
  // the same candidate polynomials with the coefficients of each stencil
  // value combined at compile time (the written form repeats a value in
  // several terms, e.g. - wn2 - 70 sqrt3 wn2)
  {
    constexpr double s3 = 1.732050807568877294;
    constexpr double c43 = 1.0 / (4 * s3), q = 1.0 / 4320;
    constexpr double a0 = 1 - s3 / 4, a1 = 1 / s3, a2 = -c43;  // w0, wn1, wn2 (and mirrored)
    constexpr double b0 = (4 + s3) / 4, b1 = -s3 / 3, b2 = s3 / 12;
    constexpr double e0 = 4314 * q, e1 = (4 + 500 * s3) * q, e2 = (-1 - 70 * s3) * q, e3 = (4 - 500 * s3) * q,
                     e4 = (-1 + 70 * s3) * q;
    constexpr double k0 = -9 + s3, k1 = -2 * (-6 + s3), k2 = -3 + s3;      // px[0] of point 1
    constexpr double m0 = -2 * s3, m1 = 3 + s3, m2 = -3 + s3;             // px[1] of point 1
    constexpr double n0 = 9 + s3, n1 = -2 * (6 + s3), n2 = 3 + s3;        // px[2] of point 1
    constexpr double t0 = 48 * s3, t1 = -72 - 26 * s3, t2 = 9 + 2 * s3, t3 = 72 - 26 * s3, t4 = -9 + 2 * s3;
    p[0] = w0 + df[0] * ((a0 * w0 + a1 * wn1 + a2 * wn2) - w0);
    p[1] = w0 + df[1] * ((w0 + c43 * wn1 - c43 * wp1) - w0);
    p[2] = w0 + df[2] * ((b0 * w0 + b1 * wp1 + b2 * wp2) - w0);
    p[3] = w0 + df[3] * ((e0 * w0 + e1 * wn1 + e2 * wn2 + e3 * wp1 + e4 * wp2) - w0);
    px[0] = -df[0] * ((k0 * w0 + k1 * wn1 + k2 * wn2) * r6h);
    px[1] = -df[1] * ((m0 * w0 + m1 * wn1 + m2 * wp1) * r6h);
    px[2] = -df[2] * ((n0 * w0 + n1 * wp1 + n2 * wp2) * r6h);
    px[3] = df[3] * ((t0 * w0 + t1 * wn1 + t2 * wn2 + t3 * wp1 + t4 * wp2) * r108h);
    g1 = 0.0;
    g1x = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      g1 += fw[k] * p[k];
      g1x += fw[k] * px[k];
    }
    p[0] = w0 + df[0] * ((b0 * w0 + b1 * wn1 + b2 * wn2) - w0);
    p[1] = w0 + df[1] * ((w0 - c43 * wn1 + c43 * wp1) - w0);
    p[2] = w0 + df[2] * ((a0 * w0 + a1 * wp1 + a2 * wp2) - w0);
    p[3] = w0 + df[3] * ((e0 * w0 + e3 * wn1 + e4 * wn2 + e1 * wp1 + e2 * wp2) - w0);
    px[0] = df[0] * ((n0 * w0 + n1 * wn1 + n2 * wn2) * r6h);
    px[1] = df[1] * ((m0 * w0 + m2 * wn1 + m1 * wp1) * r6h);
    px[2] = df[2] * ((k0 * w0 + k1 * wp1 + k2 * wp2) * r6h);
    px[3] = -df[3] * ((t0 * w0 + t3 * wn1 + t4 * wn2 + t1 * wp1 + t2 * wp2) * r108h);
    g2 = 0.0;
    g2x = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      g2 += fw[k] * p[k];
      g2x += fw[k] * px[k];
    }
    return;
  }
#endif
  p[0] = w0 + df[0] * (w0 - (sqrt3 * w0) / 4 + (4 * wn1 - wn2) GKS_D4S3 - w0);
  p[1] = w0 + df[1] * (w0 + (wn1 - wp1) GKS_D4S3 - w0);
  p[2] = w0 + df[2] * ((1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wp1 + wp2)) - w0);
  p[3] = w0 + df[3] * ((4314 * w0 + (4 + 500 * sqrt3) * wn1 - wn2 - 70 * sqrt3 * wn2 + 4 * wp1 -
                        500 * sqrt3 * wp1 - wp2 + 70 * sqrt3 * wp2) GKS_D4320 - w0);
  px[0] = -df[0] * ((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wn1 + (-3 + sqrt3) * wn2) GKS_D6H;
  px[1] = -df[1] * (-2 * sqrt3 * w0 + (3 + sqrt3) * wn1 + (-3 + sqrt3) * wp1) GKS_D6H;
  px[2] = -df[2] * ((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wp1 + (3 + sqrt3) * wp2) GKS_D6H;
  px[3] = df[3] * (48 * sqrt3 * w0 - 72 * wn1 - 26 * sqrt3 * wn1 + 9 * wn2 + 2 * sqrt3 * wn2 + 72 * wp1 -
                   26 * sqrt3 * wp1 - 9 * wp2 + 2 * sqrt3 * wp2) GKS_D108H;
  g1 = 0.0;
  g1x = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    g1 += fw[k] * p[k];
    g1x += fw[k] * px[k];
  }
  p[0] = w0 + df[0] * ((1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wn1 + wn2)) - w0);
  p[1] = w0 + df[1] * (w0 + (-wn1 + wp1) GKS_D4S3 - w0);
  p[2] = w0 + df[2] * (w0 - (sqrt3 * w0) / 4 + (4 * wp1 - wp2) GKS_D4S3 - w0);
  p[3] = w0 + df[3] * ((4314 * w0 + (4 - 500 * sqrt3) * wn1 - wn2 + 70 * sqrt3 * wn2 + 4 * wp1 +
                        500 * sqrt3 * wp1 - wp2 - 70 * sqrt3 * wp2) GKS_D4320 - w0);
  px[0] = df[0] * ((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wn1 + (3 + sqrt3) * wn2) GKS_D6H;
  px[1] = df[1] * (-2 * sqrt3 * w0 + (-3 + sqrt3) * wn1 + (3 + sqrt3) * wp1) GKS_D6H;
  px[2] = df[2] * ((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wp1 + (-3 + sqrt3) * wp2) GKS_D6H;
  px[3] = -df[3] * ((48 * sqrt3 * w0 + 72 * wn1 - 26 * sqrt3 * wn1 - 9 * wn2 + 2 * sqrt3 * wn2 - 72 * wp1 -
                     26 * sqrt3 * wp1 + 9 * wp2 + 2 * sqrt3 * wp2) GKS_D108H);
  g2 = 0.0;
  g2x = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    g2 += fw[k] * p[k];
    g2x += fw[k] * px[k];
  }
}
#undef GKS_D4S3
#undef GKS_D4320
#undef GKS_D6H
#undef GKS_D108H

// DF factors: reconstruction.cpp:1733-1742
GKS_DI void df5(double* df, const double* a) {
  const double s[4] = {a[0] + a[2], a[1] + a[3], a[2] + a[4], a[0] + a[2] + a[4]};
#pragma unroll
  for (int i = 0; i < 4; i++) df[i] = s[i] < 2.0 ? 1.0 : GKS_DIV(2.0, s[i]);
}
// Check_Order_Reduce_by_Lambda_2D: reconstruction.cpp:1439-1449
GKS_DI bool reduce_order(const double* c, const Gas& g) {
#if GKS_FAST
  const double ir = frcp(c[0]);
  double lam = lambda2(c[0], c[1] * ir, c[2] * ir, c[3], g);
#else
  double lam = lambda2(c[0], c[1] / c[0], c[2] / c[0], c[3], g);
#endif
  return (lam <= 0.0 || !(lam == lam));
}

// Calculate_alpha_k: reconstruction.cpp:1451-1464 (inputs primitive)
GKS_DI double alpha_k(const double* pl, const double* pr, const Gas& g) {
  double cl = sqrt(GKS_DIV(g.gamma * pl[3], pl[0]));
  double cr = sqrt(GKS_DIV(g.gamma * pr[3], pr[0]));
  double mln = GKS_DIV(pl[1], cl), mlt = GKS_DIV(pl[2], cl);
  double mrn = GKS_DIV(pr[1], cr), mrt = GKS_DIV(pr[2], cr);
  double A = GKS_DIV(fabs(pl[3] - pr[3]), pl[3]) + GKS_DIV(fabs(pl[3] - pr[3]), pr[3]);
  double dn = mln - mrn, dtg = mlt - mrt;
  A = A + dn * dn + dtg * dtg;
  return A;
}

}  // namespace GKS_VARIANT
}  // namespace gksd
