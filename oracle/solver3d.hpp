// solver3d.hpp — CPU restatement of the 3D path (TEST INFRASTRUCTURE ONLY:
// the checker of the HIP 3D kernels; the product never links it).
//
// The reference's 3D path (CubicTube, 3Dproblem.cpp:3-172) cannot run as
// shipped: Interface3d::gauss is never allocated (basic_function.cpp:
// 1896-1904) but written by WENO5_AO_of_point_value
// (reconstruction.cpp:4147-4161), and GKS3D feeds the pressure where the
// Maxwellian's lambda belongs (flux_function.cpp:832-838: Convar_to_primvar_3D
// returns (rho, u, v, w, p), and p goes into MMDF3d_left/right, A, Get_Tau).
// This restatement keeps the reference's 3D algorithm and operation order —
// face values by WENO5-AO across the face (reconstruction.cpp:3774-3923),
// line values along the first tangential direction
// (WENO5_AO_of_line_value, :3925-4145), Gauss-point values along the second
// (WENO5_AO_of_point_value, :4147-4399), GKS3D with gks2nd and the S2O4
// (dt, dt/2) combination (flux_function.cpp:813-1163), Update_gauss
// (time_advance.cpp:850-999), Get_CFL / Dtx_3d (:793-848) — with lambda
// computed from the conservative state (Lambda, basic_function.cpp:1179-1182)
// as the 2D and 1D fluxes do.  Two more departures: the order-reduction
// fallback of the line and point values is decided afresh at every
// reconstruction (the reference's Recon3d::order_reduce, once set, is never
// cleared), and the CFL takes the maximum speed and the minimum density over
// the interior (order-free) instead of the sequential Dtx_3d scan.  Periodic
// boundaries are added (the reference has free boundaries only in 3D).
//
// Pinned (tests/test_oracle_3d.py) against the reference's own compiled 3D
// code through oracle/ref_harness.cpp: the face, line and point leaves
// (WENO5_AO, WENO5_AO_for_line_value, WENO5_AO_for_point_value), GKS3D on
// states whose pressure equals lambda (where its bug is silent), and the
// whole driver loop with the Gauss-point records allocated and this file's
// flux in flux_function_3d; further against the reference's GKS2D on planar
// data and its 2D non-DF run on z-invariant data.
//
// Layout: cell (i, j, k) at (i*ny + j)*nz + k (z fastest, as the
// reference); faces of each direction at (i*(ny+1) + j)*(nz+1) + k.
// Gauss point q = 2 * line + point (SetGuassPoint_3D, fluid_mesh.cpp:80-88;
// weights 1/4).
#pragma once
// The leaves are templates on the number type (double here; a counting
// scalar in oracle/flopcount3d.cpp); math calls are unqualified so that
// argument-dependent lookup finds the counting overloads.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc3d {

struct Gas3 {
  int K;
  double gamma, mu, nu, c1, c2;
  int tau_type;  // 0 Euler, 1 NS
};

// ------------------------------------------------------------ gas algebra
// Convar_to_primvar_3D / Pressure: basic_function.cpp:819-826, 1162-1167
template <class Real>
inline void convar_to_primvar(Real* p, const Real* c, const Gas3& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = c[3] / c[0];
  p[4] = (g.gamma - 1.0) * (c[4] - 0.5 * c[1] * c[1] / c[0] - 0.5 * c[2] * c[2] / c[0] - 0.5 * c[3] * c[3] / c[0]);
}
// Lambda (3D): basic_function.cpp:1179-1182
template <class Real>
inline Real lambda3(Real rho, Real u, Real v, Real w, Real rhoE, const Gas3& g) {
  return (g.K + 3.0) * 0.25 * (rho / (rhoE - 0.5 * rho * (u * u + v * v + w * w)));
}
// (rho, u, v, w, lambda), the state GKS3D's moments need
template <class Real>
inline void convar_to_ulambda(Real* p, const Real* c, const Gas3& g) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = c[3] / c[0];
  p[4] = lambda3(c[0], p[1], p[2], p[3], c[4], g);
}
// Primvar_to_convar_3D: basic_function.cpp:859-866, Q_densityE :1131-1135
template <class Real>
inline void primvar_to_convar(Real* c, const Real* p, const Gas3& g) {
  c[0] = p[0];
  c[1] = p[0] * p[1];
  c[2] = p[0] * p[2];
  c[3] = p[0] * p[3];
  c[4] = p[0] * (p[4] / p[0] / (g.gamma - 1) + 0.5 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]));
}
// Ydirection / Zdirection: basic_function.cpp:1208-1224
template <class Real>
inline void ydir(Real* o, const Real* f) {
  o[0] = f[0];
  o[1] = f[2];
  o[2] = -f[1];
  o[3] = f[3];
  o[4] = f[4];
}
template <class Real>
inline void zdir(Real* o, const Real* f) {
  o[0] = f[0];
  o[1] = f[3];
  o[2] = f[2];
  o[3] = -f[1];
  o[4] = f[4];
}
// Char_base_3D: basic_function.cpp:994-1035
template <class Real>
inline void char_base(Real* s, const Real* prim, const Gas3& g) {
  const double Gamma = g.gamma;
  Real c = sqrt(Gamma * prim[4] / prim[0]);
  Real overc = 1.0 / c;
  Real alfa = (Gamma - 1.) * 0.5 * overc * overc;
  Real u = prim[1], v = prim[2], w = prim[3];
  const double r_over = 1 / (Gamma - 1);
  Real H = 0.5 * (u * u + v * v + w * w) + 0.5 / alfa;
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
  for (int i = 0; i < 25; i++) s[i] = s[i] * alfa;
}
// Convar_to_char_3D: basic_function.cpp:1037-1050
template <class Real>
inline void to_char(Real* ch, const Real* s, const Real* c) {
  for (int i = 0; i < 5; i++) ch[i] = 0.0;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) ch[i] = ch[i] + s[i * 5 + j] * c[j];
}
// Char_to_convar_3D: basic_function.cpp:1052-1093
template <class Real>
inline void from_char(Real* c, const Real* prim, const Real* ch, const Gas3& g) {
  const double Gamma = g.gamma;
  Real cs = sqrt(Gamma * prim[4] / prim[0]);
  Real alfa = (Gamma - 1.) / (2.0 * cs * cs);
  Real u = prim[1], v = prim[2], w = prim[3];
  Real H = 0.5 * (u * u + v * v + w * w) + 0.5 / alfa;
  Real s[5][5] = {{1., 1.0, 0.0, 0.0, 1.0},
                    {u - cs, u, 0, 0, u + cs},
                    {v, v, 1, 0, v},
                    {w, w, 0, 1, w},
                    {H - u * cs, 0.5 * (u * u + v * v + w * w), v, w, H + u * cs}};
  for (int i = 0; i < 5; i++) {
    c[i] = 0;
    for (int j = 0; j < 5; j++) c[i] = c[i] + s[i][j] * ch[j];
  }
}
// Check_Order_Reduce_by_Lambda_3D: reconstruction.cpp:3755-3771
template <class Real>
inline bool reduce_order(const Real* c, const Gas3& g) {
  Real l = lambda3(c[0], c[1] / c[0], c[2] / c[0], c[3] / c[0], c[4], g);
  return l <= 0.0 || l != l;
}

// ------------------------------------------------------------ WENO5-AO
// pow(x, 2) of the reference: the square, correctly rounded (x * x)
template <class Real>
inline Real sq(const Real& x) {
  return x * x;
}
// weno_5th_ao_left / right (wenotype wenoz), reconstruction.cpp:161-358:
// value and first derivative at the cell's left / right face
template <class Real>
inline void weno5_weights_nodf(Real* fw, Real wn2, Real wn1, Real w0, Real wp1, Real wp2, double eps) {
  const double dhi = 0.85, dlo = 0.85;
  Real beta[4], ww[4], alpha[4];
  double d[4];
  d[0] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[1] = (1.0 - dhi) * dlo;
  d[2] = (1.0 - dhi) * (1.0 - dlo) / 2.0;
  d[3] = dhi;
  beta[0] = 13.0 / 12.0 * sq((wn2 - 2.0 * wn1 + w0)) + 0.25 * sq((wn2 - 4.0 * wn1 + 3.0 * w0));
  beta[1] = 13.0 / 12.0 * sq((wn1 - 2.0 * w0 + wp1)) + 0.25 * sq((wn1 - wp1));
  beta[2] = 13.0 / 12.0 * sq((w0 - 2.0 * wp1 + wp2)) + 0.25 * sq((3.0 * w0 - 4.0 * wp1 + wp2));
  beta[3] = (1.0 / 5040.0) * (231153.0 * w0 * w0 + 104963.0 * wn1 * wn1 + 6908.0 * wn2 * wn2 -
                              38947.0 * wn2 * wp1 + 104963.0 * wp1 * wp1 +
                              wn1 * (-51001.0 * wn2 + 179098.0 * wp1 - 38947.0 * wp2) -
                              3.0 * w0 * (99692.0 * wn1 - 22641.0 * wn2 + 99692.0 * wp1 - 22641.0 * wp2) +
                              8209.0 * wn2 * wp2 - 51001.0 * wp1 * wp2 + 6908.0 * wp2 * wp2);
  Real tau5 = 1.0 / 3.0 * (fabs(beta[3] - beta[0]) + fabs(beta[3] - beta[1]) + fabs(beta[3] - beta[2]));
  Real sum_alpha = 0.0;
  for (int i = 0; i < 4; i++) {
    Real global_div = tau5 / (beta[i] + eps);
    alpha[i] = d[i] * (1.0 + global_div * global_div);
    sum_alpha += alpha[i];
  }
  for (int k = 0; k < 4; k++) ww[k] = alpha[k] / sum_alpha;
  fw[3] = ww[3] / d[3];
  for (int k = 0; k < 3; k++) fw[k] = ww[k] - ww[3] / d[3] * d[k];
}
template <class Real>
inline void weno5_left(Real& var, Real& der1, Real wn2, Real wn1, Real w0, Real wp1, Real wp2, double h) {
  Real fw[4], p[4], px[4];
  weno5_weights_nodf(fw, wn2, wn1, w0, wp1, wp2, 1e-10);
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
template <class Real>
inline void weno5_right(Real& var, Real& der1, Real wn2, Real wn1, Real w0, Real wp1, Real wp2, double h) {
  Real fw[4], p[4], px[4];
  weno5_weights_nodf(fw, wn2, wn1, w0, wp1, wp2, 1e-10);
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
// weno_5th_ao_2gauss (no DF), reconstruction.cpp:2393-2551: values (and,
// with order 1, first derivatives) at the two Gauss points -+1/(2 sqrt 3)
template <class Real>
inline void weno5_2gauss(Real& g1, Real& g1x, Real& g2, Real& g2x, Real wn2, Real wn1, Real w0,
                         Real wp1, Real wp2, double h, int order) {
  Real fw[4], p[4], px[4];
  weno5_weights_nodf(fw, wn2, wn1, w0, wp1, wp2, 1e-6);
  g1 = 0;
  g1x = 0;
  g2 = 0;
  g2x = 0;
  const double sqrt3 = 1.732050807568877294;
  p[0] = w0 - (sqrt3 * w0) / 4 + (4 * wn1 - wn2) / (4 * sqrt3);
  p[1] = w0 + (wn1 - wp1) / (4 * sqrt3);
  p[2] = (1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wp1 + wp2));
  p[3] = (4314 * w0 + (4 + 500 * sqrt3) * wn1 - wn2 - 70 * sqrt3 * wn2 + 4 * wp1 - 500 * sqrt3 * wp1 - wp2 +
          70 * sqrt3 * wp2) /
         4320;
  for (int k = 0; k < 4; k++) g1 += fw[k] * p[k];
  if (order > 0) {
    px[0] = -((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wn1 + (-3 + sqrt3) * wn2) / (6 * h);
    px[1] = -(-2 * sqrt3 * w0 + (3 + sqrt3) * wn1 + (-3 + sqrt3) * wp1) / (6 * h);
    px[2] = -((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wp1 + (3 + sqrt3) * wp2) / (6 * h);
    px[3] = (48 * sqrt3 * w0 - 72 * wn1 - 26 * sqrt3 * wn1 + 9 * wn2 + 2 * sqrt3 * wn2 + 72 * wp1 - 26 * sqrt3 * wp1 -
             9 * wp2 + 2 * sqrt3 * wp2) /
            (108 * h);
    for (int k = 0; k < 4; k++) g1x += fw[k] * px[k];
  }
  p[0] = (1.0 / 12.0) * (3 * (4 + sqrt3) * w0 + sqrt3 * (-4 * wn1 + wn2));
  p[1] = w0 + (-wn1 + wp1) / (4 * sqrt3);
  p[2] = w0 - (sqrt3 * w0) / 4 + (4 * wp1 - wp2) / (4 * sqrt3);
  p[3] = (4314 * w0 + (4 - 500 * sqrt3) * wn1 - wn2 + 70 * sqrt3 * wn2 + 4 * wp1 + 500 * sqrt3 * wp1 - wp2 -
          70 * sqrt3 * wp2) /
         4320;
  for (int k = 0; k < 4; k++) g2 += fw[k] * p[k];
  if (order > 0) {
    px[0] = ((9 + sqrt3) * w0 - 2 * (6 + sqrt3) * wn1 + (3 + sqrt3) * wn2) / (6 * h);
    px[1] = (-2 * sqrt3 * w0 + (-3 + sqrt3) * wn1 + (3 + sqrt3) * wp1) / (6 * h);
    px[2] = ((-9 + sqrt3) * w0 - 2 * (-6 + sqrt3) * wp1 + (-3 + sqrt3) * wp2) / (6 * h);
    px[3] = -((48 * sqrt3 * w0 + 72 * wn1 - 26 * sqrt3 * wn1 - 9 * wn2 + 2 * sqrt3 * wn2 - 72 * wp1 -
               26 * sqrt3 * wp1 + 9 * wp2 + 2 * sqrt3 * wp2) /
              (108 * h));
    for (int k = 0; k < 4; k++) g2x += fw[k] * px[k];
  }
}

// WENO5_AO (3D, face values), reconstruction.cpp:3803-3923: w[5][5] the
// stencil in the face frame; lo = value + normal derivative at the cell's
// left face, ro at its right face (convar[5], der1x[5] each)
template <class Real>
inline void weno5_face(Real* lo, Real* ro, const Real (*w)[5], double h, int recon_var, const Gas3& g) {
  Real base_left[5], base_right[5], pn1[5], p0[5], pp1[5];
  convar_to_primvar(pn1, w[1], g);
  convar_to_primvar(p0, w[2], g);
  convar_to_primvar(pp1, w[3], g);
  for (int i = 0; i < 5; i++) {
    base_left[i] = 0.5 * (pn1[i] + p0[i]);
    base_right[i] = 0.5 * (pp1[i] + p0[i]);
  }
  for (int side = 0; side < 2; side++) {
    const Real* base = side == 0 ? base_left : base_right;
    Real r[5][5], var[5], der1[5];
    if (recon_var == 0) {
      for (int s = 0; s < 5; s++)
        for (int i = 0; i < 5; i++) r[s][i] = w[s][i];
    } else {
      Real sm[25];
      char_base(sm, base, g);
      for (int s = 0; s < 5; s++) to_char(r[s], sm, w[s]);
    }
    for (int i = 0; i < 5; i++) {
      if (side == 0) weno5_left(var[i], der1[i], r[0][i], r[1][i], r[2][i], r[3][i], r[4][i], h);
      else weno5_right(var[i], der1[i], r[0][i], r[1][i], r[2][i], r[3][i], r[4][i], h);
    }
    Real* out = side == 0 ? lo : ro;
    if (recon_var == 0) {
      for (int i = 0; i < 5; i++) {
        out[i] = var[i];
        out[5 + i] = der1[i];
      }
    } else {
      from_char(out, base, var, g);
      from_char(out + 5, base, der1, g);
    }
  }
  const bool rl = reduce_order(ro, g);
  const bool rr = reduce_order(lo, g);
  if (rl || rr) {
    for (int m = 0; m < 5; m++) {
      lo[m] = w[2][m];
      ro[m] = w[2][m];
      lo[5 + m] = 0.0;
      ro[5 + m] = 0.0;
    }
  }
}

// One side of WENO5_AO_for_line_value (reconstruction.cpp:3955-4145) or
// WENO5_AO_for_point_value (:4163-4399): the stencil s[5] of NC-component
// records (convar, then NC/5 - 1 derivatives), the first field reconstructed
// with order 1 (values + the new tangential derivative), the others with
// order 0.  out[2][NC + 5]: per Gauss point convar, the carried derivatives,
// then the new derivative.
template <int NC, class Real>
inline void weno5_tangential_side(Real (*out)[NC + 5], const Real (*s)[NC], double h, int recon_var,
                                  const Gas3& g) {
  constexpr int NF = NC / 5;
  Real base[5];
  convar_to_primvar(base, s[2], g);
  Real sm[25];
  if (recon_var != 0) char_base(sm, base, g);
  for (int fld = 0; fld < NF; fld++) {
    Real r[5][5];
    for (int q = 0; q < 5; q++)
      for (int i = 0; i < 5; i++) {
        if (recon_var == 0) r[q][i] = s[q][5 * fld + i];
      }
    if (recon_var != 0) {
      for (int q = 0; q < 5; q++) {
        Real c[5];
        for (int i = 0; i < 5; i++) c[i] = s[q][5 * fld + i];
        to_char(r[q], sm, c);
      }
    }
    Real var[2][5], der[2][5];
    for (int i = 0; i < 5; i++)
      weno5_2gauss(var[0][i], der[0][i], var[1][i], der[1][i], r[0][i], r[1][i], r[2][i], r[3][i], r[4][i], h,
                   fld == 0 ? 1 : 0);
    for (int q = 0; q < 2; q++) {
      if (recon_var == 0) {
        for (int i = 0; i < 5; i++) {
          out[q][5 * fld + i] = var[q][i];
          if (fld == 0) out[q][NC + i] = der[q][i];
        }
      } else {
        from_char(out[q] + 5 * fld, base, var[q], g);
        if (fld == 0) from_char(out[q] + NC, base, der[q], g);
      }
    }
  }
}

// ------------------------------------------------------------ GKS3D
// MMDF3d_left / MMDF3d_right / MMDF3d: basic_function.cpp:189-393
template <class Real>
struct MM3 {
  Real uh[7], uw[7], vw[6], ww[6], xi2, xi4;  // uh: u+ (left) or u- (right)
};
template <class Real>
inline Real alpha_fn(Real lambda, Real u) { return erfc(sqrt(lambda) * u); }
template <class Real>
inline Real beta_fn(Real lambda, Real u) {
  const double pi = 3.14159265358979323846;
  return exp(-lambda * u * u) / sqrt(pi * lambda);
}
template <class Real>
inline void mmdf3_side(MM3<Real>& m, int side, Real u, Real v, Real w, Real lambda, const Gas3& g) {
  m.uw[0] = 1;
  m.uw[1] = u;
  Real ol = 1.0 / lambda;
  if (side == 0) {
    m.uh[0] = 0.5 * alpha_fn(lambda, -u);
    m.uh[1] = u * m.uh[0] + 0.5 * beta_fn(lambda, u);
  } else {
    m.uh[0] = 0.5 * alpha_fn(lambda, u);
    m.uh[1] = u * m.uh[0] - 0.5 * beta_fn(lambda, u);
  }
  for (int i = 2; i <= 6; i++) m.uw[i] = u * m.uw[i - 1] + 0.5 * (i - 1) * ol * m.uw[i - 2];
  for (int i = 2; i <= 6; i++) m.uh[i] = u * m.uh[i - 1] + 0.5 * (i - 1) * ol * m.uh[i - 2];
  m.vw[0] = 1;
  m.vw[1] = v;
  m.ww[0] = 1;
  m.ww[1] = w;
  for (int i = 2; i <= 5; i++) {
    m.vw[i] = v * m.vw[i - 1] + 0.5 * (i - 1) * ol * m.vw[i - 2];
    m.ww[i] = w * m.ww[i - 1] + 0.5 * (i - 1) * ol * m.ww[i - 2];
  }
  m.xi2 = 0.5 * g.K * ol;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) * ol * ol;
}
// whole-space moments of MMDF3d (divisions by lambda, as written there)
template <class Real>
inline void mmdf3_whole(MM3<Real>& m, Real u, Real v, Real w, Real lambda, const Gas3& g) {
  m.uw[0] = 1;
  m.uw[1] = u;
  for (int i = 2; i <= 6; i++) m.uw[i] = u * m.uw[i - 1] + 0.5 * (i - 1) / lambda * m.uw[i - 2];
  m.vw[0] = 1;
  m.vw[1] = v;
  m.ww[0] = 1;
  m.ww[1] = w;
  for (int i = 2; i <= 5; i++) {
    m.vw[i] = v * m.vw[i - 1] + 0.5 * (i - 1) / lambda * m.vw[i - 2];
    m.ww[i] = w * m.ww[i - 1] + 0.5 * (i - 1) / lambda * m.ww[i - 2];
  }
  m.xi2 = 0.5 * g.K / lambda;
  m.xi4 = 0.25 * (g.K * g.K + 2 * g.K) / (lambda * lambda);
  for (int i = 0; i < 7; i++) m.uh[i] = 0.0;
}
// upvwxi / unvwxi (half) and uvwxi (whole) table entries
template <class Real>
struct TabH {
  const MM3<Real>& m;
  Real operator()(int i, int j, int k, int x) const {
    return x == 0 ? m.uh[i] * m.vw[j] * m.ww[k]
                  : (x == 1 ? m.uh[i] * m.vw[j] * m.ww[k] * m.xi2 : m.uh[i] * m.vw[j] * m.ww[k] * m.xi4);
  }
};
template <class Real>
struct TabW3 {
  const MM3<Real>& m;
  Real operator()(int i, int j, int k, int x) const {
    return x == 0 ? m.uw[i] * m.vw[j] * m.ww[k]
                  : (x == 1 ? m.uw[i] * m.vw[j] * m.ww[k] * m.xi2 : m.uw[i] * m.vw[j] * m.ww[k] * m.xi4);
  }
};
// GL_address / GR_address / G_address (3D): basic_function.h:225-567
template <class T, class Real>
inline void contract3(int u, int v, int w, int x, Real* psi, const Real* a, const T& m) {
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
template <class Real>
inline void micro_slope3(Real* a, const Real* der, const Real* prim, const Gas3& g) {
  Real density = prim[0], u = prim[1], v = prim[2], w = prim[3], lambda = prim[4];
  Real overden = 1.0 / density;
  Real overlambda = 1.0 / lambda;
  Real A = 2 * der[4] * overden - (u * u + v * v + w * w + 0.5 * (g.K + 3) * overlambda) * der[0] * overden;
  Real B = (der[1] - u * der[0]) * overden;
  Real C = (der[2] - v * der[0]) * overden;
  Real D = (der[3] - w * der[0]) * overden;
  a[4] = 4 * lambda * lambda / (g.K + 3) * (A - 2 * u * B - 2 * v * C - 2 * w * D);
  a[3] = 2 * lambda * D - w * a[4];
  a[2] = 2 * lambda * C - v * a[4];
  a[1] = 2 * lambda * B - u * a[4];
  a[0] = der[0] * overden - u * a[1] - v * a[2] - w * a[3] - 0.5 * a[4] * (u * u + v * v + w * w + 0.5 * (g.K + 3) * overlambda);
}
// Get_Tau_NS / Get_Tau: basic_function.cpp:638-708
template <class Real>
inline Real tau_ns3(Real rho0, Real l0, const Gas3& g) {
  if (g.tau_type == 0) return 0.0;
  if (g.mu > 0.0) return 2.0 * g.mu * l0 / rho0;
  if (g.nu > 0.0) return 2 * g.nu * l0;
  return 0.0;
}
template <class Real>
inline Real tau_num3(Real rl, Real rr, Real r0, Real ll, Real lr, Real l0, double dt, const Gas3& g) {
  if (g.tau_type == 0) {
    if (g.c1 <= 0 && g.c2 <= 0) return 0.0;
    Real C = g.c2 * fabs(rl / ll - rr / lr) / fabs(rl / ll + rr / lr);
    return g.c1 * dt + dt * C;
  }
  Real tau_n = g.c2 * fabs(rl / ll - rr / lr) / fabs(rl / ll + rr / lr) * dt;
  if (tau_n != tau_n) tau_n = 0.0;
  if ((g.mu > 0.0 && g.nu > 0.0) || (g.mu < 0.0 && g.nu < 0.0)) return 0.0;
  if (g.mu > 0.0) return tau_n + 2.0 * g.mu * l0 / r0;
  if (g.nu > 0.0) return tau_n + 2 * g.nu * l0;
  return 0.0;
}
template <class Real>
inline void time_coe3(Real* t, Real tau, Real tn, double dt) {
  Real eta = exp(-dt / tn);
  t[0] = tn * (1 - eta);
  t[1] = tn * (eta * (dt + tn) - tn) + tau * tn * (eta - 1);
  t[2] = tau * tn * (eta - 1);
  t[3] = tn * eta + dt - tn;
  t[4] = tn * (tn - eta * (dt + tn) - tau * (eta - 1)) - dt * tau;
  t[5] = 0.5 * dt * dt - tau * tn * (eta - 1) - tau * dt;
}

// GKS3D (gks2nd_3d, S2O4): one Gauss point.  L, R: convar[5], der1x[5],
// der1y[5], der1z[5] in the face frame.  center (optional): the collision
// state and its three derivatives (20 values).
template <class Real>
inline void gks3d_point(Real* f, Real* derf, Real* center, const Real* L, const Real* R, double dt,
                        const Gas3& g) {
  Real pl[5], pr[5], p0[5];
  convar_to_ulambda(pl, L, g);
  convar_to_ulambda(pr, R, g);
  MM3<Real> ml, mr;
  mmdf3_side(ml, 0, pl[1], pl[2], pl[3], pl[4], g);
  mmdf3_side(mr, 1, pr[1], pr[2], pr[3], pr[4], g);
  const TabH<Real> tl{ml}, tr{mr};
  const TabW3<Real> wl{ml}, wr{mr};
  // Collision (basic_function.h:203-212)
  Real c0[5];
  c0[0] = pl[0] * ml.uh[0] + pr[0] * mr.uh[0];
  c0[1] = pl[0] * ml.uh[1] + pr[0] * mr.uh[1];
  c0[2] = pl[0] * ml.vw[1] * ml.uh[0] + pr[0] * mr.vw[1] * mr.uh[0];
  c0[3] = pl[0] * ml.ww[1] * ml.uh[0] + pr[0] * mr.ww[1] * mr.uh[0];
  c0[4] = 0.5 * pl[0] * (ml.uh[2] + ml.uh[0] * ml.vw[2] + ml.uh[0] * ml.ww[2] + ml.uh[0] * ml.xi2) +
          0.5 * pr[0] * (mr.uh[2] + mr.uh[0] * mr.vw[2] + mr.uh[0] * mr.ww[2] + mr.uh[0] * mr.xi2);
  Real alx[5], arx[5], aly[5], ary[5], alz[5], arz[5];
  micro_slope3(alx, L + 5, pl, g);
  micro_slope3(arx, R + 5, pr, g);
  micro_slope3(aly, L + 10, pl, g);
  micro_slope3(ary, R + 10, pr, g);
  micro_slope3(alz, L + 15, pl, g);
  micro_slope3(arz, R + 15, pr, g);
  Real c0x[5], c0y[5], c0z[5];
  {
    Real al0[5], ar0[5];
    contract3(0, 0, 0, 0, al0, alx, tl);
    contract3(0, 0, 0, 0, ar0, arx, tr);
    for (int v = 0; v < 5; v++) c0x[v] = pl[0] * al0[v] + pr[0] * ar0[v];
    contract3(0, 0, 0, 0, al0, aly, tl);
    contract3(0, 0, 0, 0, ar0, ary, tr);
    for (int v = 0; v < 5; v++) c0y[v] = pl[0] * al0[v] + pr[0] * ar0[v];
    contract3(0, 0, 0, 0, al0, alz, tl);
    contract3(0, 0, 0, 0, ar0, arz, tr);
    for (int v = 0; v < 5; v++) c0z[v] = pl[0] * al0[v] + pr[0] * ar0[v];
  }
  if (center) {
    for (int v = 0; v < 5; v++) {
      center[v] = c0[v];
      center[5 + v] = c0x[v];
      center[10 + v] = c0y[v];
      center[15 + v] = c0z[v];
    }
  }
  convar_to_ulambda(p0, c0, g);
  const Real tau = tau_ns3(p0[0], p0[4], g);
  Real t[6], F0[5], F1[5];
  time_coe3(t, tau, tau_num3(pl[0], pr[0], p0[0], pl[4], pr[4], p0[4], dt, g), dt);
  const Real unit[5] = {1, 0.0, 0.0, 0.0, 0.0};
  Real glu[5], gru[5];
  contract3(1, 0, 0, 0, glu, unit, tl);
  contract3(1, 0, 0, 0, gru, unit, tr);
  for (int i = 0; i < 5; i++) F0[i] = pl[0] * t[0] * glu[i] + pr[0] * t[0] * gru[i];
  MM3<Real> m0;
  mmdf3_whole(m0, p0[1], p0[2], p0[3], p0[4], g);
  const TabW3<Real> w0{m0};
  Real g0u[5];
  contract3(1, 0, 0, 0, g0u, unit, w0);
  for (int i = 0; i < 5; i++) F0[i] = F0[i] + p0[0] * t[3] * g0u[i];
  Real alxuul[5], arxuur[5], alyuvl[5], aryuvr[5], alzuwl[5], arzuwr[5];
  contract3(2, 0, 0, 0, alxuul, alx, tl);
  contract3(2, 0, 0, 0, arxuur, arx, tr);
  contract3(1, 1, 0, 0, alyuvl, aly, tl);
  contract3(1, 1, 0, 0, aryuvr, ary, tr);
  contract3(1, 0, 1, 0, alzuwl, alz, tl);
  contract3(1, 0, 1, 0, arzuwr, arz, tr);
  for (int i = 0; i < 5; i++)
    F0[i] = F0[i] + t[1] * (pl[0] * (alxuul[i] + alyuvl[i] + alzuwl[i]) + pr[0] * (arxuur[i] + aryuvr[i] + arzuwr[i]));
  Real alxu[5], arxu[5], alyv[5], aryv[5], alzw[5], arzw[5];
  contract3(1, 0, 0, 0, alxu, alx, wl);
  contract3(1, 0, 0, 0, arxu, arx, wr);
  contract3(0, 1, 0, 0, alyv, aly, wl);
  contract3(0, 1, 0, 0, aryv, ary, wr);
  contract3(0, 0, 1, 0, alzw, alz, wl);
  contract3(0, 0, 1, 0, arzw, arz, wr);
  Real der_AL[5], der_AR[5], Al[5], Ar[5], Alul[5], Arur[5];
  for (int i = 0; i < 5; i++) {
    der_AL[i] = -pl[0] * (alxu[i] + alyv[i] + alzw[i]);
    der_AR[i] = -pr[0] * (arxu[i] + aryv[i] + arzw[i]);
  }
  micro_slope3(Al, der_AL, pl, g);
  micro_slope3(Ar, der_AR, pr, g);
  contract3(1, 0, 0, 0, Alul, Al, tl);
  contract3(1, 0, 0, 0, Arur, Ar, tr);
  for (int i = 0; i < 5; i++) F0[i] = F0[i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);
  Real a0x[5], a0xu[5], a0xuu[5], a0y[5], a0yv[5], a0yuv[5], a0z[5], a0zw[5], a0zuw[5];
  micro_slope3(a0x, c0x, p0, g);
  contract3(1, 0, 0, 0, a0xu, a0x, w0);
  contract3(2, 0, 0, 0, a0xuu, a0x, w0);
  micro_slope3(a0y, c0y, p0, g);
  contract3(0, 1, 0, 0, a0yv, a0y, w0);
  contract3(1, 1, 0, 0, a0yuv, a0y, w0);
  micro_slope3(a0z, c0z, p0, g);
  contract3(0, 0, 1, 0, a0zw, a0z, w0);
  contract3(1, 0, 1, 0, a0zuw, a0z, w0);
  for (int i = 0; i < 5; i++) F0[i] = F0[i] + p0[0] * t[4] * (a0xuu[i] + a0yuv[i] + a0zuw[i]);
  Real derA0[5], A0[5], A0u[5];
  for (int i = 0; i < 5; i++) derA0[i] = -p0[0] * (a0xu[i] + a0yv[i] + a0zw[i]);
  micro_slope3(A0, derA0, p0, g);
  contract3(1, 0, 0, 0, A0u, A0, w0);
  for (int i = 0; i < 5; i++) F0[i] = F0[i] + p0[0] * t[5] * (A0u[i]);
  const double dt2 = 0.5 * dt;
  time_coe3(t, tau, tau_num3(pl[0], pr[0], p0[0], pl[4], pr[4], p0[4], dt2, g), dt2);
  for (int i = 0; i < 5; i++) {
    F1[i] = t[0] * (pl[0] * glu[i] + pr[0] * gru[i]);
    F1[i] = F1[i] + t[1] * (pl[0] * (alxuul[i] + alyuvl[i] + alzuwl[i]) + pr[0] * (arxuur[i] + aryuvr[i] + arzuwr[i]));
    F1[i] = F1[i] + t[2] * (pl[0] * Alul[i] + pr[0] * Arur[i]);
    F1[i] = F1[i] + p0[0] * t[3] * g0u[i];
    F1[i] = F1[i] + p0[0] * t[4] * (a0xuu[i] + a0yuv[i] + a0zuw[i]);
    F1[i] = F1[i] + p0[0] * t[5] * (A0u[i]);
  }
  for (int i = 0; i < 5; i++) {
    f[i] = (4.0 * F1[i] - F0[i]);
    derf[i] = 4.0 * (F0[i] - 2.0 * F1[i]);
  }
}

// ------------------------------------------------------------ the solver
struct Params3 {
  int NX, NY, NZ, G;
  int recon_var;
  double dx, dy, dz, cfl;
  Gas3 gas;
};

struct Solver3d {
  Params3 P;
  int nx, ny, nz;
  long long nc, nf;
  std::vector<double> U, Uold, prim;  // [5][nc]
  // per direction: face values [nf][2 side][10], line values [nf][2 line][2 side][15],
  // points [nf][4][2 side][20], stage-0 share [nf][4][5], stage combination x [nf][4][5]
  std::vector<double> fv[3], lv[3], pv[3], pc[3], xc[3];
  double t = 0.0;
  int step = 0;
  double tmax = 0.0, rhomin = 0.0;  // from the last primvar
  bool blowup = false;

  explicit Solver3d(const Params3& p) : P(p) {
    nx = P.NX + 2 * P.G;
    ny = P.NY + 2 * P.G;
    nz = P.NZ + 2 * P.G;
    nc = (long long)nx * ny * nz;
    nf = (long long)(nx + 1) * (ny + 1) * (nz + 1);
    U.assign(5 * nc, 0.0);
    Uold.assign(5 * nc, 0.0);
    prim.assign(5 * nc, 0.0);
    for (int d = 0; d < 3; d++) {
      fv[d].assign(nf * 20, 0.0);
      lv[d].assign(nf * 60, 0.0);
      pv[d].assign(nf * 160, 0.0);
      pc[d].assign(nf * 20, 0.0);
      xc[d].assign(nf * 20, 0.0);
    }
  }
  long long cell(int i, int j, int k) const { return ((long long)i * ny + j) * nz + k; }
  long long face(int i, int j, int k) const { return ((long long)i * (ny + 1) + j) * (nz + 1) + k; }
  void get_cell(double* c, long long idx) const {
    for (int v = 0; v < 5; v++) c[v] = U[v * nc + idx];
  }

  // free: boundary_condition.cpp:445-570; periodic: from i +- NX
  void boundary(int side, int kind) {
    const int G = P.G;
    for (int a = 0; a < (side < 2 ? ny : nx); a++)
      for (int b = 0; b < (side < 4 ? nz : ny); b++)
        for (int gl = 0; gl < G; gl++) {
          int i, j, k, si, sj, sk;
          if (side < 2) {
            j = a;
            k = b;
            i = side == 0 ? G - 1 - gl : nx - G + gl;
            sj = j;
            sk = k;
            si = kind == 0 ? (side == 0 ? G : nx - G - 1) : (side == 0 ? i + P.NX : i - P.NX);
          } else if (side < 4) {
            i = a;
            k = b;
            j = side == 2 ? G - 1 - gl : ny - G + gl;
            si = i;
            sk = k;
            sj = kind == 0 ? (side == 2 ? G : ny - G - 1) : (side == 2 ? j + P.NY : j - P.NY);
          } else {
            i = a;
            j = b;
            k = side == 4 ? G - 1 - gl : nz - G + gl;
            si = i;
            sj = j;
            sk = kind == 0 ? (side == 4 ? G : nz - G - 1) : (side == 4 ? k + P.NZ : k - P.NZ);
          }
          const long long dst = cell(i, j, k), src = cell(si, sj, sk);
          for (int v = 0; v < 5; v++) U[v * nc + dst] = U[v * nc + src];
        }
  }
  // Convar_to_primvar (3D): basic_function.cpp:775-804, interior cells;
  // plus the CFL inputs (maximum speed, minimum density)
  void primvar() {
    tmax = 0.0;
    rhomin = 1e300;
    for (int i = P.G; i < nx - P.G; i++)
      for (int j = P.G; j < ny - P.G; j++)
        for (int k = P.G; k < nz - P.G; k++) {
          const long long c = cell(i, j, k);
          double u[5], p[5];
          get_cell(u, c);
          convar_to_primvar(p, u, P.gas);
          for (int v = 0; v < 5; v++) prim[v * nc + c] = p[v];
          if (std::isnan(p[0]) || std::isnan(p[4]) || p[0] < 0 || p[4] < 0) blowup = true;
          double overden = 1.0 / p[0];
          double tmp = sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + sqrt(P.gas.gamma * p[4] * overden);
          if (tmp > tmax) tmax = tmp;
          if (p[0] < rhomin) rhomin = p[0];
        }
  }
  // Get_CFL / Dtx_3d (time_advance.cpp:793-848) from the maximum speed and
  // minimum density of the last primvar
  double cfl_dt(double tstop) const {
    double dx = P.dx;
    if (P.dy < dx) dx = P.dy;
    if (P.dz < dx) dx = P.dz;
    double dt = dx;
    if (tmax > P.cfl * dx / dt) dt = P.cfl * dx / tmax;
    if (dt > 0.33 * P.cfl * dx * dx * rhomin / P.gas.mu && P.gas.tau_type != 0 && P.gas.mu > 0)
      dt = 0.33 * P.cfl * dx * dx * rhomin / P.gas.mu;
    if (t + dt > tstop) dt = tstop - t + 1e-16;
    return dt;
  }
  // the CFL inputs from given primitive variables (the initial condition's)
  void cfl_from_prim(const double* pr) {
    tmax = 0.0;
    rhomin = 1e300;
    for (int i = P.G; i < nx - P.G; i++)
      for (int j = P.G; j < ny - P.G; j++)
        for (int k = P.G; k < nz - P.G; k++) {
          const long long c = cell(i, j, k);
          double p[5];
          for (int v = 0; v < 5; v++) p[v] = pr[v * nc + c];
          double overden = 1.0 / p[0];
          double tmp = sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + sqrt(P.gas.gamma * p[4] * overden);
          if (tmp > tmax) tmax = tmp;
          if (p[0] < rhomin) rhomin = p[0];
        }
  }
  void copy_old() {
    for (int i = P.G; i < nx - P.G; i++)
      for (int j = P.G; j < ny - P.G; j++)
        for (int k = P.G; k < nz - P.G; k++) {
          const long long c = cell(i, j, k);
          for (int v = 0; v < 5; v++) Uold[v * nc + c] = U[v * nc + c];
        }
  }

  // WENO5_AO_splitting_3d over the cells of Reconstruction_within_cell
  // (reconstruction.cpp:3698-3716, 3774-3801) for direction d
  void face_values(int d) {
    const int G = P.G;
    for (int i = G - 2; i < nx - G + 2; i++)
      for (int j = G - 2; j < ny - G + 2; j++)
        for (int k = G - 2; k < nz - G + 2; k++) {
          const int idx = d == 0 ? i : (d == 1 ? j : k);
          const int n = d == 0 ? nx : (d == 1 ? ny : nz);
          if (!(idx > G - 2 && idx < n - G + 1)) continue;
          double w[5][5];
          for (int s = 0; s < 5; s++) {
            const int o = s - 2;
            double u[5];
            get_cell(u, cell(i + (d == 0) * o, j + (d == 1) * o, k + (d == 2) * o));
            if (d == 0)
              for (int v = 0; v < 5; v++) w[s][v] = u[v];
            else if (d == 1)
              ydir(w[s], u);
            else
              zdir(w[s], u);
          }
          double lo[10], ro[10];
          weno5_face(lo, ro, w, d == 0 ? P.dx : (d == 1 ? P.dy : P.dz), P.recon_var, P.gas);
          const long long fa = face(i, j, k);
          const long long fb = face(i + (d == 0), j + (d == 1), k + (d == 2));
          for (int m = 0; m < 10; m++) {
            fv[d][(fa * 2 + 1) * 10 + m] = lo[m];  // the cell's left face, its right-side state
            fv[d][(fb * 2 + 0) * 10 + m] = ro[m];
          }
        }
  }
  // WENO5_AO_of_line_value (reconstruction.cpp:3718-3734, 3925-3953)
  void line_values(int d) {
    const int G = P.G;
    for (int i = G - 2; i < nx - G + 2; i++)
      for (int j = G - 2; j < ny - G + 2; j++)
        for (int k = G - 2; k < nz - G + 2; k++) {
          int fi = i, fj = j, fk = k;
          long long st[5];
          double h;
          if (d == 0) {
            fi = i + 1;
            if (!(fj > G - 1 && fj < ny - G && fi > G - 1 && fi < nx - G + 1)) continue;
            for (int s = 0; s < 5; s++) st[s] = face(fi, fj - 2 + s, fk);
            h = P.dy;
          } else if (d == 1) {
            fj = j + 1;
            if (!(fi > G - 1 && fi < nx - G && fj > G - 1 && fj < ny - G + 1)) continue;
            for (int s = 0; s < 5; s++) st[s] = face(fi + 2 - s, fj, fk);
            h = P.dx;
          } else {
            fk = k + 1;
            if (!(fj > G - 1 && fj < ny - G && fk > G - 1 && fk < nz - G + 1)) continue;
            for (int s = 0; s < 5; s++) st[s] = face(fi, fj - 2 + s, fk);
            h = P.dy;
          }
          const long long f = face(fi, fj, fk);
          double res[2][2][15];  // [side][line][15]
          for (int side = 0; side < 2; side++) {
            double s10[5][10];
            for (int s = 0; s < 5; s++)
              for (int m = 0; m < 10; m++) s10[s][m] = fv[d][(st[s] * 2 + side) * 10 + m];
            weno5_tangential_side<10>(res[side], s10, h, P.recon_var, P.gas);
          }
          bool red = false;
          for (int l = 0; l < 2; l++)
            if (reduce_order(res[0][l], P.gas) || reduce_order(res[1][l], P.gas)) red = true;
          if (red) {
            for (int side = 0; side < 2; side++)
              for (int l = 0; l < 2; l++) {
                for (int m = 0; m < 10; m++) res[side][l][m] = fv[d][(f * 2 + side) * 10 + m];
                for (int m = 10; m < 15; m++) res[side][l][m] = 0.0;
              }
          }
          for (int l = 0; l < 2; l++)
            for (int side = 0; side < 2; side++)
              for (int m = 0; m < 15; m++) lv[d][((f * 2 + l) * 2 + side) * 15 + m] = res[side][l][m];
        }
  }
  // the faces whose fluxes are computed (Calculate_flux, flux_function.cpp:760-811)
  bool flux_face(int d, int i, int j, int k) const {
    const int G = P.G;
    return i >= G && i < P.NX + G + (d == 0) && j >= G && j < P.NY + G + (d == 1) && k >= G &&
           k < P.NZ + G + (d == 2);
  }
  // WENO5_AO_of_point_value (reconstruction.cpp:3736-3752, 4147-4161) at
  // the faces that get a flux
  void point_values(int d) {
    for (int i = 0; i <= nx; i++)
      for (int j = 0; j <= ny; j++)
        for (int k = 0; k <= nz; k++) {
          if (!flux_face(d, i, j, k)) continue;
          const long long f = face(i, j, k);
          long long st[5];
          double h;
          if (d == 2) {
            for (int s = 0; s < 5; s++) st[s] = face(i + 2 - s, j, k);
            h = P.dx;
          } else {
            for (int s = 0; s < 5; s++) st[s] = face(i, j, k - 2 + s);
            h = P.dz;
          }
          for (int l = 0; l < 2; l++) {
            double res[2][2][20];  // [side][point][20]
            for (int side = 0; side < 2; side++) {
              double s15[5][15];
              for (int s = 0; s < 5; s++)
                for (int m = 0; m < 15; m++) s15[s][m] = lv[d][((st[s] * 2 + l) * 2 + side) * 15 + m];
              weno5_tangential_side<15>(res[side], s15, h, P.recon_var, P.gas);
            }
            bool red = false;
            for (int q = 0; q < 2; q++)
              if (reduce_order(res[0][q], P.gas) || reduce_order(res[1][q], P.gas)) red = true;
            if (red) {
              for (int side = 0; side < 2; side++)
                for (int q = 0; q < 2; q++) {
                  for (int m = 0; m < 15; m++) res[side][q][m] = lv[d][((f * 2 + l) * 2 + side) * 15 + m];
                  for (int m = 15; m < 20; m++) res[side][q][m] = 0.0;
                }
            }
            for (int q = 0; q < 2; q++)
              for (int side = 0; side < 2; side++)
                for (int m = 0; m < 20; m++) pv[d][((f * 4 + 2 * l + q) * 2 + side) * 20 + m] = res[side][q][m];
          }
        }
  }
  // GKS3D per Gauss point and the S2O4 stage combination of Update_gauss
  // (time_advance.cpp:867-935): stage 0 x = 0.5 f + 1/8 derf (the stage-1
  // share 1.0 f + 1/6 derf kept); stage 1 x = share + (0 f + 1/3 derf)
  void fluxes(int d, int stage, double dt) {
    for (int i = 0; i <= nx; i++)
      for (int j = 0; j <= ny; j++)
        for (int k = 0; k <= nz; k++) {
          if (!flux_face(d, i, j, k)) continue;
          const long long f = face(i, j, k);
          for (int q = 0; q < 4; q++) {
            const double* L = &pv[d][((f * 4 + q) * 2 + 0) * 20];
            const double* R = &pv[d][((f * 4 + q) * 2 + 1) * 20];
            double fo[5], dfo[5];
            gks3d_point(fo, dfo, static_cast<double*>(nullptr), L, R, dt, P.gas);
            for (int v = 0; v < 5; v++) {
              double x = 0.0;
              if (stage == 0) {
                x = x + (0.5 * fo[v] + 1.0 / 8.0 * dfo[v]);
                pc[d][(f * 4 + q) * 5 + v] = 0.0 + (1.0 * fo[v] + 1.0 / 6.0 * dfo[v]);
              } else {
                x = pc[d][(f * 4 + q) * 5 + v] + (0.0 * fo[v] + 1.0 / 3.0 * dfo[v]);
              }
              xc[d][(f * 4 + q) * 5 + v] = x;
            }
          }
        }
  }
  // the cell update of Update_gauss (time_advance.cpp:937-998)
  void update() {
    const double odx = 1 / P.dx, ody = 1 / P.dy, odz = 1 / P.dz;
    for (int i = P.G; i < nx - P.G; i++)
      for (int j = P.G; j < ny - P.G; j++)
        for (int k = P.G; k < nz - P.G; k++) {
          const long long c = cell(i, j, k), fc = face(i, j, k);
          const long long fx = face(i + 1, j, k), fy = face(i, j + 1, k), fz = face(i, j, k + 1);
          double tot[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
          for (int q = 0; q < 4; q++) {
            const double w = 0.25;
            auto X = [&](int d, long long f, int v) { return xc[d][(f * 4 + q) * 5 + v]; };
            tot[0] += w * odz * (X(2, fc, 0) - X(2, fz, 0));
            tot[1] += w * odz * (-X(2, fc, 3) + X(2, fz, 3));
            tot[2] += w * odz * (X(2, fc, 2) - X(2, fz, 2));
            tot[3] += w * odz * (X(2, fc, 1) - X(2, fz, 1));
            tot[4] += w * odz * (X(2, fc, 4) - X(2, fz, 4));
            tot[0] += w * ody * (X(1, fc, 0) - X(1, fy, 0));
            tot[1] += w * ody * (-X(1, fc, 2) + X(1, fy, 2));
            tot[2] += w * ody * (X(1, fc, 1) - X(1, fy, 1));
            tot[3] += w * ody * (X(1, fc, 3) - X(1, fy, 3));
            tot[4] += w * ody * (X(1, fc, 4) - X(1, fy, 4));
            for (int v = 0; v < 5; v++) tot[v] += w * odx * (X(0, fc, v) - X(0, fx, v));
          }
          for (int v = 0; v < 5; v++) U[v * nc + c] = Uold[v * nc + c] + tot[v];
        }
  }
  // one stage of the driver loop (3Dproblem.cpp:135-155)
  void stage(int s, double dt, const int* bc) {
    for (int side = 0; side < 6; side++) boundary(side, bc[side]);
    primvar();
    for (int d = 0; d < 3; d++) {
      face_values(d);
      line_values(d);
      point_values(d);
      fluxes(d, s, dt);
    }
    update();
  }
  // the driver loop; returns 0, or 4 on blow-up (the step is not counted)
  int run(const int* bc, double tstop, int max_steps, double* dt_hist) {
    int n = 0;
    while (t < tstop && n < max_steps) {
      if (blowup) return 4;
      copy_old();
      const double dt = cfl_dt(tstop);
      stage(0, dt, bc);
      stage(1, dt, bc);
      if (blowup) return 4;
      if (dt_hist) dt_hist[n] = dt;
      t = t + dt;
      step++;
      n++;
    }
    return 0;
  }
};

}  // namespace orc3d
