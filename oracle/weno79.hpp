// weno79.hpp — CPU restatement of the reference's WENO7-AO and WENO9-AO
// reconstructions with the DF factor (2D normal and tangential leaves).
//
// TEST INFRASTRUCTURE ONLY (see gks_core.hpp).  The candidate polynomials and
// smoothness indicators are written as tables of (coefficient, stencil
// offset) terms summed left to right, each table in the reference's term
// order, so that the sums round exactly as the reference's expressions do:
// c*w added or subtracted term by term, negative coefficients by their
// magnitude (x - c*w == x + (-c)*w in IEEE arithmetic).
//
// Reference: reconstruction.cpp
//   WENO7: weno_7th_ao_with_df_left/right :750-974, WENO7_AO_with_df :1894-2028,
//          WENO7_AO_with_df_tangent :2840-2865, weno_7th_ao_with_df_tangential
//          :2867-3018, weno_7th_ao_with_df_2gauss (3 points) :3020-3174;
//   WENO9: weno_9th_ao_with_df_left/right :1072-1364, WENO9_AO_with_df
//          :2063-2211, WENO9_AO_with_df_tangent :3176-3207,
//          weno_9th_ao_with_df_tangential :3209-3360, weno_9th_ao_with_df_2gauss
//          (4 Lobatto points) :3362-3549, Polynomial_9th :3551-3609.
#pragma once
#include "gks_core.hpp"

namespace orc {

// one term c * w[o] of a stencil sum (o = offset from the centre cell)
struct WT {
  double c;
  int o;
};

// sum of n terms in table order; unit coefficients add the value itself
template <class R> inline R tsum(const WT* t, int n, const R* w) {
  auto term = [&](const WT& x) -> R { return x.c == 1.0 ? w[x.o] : (x.c == -1.0 ? -w[x.o] : x.c * w[x.o]); };
  R s = term(t[0]);
  for (int k = 1; k < n; k++) {
    if (t[k].c == 1.0) s = s + w[t[k].o];
    else if (t[k].c == -1.0) s = s - w[t[k].o];
    else s = s + t[k].c * w[t[k].o];
  }
  return s;
}
template <class R, size_t N> inline R tsum(const WT (&t)[N], const R* w) { return tsum(t, (int)N, w); }

// a table row of up to 9 terms
struct WRow {
  int n;
  WT t[9];
};
template <class R> inline R tsum(const WRow& r, const R* w) { return tsum(r.t, r.n, w); }

// ------------------------------------------------------ smoothness indicators
// three-point Jiang-Shu indicators of the 3rd-order sub-stencils
template <class R> inline void beta3_js(R* b, const R* w) {
  R q, q2;
  q = w[-2] - 2.0 * w[-1] + w[0];
  q2 = w[-2] - 4.0 * w[-1] + 3.0 * w[0];
  b[0] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = w[-1] - 2.0 * w[0] + w[1];
  q2 = w[-1] - w[1];
  b[1] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
  q = w[0] - 2.0 * w[1] + w[2];
  q2 = 3.0 * w[0] - 4.0 * w[1] + w[2];
  b[2] = 13.0 / 12.0 * (q * q) + 0.25 * (q2 * q2);
}

// the 7th-order stencil's indicator from its four 4-point sub-indicators
// (a1, a2, a3 Taylor coefficients; 13/3 and 781/20 are integer quotients
// 4 and 39 in the reference, e.g. reconstruction.cpp:783-784)
template <class R> inline R beta7_global(const R* w) {
  static const WT A1[4][4] = {{{-19, -3}, {87, -2}, {-177, -1}, {109, 0}},
                              {{11, -2}, {-63, -1}, {33, 0}, {19, 1}},
                              {{-19, -1}, {-33, 0}, {63, 1}, {-11, 2}},
                              {{-109, 0}, {177, 1}, {-87, 2}, {19, 3}}};
  static const WT A2_0[4] = {{-1, -3}, {4, -2}, {-5, -1}, {2, 0}};
  static const WT A2_12[3] = {{1, -1}, {-2, 0}, {1, 1}};
  static const WT A2_3[4] = {{2, 0}, {-5, 1}, {4, 2}, {-1, 3}};
  static const WT A3[4][4] = {{{-1, -3}, {3, -2}, {-3, -1}, {1, 0}},
                              {{-1, -2}, {3, -1}, {-3, 0}, {1, 1}},
                              {{-1, -1}, {3, 0}, {-3, 1}, {1, 2}},
                              {{-1, 0}, {3, 1}, {-3, 2}, {1, 3}}};
  R b[4];
  for (int k = 0; k < 4; k++) {
    R a1 = tsum(A1[k], w) / 60.0;
    R a2 = (k == 0 ? tsum(A2_0, w) : (k == 3 ? tsum(A2_3, w) : tsum(A2_12, w))) / 2.0;
    R a3 = tsum(A3[k], w) / 6.0;
    R t = a1 + a3 / 10.0;
    b[k] = t * t + 4.0 * a2 * a2 + 39.0 * a3 * a3;
  }
  return 1.0 / 20.0 * (b[0] + 9.0 * b[1] + 9.0 * b[2] + b[3]) + fabs(b[0] - b[1] - b[2] + b[3]);
}

// the 9th-order stencil's five 5-point sub-indicators (real constants here,
// reconstruction.cpp:1138-1139); bs[2] doubles as the 5th-order indicator
template <class R> inline void beta9_sub(R* bs, const R* w) {
  static const WRow A1[5] = {{5, {{-462, -1}, {336, -2}, {-146, -3}, {27, -4}, {245, 0}}},
                             {5, {{-192, -1}, {66, -2}, {-11, -3}, {110, 0}, {27, 1}}},
                             {4, {{-82, -1}, {11, -2}, {82, 1}, {-11, 2}}},
                             {5, {{-27, -1}, {-110, 0}, {192, 1}, {-66, 2}, {11, 3}}},
                             {5, {{-245, 0}, {462, 1}, {-336, 2}, {146, 3}, {-27, 4}}}};
  static const WRow A2[5] = {{5, {{-240, -1}, {262, -2}, {-128, -3}, {25, -4}, {81, 0}}},
                             {5, {{10, -1}, {12, -2}, {-3, -3}, {-44, 0}, {25, 1}}},
                             {5, {{40, -1}, {-3, -2}, {-74, 0}, {40, 1}, {-3, 2}}},
                             {5, {{25, -1}, {-44, 0}, {10, 1}, {12, 2}, {-3, 3}}},
                             {5, {{81, 0}, {-240, 1}, {262, 2}, {-128, 3}, {25, 4}}}};
  static const WRow A3[5] = {{5, {{-18, -1}, {24, -2}, {-14, -3}, {3, -4}, {5, 0}}},
                             {5, {{12, -1}, {-6, -2}, {1, -3}, {-10, 0}, {3, 1}}},
                             {4, {{2, -1}, {-1, -2}, {-2, 1}, {1, 2}}},
                             {5, {{-3, -1}, {10, 0}, {-12, 1}, {6, 2}, {-1, 3}}},
                             {5, {{-5, 0}, {18, 1}, {-24, 2}, {14, 3}, {-3, 4}}}};
  static const WRow A4[5] = {{5, {{-4, -1}, {6, -2}, {-4, -3}, {1, -4}, {1, 0}}},
                             {5, {{6, -1}, {-4, -2}, {1, -3}, {-4, 0}, {1, 1}}},
                             {5, {{-4, -1}, {1, -2}, {6, 0}, {-4, 1}, {1, 2}}},
                             {5, {{1, -1}, {-4, 0}, {6, 1}, {-4, 2}, {1, 3}}},
                             {5, {{1, 0}, {-4, 1}, {6, 2}, {-4, 3}, {1, 4}}}};
  for (int k = 0; k < 5; k++) {
    R a1 = tsum(A1[k], w) / 120.0, a2 = tsum(A2[k], w) / 56.0, a3 = tsum(A3[k], w) / 12.0,
      a4 = tsum(A4[k], w) / 24.0;
    R t1 = a1 + a3 / 10.0;
    R t2 = a2 + 123.0 / 455.0 * a4;
    bs[k] = t1 * t1 + 13.0 / 3.0 * t2 * t2 + 781.0 / 20.0 * a3 * a3 + 1421461.0 / 2275.0 * a4 * a4;
  }
}

// AO weights of the P(7,5,3,3,3) / P(9,7,5,3,3,3) hierarchies, combined into
// the final weights fw[k] = w_k - w_top d_k / d_top, fw[top] = w_top / d_top.
// ORDER 7: NS = 5 sub-stencils; ORDER 9: NS = 6.
template <int ORDER, class R> inline void ao_weights(R* fw, const R* w) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  const double h85 = 0.85;
  double d[NS];
  R beta[NS];
  beta3_js(beta, w);
  if constexpr (ORDER == 7) {
    d[0] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) / 2.0;
    d[1] = (1.0 - h85) * (1.0 - h85) * h85;
    d[2] = d[0];
    d[3] = (1.0 - h85) * h85;
    d[4] = h85;
    beta[3] = 1.0 / 6.0 * (beta[0] + 4.0 * beta[1] + beta[2]) + fabs(beta[0] - beta[2]);
    beta[4] = beta7_global(w);
  } else {
    d[0] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) * (1.0 - h85) / 2.0;
    d[1] = (1.0 - h85) * (1.0 - h85) * (1.0 - h85) * h85;
    d[2] = d[0];
    d[3] = (1.0 - h85) * (1.0 - h85) * h85;
    d[4] = (1.0 - h85) * h85;
    d[NS - 1] = h85;
    R bs[5];
    beta9_sub(bs, w);
    beta[3] = bs[2];
    beta[4] = beta7_global(w);
    beta[NS - 1] = 1.0 / 70.0 * (bs[0] + 16.0 * bs[1] + 36.0 * bs[2] + 16.0 * bs[3] + bs[4]) +
                   fabs(bs[0] - bs[1] - bs[3] + bs[4]);
  }
  const R top = beta[NS - 1];
  R tsumb = fabs(top - beta[0]);
  for (int k = 1; k < NS - 1; k++) tsumb = tsumb + fabs(top - beta[k]);
  const R tau = (ORDER == 7 ? 0.25 : 0.2) * tsumb;
  R alpha[NS], sum = 0.0;
  for (int k = 0; k < NS; k++) {
    R gd = tau / (beta[k] + 1e-6);
    alpha[k] = d[k] * (1.0 + gd * gd);
    sum += alpha[k];
  }
  R ww[NS];
  for (int k = 0; k < NS; k++) ww[k] = alpha[k] / sum;
  fw[NS - 1] = ww[NS - 1] / d[NS - 1];
  for (int k = 0; k < NS - 1; k++) fw[k] = ww[k] - ww[NS - 1] / d[NS - 1] * d[k];
}

// DF factors: sums of the alphas of the cells each sub-stencil spans, at the
// sub-stencil parity (reconstruction.cpp:1897-1908, 2066-2078)
template <int ORDER, class R> inline void df_factors(R* df, const R* a /* centred */) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  R s[NS];
  s[0] = a[-2] + a[0];
  s[1] = a[-1] + a[1];
  s[2] = a[0] + a[2];
  s[3] = a[-2] + a[0] + a[2];
  s[4] = a[-3] + a[-1] + a[1] + a[3];
  if constexpr (ORDER == 9) s[5] = a[-4] + a[-2] + a[0] + a[2] + a[4];
  for (int i = 0; i < NS; i++) df[i] = s[i] < 2.0 ? R(1.0) : 2.0 / s[i];
}

// ----------------------------------------------- face values (normal sweep)
// value and first derivative at the cell's left (side 0) or right (side 1)
// face from the candidates of weno_7th/9th_ao_with_df_left/right
template <int ORDER, class R>
inline void weno79_face(R& var, R& der1, int side, const R* w, const R* df, double h) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  R fw[NS];
  ao_weights<ORDER>(fw, w);
  const R w0 = w[0];
  R p[NS], px[NS];
  if (side == 0) {
    static const WT P0[4] = {{-1.0 / 6.0, -2}, {5.0 / 6.0, -1}, {1.0 / 3.0, 0}, {-1, 0}};
    static const WT P1[4] = {{1.0 / 3.0, -1}, {5.0 / 6.0, 0}, {-1.0 / 6.0, 1}, {-1, 0}};
    static const WT P2[4] = {{11.0 / 6.0, 0}, {-7.0 / 6.0, 1}, {1.0 / 3.0, 2}, {-1, 0}};
    static const WT P3[5] = {{47, 0}, {27, -1}, {-3, -2}, {-13, 1}, {2, 2}};
    static const WT P4[7] = {{319, 0}, {214, -1}, {-38, -2}, {4, -3}, {-101, 1}, {25, 2}, {-3, 3}};
    static const WT P5[9] = {{1879, 0}, {1375, -1}, {-305, -2}, {55, -3}, {-5, -4},
                             {-641, 1}, {199, 2}, {-41, 3}, {4, 4}};
    static const WT X0[2] = {{1, 0}, {-1, -1}};
    static const WT X2[3] = {{2, 0}, {-3, 1}, {1, 2}};
    static const WT X3[4] = {{15, 0}, {-15, -1}, {1, -2}, {-1, 1}};
    static const WT X4[6] = {{245, 0}, {-245, -1}, {25, -2}, {-2, -3}, {-25, 1}, {2, 2}};
    static const WT X5[8] = {{7175, 0}, {-7175, -1}, {889, -2}, {-119, -3}, {9, -4}, {-889, 1}, {119, 2}, {-9, 3}};
    p[0] = w0 + df[0] * tsum(P0, w);
    p[1] = w0 + df[1] * tsum(P1, w);
    p[2] = w0 + df[2] * tsum(P2, w);
    p[3] = w0 + df[3] * ((1.0 / 60.0) * tsum(P3, w) - w0);
    p[4] = w0 + df[4] * ((1.0 / 420.0) * tsum(P4, w) - w0);
    px[0] = df[0] * tsum(X0, w) / h;
    px[1] = df[1] * tsum(X0, w) / h;
    px[2] = -df[2] * (tsum(X2, w) / h);
    px[3] = df[3] * tsum(X3, w) / (12.0 * h);
    px[4] = df[4] * tsum(X4, w) / (180.0 * h);
    if constexpr (ORDER == 9) {
      p[NS - 1] = w0 + df[NS - 1] * ((1.0 / 2520.0) * tsum(P5, w) - w0);
      px[NS - 1] = df[NS - 1] * tsum(X5, w) / (5040.0 * h);
    }
  } else {
    static const WT P0[4] = {{1.0 / 3.0, -2}, {-7.0 / 6.0, -1}, {11.0 / 6.0, 0}, {-1, 0}};
    static const WT P1[4] = {{-1.0 / 6.0, -1}, {5.0 / 6.0, 0}, {1.0 / 3.0, 1}, {-1, 0}};
    static const WT P2[4] = {{1.0 / 3.0, 0}, {5.0 / 6.0, 1}, {-1.0 / 6.0, 2}, {-1, 0}};
    static const WT P3[5] = {{47, 0}, {-13, -1}, {2, -2}, {27, 1}, {-3, 2}};
    static const WT P4[7] = {{319, 0}, {-101, -1}, {25, -2}, {-3, -3}, {214, 1}, {-38, 2}, {4, 3}};
    static const WT P5[9] = {{1879, 0}, {-641, -1}, {199, -2}, {-41, -3}, {4, -4},
                             {1375, 1}, {-305, 2}, {55, 3}, {-5, 4}};
    static const WT X0[3] = {{2, 0}, {-3, -1}, {1, -2}};
    static const WT X1[2] = {{-1, 0}, {1, 1}};
    static const WT X3[4] = {{-15, 0}, {1, -1}, {15, 1}, {-1, 2}};
    static const WT X4[6] = {{-245, 0}, {25, -1}, {-2, -2}, {245, 1}, {-25, 2}, {2, 3}};
    static const WT X5[8] = {{-7175, 0}, {889, -1}, {-119, -2}, {9, -3}, {7175, 1}, {-889, 2}, {119, 3}, {-9, 4}};
    p[0] = w0 + df[0] * tsum(P0, w);
    p[1] = w0 + df[1] * tsum(P1, w);
    p[2] = w0 + df[2] * tsum(P2, w);
    p[3] = w0 + df[3] * ((1.0 / 60.0) * tsum(P3, w) - w0);
    p[4] = w0 + df[4] * ((1.0 / 420.0) * tsum(P4, w) - w0);
    px[0] = df[0] * tsum(X0, w) / h;
    px[1] = df[1] * tsum(X1, w) / h;
    px[2] = df[2] * tsum(X1, w) / h;
    px[3] = df[3] * tsum(X3, w) / (12.0 * h);
    px[4] = df[4] * tsum(X4, w) / (180.0 * h);
    if constexpr (ORDER == 9) {
      p[NS - 1] = w0 + df[NS - 1] * ((1.0 / 2520.0) * tsum(P5, w) - w0);
      px[NS - 1] = df[NS - 1] * tsum(X5, w) / (5040.0 * h);
    }
  }
  var = 0.0;
  der1 = 0.0;
  for (int k = 0; k < NS; k++) {
    var += fw[k] * p[k];
    der1 += fw[k] * px[k];
  }
}

// WENO7_AO_with_df / WENO9_AO_with_df: one cell, one direction, states in
// the x-like frame; w[-r..r] -> line state at the cell's left face (`left`)
// and right face (`right`), convar + der1x
template <int ORDER, class R>
inline void weno79_normal(Pt<R>& left, Pt<R>& right, const R* alpha /* centred */, const R* const* wv /* centred */,
                          double h, int recon_var, const Gas& g) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  constexpr int r = (ORDER - 1) / 2;
  R df[NS];
  df_factors<ORDER>(df, alpha);
  R pn1[4], p0[4], pp1[4], bl[4], br[4];
  convar_to_primvar_2d(pn1, wv[-1], g);
  convar_to_primvar_2d(p0, wv[0], g);
  convar_to_primvar_2d(pp1, wv[1], g);
  for (int i = 0; i < 4; i++) {
    bl[i] = 0.5 * (pn1[i] + p0[i]);
    br[i] = 0.5 * (pp1[i] + p0[i]);
  }
  R c[2 * r + 1][4];
  for (int side = 0; side < 2; side++) {
    const R* base = side == 0 ? bl : br;
    for (int s = -r; s <= r; s++) {
      if (recon_var == 0)
        for (int i = 0; i < 4; i++) c[s + r][i] = wv[s][i];
      else
        convar_to_char(c[s + r], base, wv[s], g);
    }
    R var[4], d1[4];
    for (int i = 0; i < 4; i++) {
      R line[2 * r + 1];
      for (int s = 0; s < 2 * r + 1; s++) line[s] = c[s][i];
      weno79_face<ORDER>(var[i], d1[i], side, line + r, df, h);
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
      right.convar[m] = wv[0][m];
      left.convar[m] = wv[0][m];
      right.der1x[m] = 0.0;
      left.der1x[m] = 0.0;
    }
  }
}

// ------------------------------------------- Gauss-point values (tangential)
// WENO7: 3 points, the face centre then -/+ sqrt(3/5) of the half face
// (weno_7th_ao_with_df_2gauss); values g[q] and tangential derivatives gx[q]
template <class R>
inline void weno7_gauss(R* g, R* gx, const R* w, const R* df, double h) {
  R fw[5];
  ao_weights<7>(fw, w);
  const R w0 = w[0];
  const double s = 3.872983346207417;  // sqrt(15) as the reference writes it
  R p[5], px[5];
  auto comb = [&](R& v, R& vx) {
    v = 0.0;
    vx = 0.0;
    for (int k = 0; k < 5; k++) {
      v += fw[k] * p[k];
      vx += fw[k] * px[k];
    }
  };
  {  // centre
    static const WT P0[3] = {{23, 0}, {2, -1}, {-1, -2}};
    static const WT P1[3] = {{26, 0}, {-1, -1}, {-1, 1}};
    static const WT P2[3] = {{23, 0}, {2, 1}, {-1, 2}};
    static const WT P3[5] = {{2134, 0}, {-116, -1}, {9, -2}, {-116, 1}, {9, 2}};
    static const WT P4[7] = {{121004, 0}, {-7621, -1}, {954, -2}, {-75, -3}, {-7621, 1}, {954, 2}, {-75, 3}};
    static const WT X0[3] = {{3, 0}, {-4, -1}, {1, -2}};
    static const WT X1[2] = {{-1, -1}, {1, 1}};
    static const WT X2[3] = {{-3, 0}, {4, 1}, {-1, 2}};
    static const WT X3[4] = {{-34, -1}, {5, -2}, {34, 1}, {-5, 2}};
    static const WT X4[6] = {{-9455, -1}, {2236, -2}, {-259, -3}, {9455, 1}, {-2236, 2}, {259, 3}};
    p[0] = w0 + df[0] * (tsum(P0, w) / 24.0 - w0);
    p[1] = w0 + df[1] * (tsum(P1, w) / 24.0 - w0);
    p[2] = w0 + df[2] * (tsum(P2, w) / 24.0 - w0);
    p[3] = w0 + df[3] * (tsum(P3, w) / 1920.0 - w0);
    p[4] = w0 + df[4] * (tsum(P4, w) / 107520.0 - w0);
    px[0] = df[0] * tsum(X0, w) / (2.0 * h);
    px[1] = df[1] * tsum(X1, w) / (2.0 * h);
    px[2] = df[2] * tsum(X2, w) / (2.0 * h);
    px[3] = df[3] * tsum(X3, w) / (48.0 * h);
    px[4] = df[4] * tsum(X4, w) / (11520.0 * h);
    comb(g[0], gx[0]);
  }
  for (int q = 1; q < 3; q++) {
    // the two off-centre points are mirror images of each other; each
    // coefficient is formed exactly as the reference forms it
    if (q == 1) {
      const WT A0[3] = {{62 - 9 * s, 0}, {4 * (-1 + 3 * s), -1}, {2 - 3 * s, -2}};
      const WT A1[3] = {{56, 0}, {2 + 3 * s, -1}, {2 - 3 * s, 1}};
      const WT A2[3] = {{62 + 9 * s, 0}, {-(4 * (1 + 3 * s)), 1}, {2 + 3 * s, 2}};
      const WT A3[8] = {{2186, 0}, {4 * (29 + 41 * s), -1}, {-9, -2}, {-(22 * s), -2},
                        {116, 1}, {-(164 * s), 1}, {-9, 2}, {22 * s, 2}};
      const WT A4[12] = {{4534440, 0}, {5 * (57144 + 78421 * s), -1}, {-35748, -2}, {-(84364 * s), -2},
                         {2808, -3}, {9541 * s, -3}, {285720, 1}, {-(392105 * s), 1},
                         {-35748, 2}, {84364 * s, 2}, {2808, 3}, {-(9541 * s), 3}};
      const WT B0[3] = {{-(-15 + s), 0}, {2 * (-10 + s), -1}, {-(-5 + s), -2}};
      const WT B1[3] = {{2 * s, 0}, {-(5 + s), -1}, {-(-5 + s), 1}};
      const WT B2[3] = {{-(15 + s), 0}, {2 * (10 + s), 1}, {-(5 + s), 2}};
      const WT B3[8] = {{78 * s, 0}, {-(2 * (95 + 21 * s)), -1}, {20, -2}, {3 * s, -2},
                        {190, 1}, {-(42 * s), 1}, {-20, 2}, {3 * s, 2}};
      const WT B4[12] = {{103860 * s, 0}, {-(5 * (49925 + 11619 * s)), -1}, {41300, -2}, {6678 * s, -2},
                         {-4325, -3}, {-(513 * s), -3}, {249625, 1}, {-(58095 * s), 1},
                         {-41300, 2}, {6678 * s, 2}, {4325, 3}, {-(513 * s), 3}};
      p[0] = w0 + df[0] * (tsum(A0, w) / 60.0 - w0);
      p[1] = w0 + df[1] * (tsum(A1, w) / 60.0 - w0);
      p[2] = w0 + df[2] * (tsum(A2, w) / 60.0 - w0);
      p[3] = w0 + df[3] * (tsum(A3, w) / 2400.0 - w0);
      p[4] = w0 + df[4] * (tsum(A4, w) / 5040000.0 - w0);
      px[0] = df[0] * tsum(B0, w) / (10.0 * h);
      px[1] = df[1] * tsum(B1, w) / (h * 10.0);
      px[2] = df[2] * tsum(B2, w) / (10.0 * h);
      px[3] = df[3] * tsum(B3, w) / (300.0 * h);
      px[4] = df[4] * tsum(B4, w) / (360000.0 * h);
    } else {
      const WT A0[3] = {{62 + 9 * s, 0}, {-(4 * (1 + 3 * s)), -1}, {2 + 3 * s, -2}};
      const WT A1[3] = {{56, 0}, {2 - 3 * s, -1}, {2 + 3 * s, 1}};
      const WT A2[3] = {{62 - 9 * s, 0}, {4 * (-1 + 3 * s), 1}, {2 - 3 * s, 2}};
      const WT A3[8] = {{2186, 0}, {-(4 * (-29 + 41 * s)), -1}, {-9, -2}, {22 * s, -2},
                        {116, 1}, {164 * s, 1}, {-9, 2}, {-(22 * s), 2}};
      const WT A4[12] = {{4534440, 0}, {-(5 * (-57144 + 78421 * s)), -1}, {-35748, -2}, {84364 * s, -2},
                         {2808, -3}, {-(9541 * s), -3}, {285720, 1}, {392105 * s, 1},
                         {-35748, 2}, {-(84364 * s), 2}, {2808, 3}, {9541 * s, 3}};
      const WT B0[3] = {{15 + s, 0}, {-(2 * (10 + s)), -1}, {5 + s, -2}};
      const WT B1[3] = {{-(2 * s), 0}, {-(5 - s), -1}, {5 + s, 1}};
      const WT B2[3] = {{-15 + s, 0}, {-(2 * (-10 + s)), 1}, {-5 + s, 2}};
      const WT B3[8] = {{-(78 * s), 0}, {2 * (-95 + 21 * s), -1}, {20, -2}, {-(3 * s), -2},
                        {190, 1}, {42 * s, 1}, {-20, 2}, {-(3 * s), 2}};
      const WT B4[12] = {{-(103860 * s), 0}, {5 * (-49925 + 11619 * s), -1}, {41300, -2}, {-(6678 * s), -2},
                         {-4325, -3}, {513 * s, -3}, {249625, 1}, {58095 * s, 1},
                         {-41300, 2}, {-(6678 * s), 2}, {4325, 3}, {513 * s, 3}};
      p[0] = w0 + df[0] * (tsum(A0, w) / 60.0 - w0);
      p[1] = w0 + df[1] * (tsum(A1, w) / 60.0 - w0);
      p[2] = w0 + df[2] * (tsum(A2, w) / 60.0 - w0);
      p[3] = w0 + df[3] * (tsum(A3, w) / 2400.0 - w0);
      p[4] = w0 + df[4] * (tsum(A4, w) / 5040000.0 - w0);
      px[0] = df[0] * tsum(B0, w) / (10.0 * h);
      px[1] = df[1] * tsum(B1, w) / (h * 10.0);
      px[2] = df[2] * tsum(B2, w) / (10.0 * h);
      px[3] = df[3] * tsum(B3, w) / (300.0 * h);
      px[4] = df[4] * tsum(B4, w) / (360000.0 * h);
    }
    comb(g[q], gx[q]);
  }
}

// c * x^n formed as ((c*x)*x)... (the reference writes the powers out)
template <class R> inline R mono(R c, R x, int n) {
  for (int k = 0; k < n; k++) c = c * x;
  return c;
}

// Polynomial_9th: the six candidates (values and derivatives) at the local
// coordinate x in [-1, 0] of the face (reconstruction.cpp:3551-3609)
template <class R> inline void poly9(R* p, R* px, const R* df, const R* w, R x, double h) {
  const R w0 = w[0];
  auto cand = [&](int k, const R* cf, int nc) {
    R v = cf[0];
    for (int n = 1; n < nc; n++) v = v + mono(cf[n], x, n);
    p[k] = w0 + df[k] * (v - w0);
    R dv = cf[1];
    for (int n = 2; n < nc; n++) dv = dv + mono(R(double(n)) * cf[n], x, n - 1);
    px[k] = df[k] * dv / h;
  };
  {
    static const WT A[3][3] = {{{11, 0}, {-7, -1}, {2, -2}}, {{5, 0}, {-1, -1}, {2, 1}}, {{2, 0}, {5, 1}, {-1, 2}}};
    static const WT B0[3] = {{2, 0}, {-3, -1}, {1, -2}};
    static const WT B12[2] = {{-1, 0}, {1, 1}};
    static const WT C[3][3] = {{{1, 0}, {-2, -1}, {1, -2}}, {{-2, 0}, {1, -1}, {1, 1}}, {{1, 0}, {-2, 1}, {1, 2}}};
    for (int k = 0; k < 3; k++) {
      R cf[3] = {tsum(A[k], w) / 6.0, k == 0 ? tsum(B0, w) : tsum(B12, w), tsum(C[k], w) / 2.0};
      cand(k, cf, 3);
    }
  }
  {
    static const WRow T[5] = {{5, {{47, 0}, {-13, -1}, {2, -2}, {27, 1}, {-3, 2}}},
                              {4, {{-15, 0}, {1, -1}, {15, 1}, {-1, 2}}},
                              {5, {{-8, 0}, {6, -1}, {-1, -2}, {2, 1}, {1, 2}}},
                              {4, {{3, 0}, {-1, -1}, {-3, 1}, {1, 2}}},
                              {5, {{6, 0}, {-4, -1}, {1, -2}, {-4, 1}, {1, 2}}}};
    static const double den[5] = {60.0, 12.0, 8.0, 6.0, 24.0};
    R cf[5];
    for (int n = 0; n < 5; n++) cf[n] = tsum(T[n], w) / den[n];
    cand(3, cf, 5);
  }
  {
    static const WRow T[7] = {{7, {{319, 0}, {-101, -1}, {25, -2}, {-3, -3}, {214, 1}, {-38, 2}, {4, 3}}},
                              {6, {{-245, 0}, {25, -1}, {-2, -2}, {245, 1}, {-25, 2}, {2, 3}}},
                              {7, {{-230, 0}, {210, -1}, {-57, -2}, {7, -3}, {15, 1}, {63, 2}, {-8, 3}}},
                              {6, {{28, 0}, {-11, -1}, {1, -2}, {-28, 1}, {11, 2}, {-1, 3}}},
                              {7, {{46, 0}, {-39, -1}, {15, -2}, {-2, -3}, {-24, 1}, {3, 2}, {1, 3}}},
                              {6, {{-10, 0}, {5, -1}, {-1, -2}, {10, 1}, {-5, 2}, {1, 3}}},
                              {7, {{-20, 0}, {15, -1}, {-6, -2}, {1, -3}, {15, 1}, {-6, 2}, {1, 3}}}};
    static const double den[7] = {420.0, 180.0, 240.0, 36.0, 144.0, 120.0, 720.0};
    R cf[7];
    for (int n = 0; n < 7; n++) cf[n] = tsum(T[n], w) / den[n];
    cand(4, cf, 7);
  }
  {
    static const WRow T[9] = {
        {9, {{1879, 0}, {-641, -1}, {199, -2}, {-41, -3}, {4, -4}, {1375, 1}, {-305, 2}, {55, 3}, {-5, 4}}},
        {8, {{-7175, 0}, {889, -1}, {-119, -2}, {9, -3}, {7175, 1}, {-889, 2}, {119, 3}, {-9, 4}}},
        {9, {{-27895, 0}, {28679, -1}, {-9835, -2}, {2081, -3}, {-205, -4}, {-2065, 1}, {11459, 2}, {-2455, 3}, {236, 4}}},
        {8, {{1365, 0}, {-587, -1}, {89, -2}, {-7, -3}, {-1365, 1}, {587, 2}, {-89, 3}, {7, 4}}},
        {9, {{1174, 0}, {-1160, -1}, {556, -2}, {-128, -3}, {13, -4}, {-464, 1}, {-68, 2}, {88, 3}, {-11, 4}}},
        {8, {{-75, 0}, {41, -1}, {-11, -2}, {1, -3}, {75, 1}, {-41, 2}, {11, 3}, {-1, 4}}},
        {9, {{-380, 0}, {334, -1}, {-170, -2}, {46, -3}, {-5, -4}, {250, 1}, {-86, 2}, {10, 3}, {1, 4}}},
        {8, {{35, 0}, {-21, -1}, {7, -2}, {-1, -3}, {-35, 1}, {21, 2}, {-7, 3}, {1, 4}}},
        {9, {{70, 0}, {-56, -1}, {28, -2}, {-8, -3}, {1, -4}, {-56, 1}, {28, 2}, {-8, 3}, {1, 4}}}};
    static const double den[9] = {2520.0, 5040.0, 30240.0, 1440.0, 3456.0, 480.0, 8640.0, 5040.0, 40320.0};
    R cf[9];
    for (int n = 0; n < 9; n++) cf[n] = tsum(T[n], w) / den[n];
    cand(5, cf, 9);
  }
}

// WENO9: 4 Lobatto points of the face, at x = -1, 0, -1/2 -/+ sqrt(5)/10
// (weno_9th_ao_with_df_2gauss)
template <class R>
inline void weno9_gauss(R* g, R* gx, const R* w, const R* df, double h) {
  R fw[6];
  ao_weights<9>(fw, w);
  const double sqrt5 = 2.236067977499790;
  const double xs[4] = {-1.0, 0.0, -0.5 - sqrt5 / 10.0, -0.5 + sqrt5 / 10.0};
  for (int q = 0; q < 4; q++) {
    R p[6], px[6];
    poly9(p, px, df, w, R(xs[q]), h);
    g[q] = 0.0;
    gx[q] = 0.0;
    for (int k = 0; k < 6; k++) {
      g[q] += fw[k] * p[k];
      gx[q] += fw[k] * px[k];
    }
  }
}

template <int ORDER, class R>
inline void weno79_gauss(R* g, R* gx, const R* w, const R* df, double h) {
  if constexpr (ORDER == 7) weno7_gauss(g, gx, w, df, h);
  else weno9_gauss(g, gx, w, df, h);
}

// weno_7th/9th_ao_with_df_tangential: both sides of one face.  lineL/lineR
// hold the 2r+1 line states along the face (index r = the face itself);
// gpL/gpR receive NG = 3 (WENO7) or 4 (WENO9) Gauss-point states.  WENO9 in
// characteristic variables keeps the reference's right-side projection of
// only the inner seven points: the outermost two reuse the left side's
// projections (reconstruction.cpp:3306-3312).
template <int ORDER, class R>
inline void weno79_tangent_face(Pt<R>* gpL, Pt<R>* gpR, const Pt<R>* lineL, const Pt<R>* lineR, const R* a1,
                                const R* a2, double h, int recon_var, const Gas& g) {
  constexpr int NS = ORDER == 7 ? 5 : 6;
  constexpr int r = (ORDER - 1) / 2;
  constexpr int NG = ORDER == 7 ? 3 : 4;
  R c[2 * r + 1][4];  // projected line values, kept across the two sides
  for (int side = 0; side < 2; side++) {
    const Pt<R>* line = side == 0 ? lineL : lineR;
    Pt<R>* gp = side == 0 ? gpL : gpR;
    R df[NS];
    df_factors<ORDER>(df, (side == 0 ? a1 : a2) + r);
    R base[4];
    convar_to_primvar_2d(base, line[r].convar, g);
    if (recon_var == 0) {
      for (int i = 0; i < 4; i++) {
        R w[2 * r + 1], v[NG], vx[NG];
        for (int s = 0; s < 2 * r + 1; s++) w[s] = line[s].convar[i];
        weno79_gauss<ORDER>(v, vx, w + r, df, h);
        for (int q = 0; q < NG; q++) {
          gp[q].convar[i] = v[q];
          gp[q].der1y[i] = vx[q];
        }
      }
    } else {
      for (int s = 0; s < 2 * r + 1; s++) {
        if (ORDER == 9 && side == 1 && (s == 0 || s == 2 * r)) continue;
        convar_to_char(c[s], base, line[s].convar, g);
      }
      R var[NG][4], d1[NG][4];
      for (int i = 0; i < 4; i++) {
        R w[2 * r + 1], v[NG], vx[NG];
        for (int s = 0; s < 2 * r + 1; s++) w[s] = c[s][i];
        weno79_gauss<ORDER>(v, vx, w + r, df, h);
        for (int q = 0; q < NG; q++) {
          var[q][i] = v[q];
          d1[q][i] = vx[q];
        }
      }
      for (int q = 0; q < NG; q++) {
        char_to_convar(gp[q].convar, base, var[q], g);
        char_to_convar(gp[q].der1y, base, d1[q], g);
      }
    }
    for (int i = 0; i < 4; i++) {
      R w[2 * r + 1], v[NG], vx[NG];
      for (int s = 0; s < 2 * r + 1; s++) w[s] = line[s].der1x[i];
      weno79_gauss<ORDER>(v, vx, w + r, df, h);
      for (int q = 0; q < NG; q++) gp[q].der1x[i] = v[q];
    }
  }
  for (int side = 0; side < 2; side++) {
    const Pt<R>* line = side == 0 ? lineL : lineR;
    Pt<R>* gp = side == 0 ? gpL : gpR;
    for (int q = 0; q < NG; q++) {
      if (reduce_order(gp[q].convar, g)) {
        for (int v = 0; v < 4; v++) {
          gp[q].convar[v] = line[r].convar[v];
          gp[q].der1x[v] = line[r].der1x[v];
          gp[q].der1y[v] = 0.0;
        }
      }
    }
  }
}

}  // namespace orc
