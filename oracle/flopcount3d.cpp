// flopcount3d.cpp — fixes the algorithmic FP64 operation counts of the 3D
// path's units (one cell's face values in one direction, one face side's
// line values, one face side's point values along one line, one Gauss
// point's GKS3D) by running the 3D restatement's leaves on a counting
// scalar.  TEST INFRASTRUCTURE ONLY.
//
// Counted as in flopcount.cpp: every +, -, *, / and sqrt between runtime
// values (1 each); exp and erfc separately; not counted:
// negation, fabs, comparisons, arithmetic the compiler folds between
// constants.  Output: one line of JSON.
#include <cmath>
#include <cstdio>

namespace cnt {
static long long n_arith = 0, n_exp = 0, n_erfc = 0;
}
struct CD {
  double v;
  CD() : v(0.0) {}
  CD(double x) : v(x) {}
  CD& operator+=(const CD& o) {
    cnt::n_arith++;
    v += o.v;
    return *this;
  }
};
inline CD operator+(const CD& a, const CD& b) { cnt::n_arith++; return CD(a.v + b.v); }
inline CD operator-(const CD& a, const CD& b) { cnt::n_arith++; return CD(a.v - b.v); }
inline CD operator*(const CD& a, const CD& b) { cnt::n_arith++; return CD(a.v * b.v); }
inline CD operator/(const CD& a, const CD& b) { cnt::n_arith++; return CD(a.v / b.v); }
inline CD operator-(const CD& a) { return CD(-a.v); }
inline bool operator<(const CD& a, const CD& b) { return a.v < b.v; }
inline bool operator>(const CD& a, const CD& b) { return a.v > b.v; }
inline bool operator<=(const CD& a, const CD& b) { return a.v <= b.v; }
inline bool operator!=(const CD& a, const CD& b) { return a.v != b.v; }
// found by argument-dependent lookup from the templated leaves
inline CD sqrt(const CD& a) { cnt::n_arith++; return CD(std::sqrt(a.v)); }
inline CD exp(const CD& a) { cnt::n_exp++; return CD(std::exp(a.v)); }
inline CD erfc(const CD& a) { cnt::n_erfc++; return CD(std::erfc(a.v)); }
inline CD fabs(const CD& a) { return CD(a.v < 0 ? -a.v : a.v); }

#include "solver3d.hpp"

using namespace orc3d;

static long long snap() { return cnt::n_arith; }

int main() {
  Gas3 g{2, 1.4, -1.0, -1.0, 0.05, 1.0, 0};  // CubicTube's gas, Euler collision time
  // a smooth stencil of five states (no order reduction)
  const double prim[5][5] = {{1.00, 0.30, -0.20, 0.10, 1.00},
                             {1.05, 0.32, -0.18, 0.12, 1.04},
                             {1.12, 0.35, -0.15, 0.15, 1.10},
                             {1.18, 0.37, -0.13, 0.16, 1.15},
                             {1.22, 0.38, -0.12, 0.18, 1.19}};
  CD w[5][5];
  for (int s = 0; s < 5; s++) {
    CD p[5];
    for (int v = 0; v < 5; v++) p[v] = prim[s][v];
    primvar_to_convar(w[s], p, g);
  }
  const double h = 0.01;
  long long c0 = snap();
  CD lo[10], ro[10];
  weno5_face(lo, ro, w, h, 1, g);
  const long long face_cell = snap() - c0;
  // line values: stencil of 5 face records (convar, der1x), one side
  CD s10[5][10], s15[5][15];
  for (int q = 0; q < 5; q++)
    for (int m = 0; m < 10; m++) s10[q][m] = m < 5 ? w[q][m] : 0.1 * w[q][m - 5];
  for (int q = 0; q < 5; q++)
    for (int m = 0; m < 15; m++) s15[q][m] = m < 5 ? w[q][m] : (m < 10 ? 0.1 * w[q][m - 5] : 0.05 * w[q][m - 10]);
  CD out10[2][15], out15[2][20];
  c0 = snap();
  weno5_tangential_side<10>(out10, s10, h, 1, g);
  const long long line_side = snap() - c0;
  c0 = snap();
  weno5_tangential_side<15>(out15, s15, h, 1, g);
  const long long point_side = snap() - c0;
  // GKS3D at one Gauss point (slopes of the reconstruction size)
  CD L[20], R[20], f[5], d[5];
  for (int v = 0; v < 5; v++) {
    L[v] = w[1][v];
    R[v] = w[2][v];
    for (int k = 1; k < 4; k++) {
      L[5 * k + v] = 0.3 * w[1][v];
      R[5 * k + v] = -0.2 * w[2][v];
    }
  }
  long long e0 = cnt::n_exp, r0 = cnt::n_erfc;
  c0 = snap();
  gks3d_point(f, d, static_cast<CD*>(nullptr), L, R, 1e-3, g);
  const long long flux_point = snap() - c0;
  const long long flux_exp = cnt::n_exp - e0, flux_erfc = cnt::n_erfc - r0;
  // one interior cell's S2O4 step: 2 stages x 3 directions x (face values
  // of the cell, 2 line sides, 2 lines x 2 point sides, 4 Gauss points) —
  // one face per cell and direction; the update and the primitive sweep
  // add a few tens
  const double cell_step = 2.0 * 3.0 * (face_cell + 2.0 * line_side + 4.0 * point_side + 4.0 * flux_point);
  printf("{\"face_cell_dir\": %lld, \"line_face_side\": %lld, \"point_face_line_side\": %lld, "
         "\"flux_point\": %lld, \"flux_point_exp\": %lld, \"flux_point_erfc\": %lld, \"cell_step\": %.0f}\n",
         face_cell, line_side, point_side, flux_point, flux_exp, flux_erfc, cell_step);
  return 0;
}
