// flopcount.cpp — fixes the algorithmic FP64 operation count of one cell
// update (one full S2O4 step of one interior cell) by running the oracle's
// sweeps on a counting scalar.  TEST INFRASTRUCTURE ONLY.
//
// Counted: every +, -, *, / and sqrt between runtime values (1 each);
// exp and erfc are counted separately.  Not counted: negation, fabs,
// comparisons, and operations the compiler folds between constants.
// Output (one line of JSON): flops and transcendental calls per cell update
// and per (face, Gauss point) of the flux leaf.
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace cnt {
static long long n_arith = 0, n_exp = 0, n_erfc = 0, n_sqrt = 0;
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
  CD& operator*=(const CD& o) {
    cnt::n_arith++;
    v *= o.v;
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
inline bool operator==(const CD& a, const CD& b) { return a.v == b.v; }
inline bool operator!=(const CD& a, const CD& b) { return a.v != b.v; }

namespace orc {
inline CD sqrt(const CD& a) { cnt::n_arith++; cnt::n_sqrt++; return CD(std::sqrt(a.v)); }
inline CD exp(const CD& a) { cnt::n_exp++; return CD(std::exp(a.v)); }
inline CD erfc(const CD& a) { cnt::n_erfc++; return CD(std::erfc(a.v)); }
inline CD fabs(const CD& a) { return CD(a.v < 0 ? -a.v : a.v); }
inline double as_double(const CD& a) { return a.v; }
}  // namespace orc

#include "solver2d.hpp"

namespace orc {
template <> double Solver2d<CD>::to_double(const CD& r) { return r.v; }
}  // namespace orc

int main(int argc, char** argv) {
  int N = argc > 1 ? atoi(argv[1]) : 32;
  gks2d_params p = {};
  p.nodex = N;
  p.nodey = N;
  p.ghost = 5;
  p.gausspoint = 2;
  p.weno_order = 5;
  p.recon_variable = GKS_VAR_CHARACTERISTIC;
  p.use_df = 1;
  p.tau_type = GKS_TAU_EULER;
  p.K = 3;
  p.gamma = 1.4;
  p.mu = -1.0;
  p.nu = -1.0;
  p.pr = 1.0;
  p.c1_euler = 0.05;
  p.c2_euler = 1.0;
  p.left = 0.0;
  p.right = 1.0;
  p.down = 0.0;
  p.up = 1.0;
  p.cfl = 0.5;
  orc::Solver2d<CD> s(p);
  // four-quadrant Riemann state (smooth enough that no order reduction fires)
  const double zone[4][4] = {{0.138, 1.206, 1.206, 0.029},
                             {0.5323, 0.0, 1.206, 0.3},
                             {1.5, 0.0, 0.0, 1.5},
                             {0.5323, 1.206, 0.0, 0.3}};
  for (int i = 0; i < s.nx; i++)
    for (int j = 0; j < s.ny; j++) {
      int q;
      bool xl = (i - 5 + 1) < 0.5 * N, yl = (j - 5 + 1) < 0.5 * N;
      q = (xl && yl) ? 0 : (!xl && yl) ? 1 : (!xl && !yl) ? 2 : 3;
      CD pr[4] = {zone[q][0], zone[q][1], zone[q][2], zone[q][3]}, c[4];
      orc::primvar_to_convar_2d(c, pr, s.g);
      for (int v = 0; v < 4; v++) {
        s.U[v][i * s.ny + j] = c[v];
        s.prim[v][i * s.ny + j] = pr[v];
      }
    }
  int bc[4] = {0, 0, 0, 0};
  // one warm step so alpha is populated, then count one full step
  s.run(bc, 1.0, 1, nullptr);
  long long a0 = cnt::n_arith, e0 = cnt::n_exp, r0 = cnt::n_erfc;
  s.copy_old();
  CD dt = s.cfl(1.0);
  long long a_before = cnt::n_arith, e_before = cnt::n_exp, r_before = cnt::n_erfc;
  (void)a0; (void)e0; (void)r0;
  long long part[5] = {0, 0, 0, 0, 0};
  for (int st = 0; st < 2; st++) {
    for (int sd = 0; sd < 4; sd++) s.boundary(sd, bc[sd]);
    long long c0 = cnt::n_arith;
    s.primvar();
    long long c1 = cnt::n_arith;
    s.reconstruct();
    long long c2 = cnt::n_arith;
    s.flux(st, dt.v);
    long long c3 = cnt::n_arith;
    s.update(st);
    long long c4 = cnt::n_arith;
    s.update_alpha();
    long long c5 = cnt::n_arith;
    part[0] += c1 - c0;
    part[1] += c2 - c1;
    part[2] += c3 - c2;
    part[3] += c4 - c3;
    part[4] += c5 - c4;
  }
  double cells = (double)N * N;
  double arith = (double)(cnt::n_arith - a_before) / cells;
  double ex = (double)(cnt::n_exp - e_before) / cells;
  double er = (double)(cnt::n_erfc - r_before) / cells;
  // flux leaf alone: one point
  orc::Pt<CD> L, R, C;
  for (int v = 0; v < 4; v++) {
    L.convar[v] = s.xgL[(size_t)s.face(10, 10) * 2].convar[v];
    R.convar[v] = s.xgR[(size_t)s.face(10, 10) * 2].convar[v];
    L.der1x[v] = 0.1; L.der1y[v] = 0.1; R.der1x[v] = 0.1; R.der1y[v] = 0.1;
  }
  long long f0 = cnt::n_arith, fe0 = cnt::n_exp, fr0 = cnt::n_erfc;
  orc::MM2<CD> ml, mr;
  orc::center_collision(C, ml, mr, L, R, s.g);
  CD f[4], d[4];
  orc::gks2d_flux(f, d, L, R, C, ml, mr, 1e-4, s.g);
  const long long fp_arith = cnt::n_arith - f0, fp_exp = cnt::n_exp - fe0, fp_erfc = cnt::n_erfc - fr0;
  // per-kernel units on developed data (face (10,10) of the last stage)
  long long u0 = cnt::n_arith;
  {
    orc::Pt<CD> lo, ro;
    CD w[5][4], a[5] = {0.0, 0.0, 3.0, 0.0, 0.0};
    for (int q = 0; q < 5; q++) s.get_cell(w[q], (8 + q) * s.ny + 10);
    orc::weno5_df_normal(lo, ro, a, w[0], w[1], w[2], w[3], w[4], s.dx, 1, s.g);
  }
  long long u1 = cnt::n_arith;
  {
    orc::Pt<CD> line[5], out[2];
    CD a[5] = {0.0, 0.0, 3.0, 0.0, 0.0};
    for (int q = 0; q < 5; q++) line[q] = s.xlL[s.face(10, 8 + q)];
    orc::weno5_df_tangent_side(out, line, a, s.dy, 1, s.g);
  }
  long long u2 = cnt::n_arith;
  {
    CD pl[4], pr[4];
    orc::convar_to_primvar_2d(pl, L.convar, s.g);
    orc::convar_to_primvar_2d(pr, R.convar, s.g);
    CD a = orc::alpha_k(pl, pr, s.g);
    (void)a;
  }
  long long u3 = cnt::n_arith;
  printf("{\"normal_cell_dir\": %lld, \"tangent_face_side\": %lld, \"alpha_point\": %lld, ",
         u1 - u0, u2 - u1, u3 - u2);
  printf("\"N\": %d, \"flops_per_cell_update\": %.1f, \"exp_per_cell_update\": %.2f, "
         "\"erfc_per_cell_update\": %.2f, \"primvar\": %.1f, \"reconstruct\": %.1f, \"flux\": %.1f, "
         "\"update\": %.1f, \"alpha\": %.1f, \"flux_point_flops\": %lld, \"flux_point_exp\": %lld, "
         "\"flux_point_erfc\": %lld}\n",
         N, arith, ex, er, part[0] / cells, part[1] / cells, part[2] / cells, part[3] / cells,
         part[4] / cells, fp_arith, fp_exp, fp_erfc);
  return 0;
}
