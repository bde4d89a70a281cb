// solver2d.hpp — CPU restatement of the reference's 2D sweeps (test
// infrastructure only; see gks_core.hpp).  Data live in SoA arrays indexed
// like the reference (cell i*ny+j, face i*(ny+1)+j for both directions) so the
// loop ranges below can be read against the reference's line by line.
#pragma once
#include <cstring>
#include <vector>

#include "../include/gks2d.h"
#include "gks_core.hpp"
#include "weno79.hpp"

namespace orc {

inline Gas gas_from(const gks2d_params& p) {
  Gas g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.pr = p.pr;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  g.prandtl_fix = p.prandtl_fix;
  return g;
}

template <class R> struct Solver2d {
  gks2d_params P;
  Gas g;
  int nx, ny, G, gp;
  double dx, dy, area;
  double gw[4];
  std::vector<R> U[4], Uold[4], prim[4];
  std::vector<R> ax, ay;
  // line states: lL = line.left (state left of the face), lR = line.right
  std::vector<Pt<R>> xlL, xlR, ylL, ylR;
  // Gauss-point states, index face*gp + q
  std::vector<Pt<R>> xgL, xgR, ygL, ygR;
  // fluxes per stage: [(face*gp+q)*8 + (0..3 f, 4..7 derf)]
  std::vector<R> xfl[2], yfl[2];
  bool blowup = false;
  double t = 0.0;
  int step = 0;

  int cell(int i, int j) const { return i * ny + j; }
  int face(int i, int j) const { return i * (ny + 1) + j; }

  explicit Solver2d(const gks2d_params& p) : P(p), g(gas_from(p)) {
    G = p.ghost;
    nx = p.nodex + 2 * G;
    ny = p.nodey + 2 * G;
    gp = p.gausspoint;
    dx = (p.right - p.left) / p.nodex;
    dy = (p.up - p.down) / p.nodey;
    area = dx * dy;
    // SetGuassPoint, fluid_mesh.cpp:9-45
    if (gp == 2) {
      gw[0] = 0.5;
      gw[1] = 0.5;
    } else if (gp == 3) {
      gw[0] = 4.0 / 9.0;
      gw[1] = 0.5 - 0.5 * gw[0];
      gw[2] = gw[1];
    } else if (gp == 4) {
      gw[0] = 1.0 / 12.0;
      gw[1] = 1.0 / 12.0;
      gw[2] = 5.0 / 12.0;
      gw[3] = 5.0 / 12.0;
    } else {
      gw[0] = 1.0;
    }
    size_t nc = (size_t)nx * ny, nf = (size_t)(nx + 1) * (ny + 1);
    for (int v = 0; v < 4; v++) {
      U[v].assign(nc, R(0.0));
      Uold[v].assign(nc, R(0.0));
      prim[v].assign(nc, R(0.0));
    }
    ax.assign(nc, R(0.0));
    ay.assign(nc, R(0.0));
    Pt<R> z;
    for (int k = 0; k < 4; k++) z.convar[k] = z.der1x[k] = z.der1y[k] = R(0.0);
    xlL.assign(nf, z);
    xlR.assign(nf, z);
    ylL.assign(nf, z);
    ylR.assign(nf, z);
    xgL.assign(nf * gp, z);
    xgR.assign(nf * gp, z);
    ygL.assign(nf * gp, z);
    ygR.assign(nf * gp, z);
    for (int s = 0; s < 2; s++) {
      xfl[s].assign(nf * gp * 8, R(0.0));
      yfl[s].assign(nf * gp * 8, R(0.0));
    }
  }

  void get_cell(R* c, int idx) const {
    for (int v = 0; v < 4; v++) c[v] = U[v][idx];
  }

  // ------------------------------------------------------------- boundaries
  // boundary_condition.cpp:66-186 (free, periodic) and :367-442 (reflective)
  void boundary(int side, int kind) {
    if (side == GKS_SIDE_LEFT || side == GKS_SIDE_RIGHT) {
      for (int gi = 0; gi < G; gi++) {
        int i = side == GKS_SIDE_LEFT ? G - 1 - gi : nx - G + gi;
        int src;
        if (kind == GKS_BC_FREE) src = side == GKS_SIDE_LEFT ? i + 1 : i - 1;
        else if (kind == GKS_BC_PERIODIC) src = side == GKS_SIDE_LEFT ? i + P.nodex : i - P.nodex;
        else src = side == GKS_SIDE_LEFT ? 2 * G - 1 - i : 2 * nx - 2 * G - 1 - i;
        for (int j = 0; j < ny; j++)
          for (int v = 0; v < 4; v++) {
            R val = U[v][cell(src, j)];
            if (kind == GKS_BC_REFLECT && v == 1) val = -val;
            U[v][cell(i, j)] = val;
          }
      }
    } else {
      for (int gj = 0; gj < G; gj++) {
        int j = side == GKS_SIDE_DOWN ? G - 1 - gj : ny - G + gj;
        int src;
        if (kind == GKS_BC_FREE) src = side == GKS_SIDE_DOWN ? j + 1 : j - 1;
        else if (kind == GKS_BC_PERIODIC) src = side == GKS_SIDE_DOWN ? j + P.nodey : j - P.nodey;
        else src = side == GKS_SIDE_DOWN ? 2 * G - 1 - j : 2 * ny - 2 * G - 1 - j;
        for (int i = 0; i < nx; i++)
          for (int v = 0; v < 4; v++) {
            R val = U[v][cell(i, src)];
            if (kind == GKS_BC_REFLECT && v == 2) val = -val;
            U[v][cell(i, j)] = val;
          }
      }
    }
  }

  // Convar_to_Primvar: basic_function.cpp:752-773
  void primvar() {
    bool bad = false;
#pragma omp parallel for reduction(|| : bad)
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        R c[4], p[4];
        get_cell(c, cell(i, j));
        convar_to_primvar_2d(p, c, g);
        for (int v = 0; v < 4; v++) prim[v][cell(i, j)] = p[v];
        if (p[0] < 0 || p[3] < 0 || p[0] != p[0] || p[3] != p[3]) bad = true;
      }
    if (bad) blowup = true;
  }

  // CopyFluid_new_to_old: basic_function.cpp:1571-1584
  void copy_old() {
#pragma omp parallel for
    for (int i = G; i < G + P.nodex; i++)
      for (int j = G; j < G + P.nodey; j++)
        for (int v = 0; v < 4; v++) Uold[v][cell(i, j)] = U[v][cell(i, j)];
  }

  // Dtx (2D): time_advance.cpp:292-305
  R dtx(R dtx_in, double h, R rho, R u, R v, R p) const {
    R d = dtx_in;
    R tmp = sqrt(u * u + v * v) + sqrt(g.gamma * p / rho);
    if (tmp > P.cfl * h / d) d = P.cfl * h / tmp;
    if (d > 0.25 * P.cfl * h * h / g.mu && g.tau_type == GKS_TAU_NS && g.mu > 0)
      d = 0.25 * P.cfl * h * h / g.mu;
    return d;
  }
  // Get_CFL (2D): time_advance.cpp:264-290, reads the stored primvar
  R cfl(double tstop) const {
    R d = dx;
    for (int i = G; i < nx - G; i++)
      for (int j = G; j < ny - G; j++) {
        int c = cell(i, j);
        d = dtx(d, dx, prim[0][c], prim[1][c], prim[2][c], prim[3][c]);
        d = dtx(d, dy, prim[0][c], prim[1][c], prim[2][c], prim[3][c]);
      }
    if (t + d > tstop) d = tstop - t + 1e-16;
    return d;
  }

  // Reconstruction_within_cell: reconstruction.cpp:1415-1437 with the
  // WENO5_AO_with_df_normal (:1701-1729) and _tangent (:2553-2576) leaves.
  void reconstruct() {
    if (P.weno_order == 7 && gp == 3) return reconstruct79<7>();
    if (P.weno_order == 9 && gp == 4) return reconstruct79<9>();
    if (P.weno_order != 5 || gp != 2) return;
#pragma omp parallel for
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        if (i > G - 2 && i < nx - G + 1) {
          R w[5][4], a[5];
          for (int s = 0; s < 5; s++) {
            get_cell(w[s], cell(i - 2 + s, j));
            a[s] = ax[cell(i - 2 + s, j)];
          }
          weno5_df_normal(xlR[face(i, j)], xlL[face(i + 1, j)], a, w[0], w[1], w[2], w[3], w[4], dx,
                          P.recon_variable, g);
        }
        if (j > G - 2 && j < ny - G + 1) {
          R w[5][4], a[5], tmp[4];
          for (int s = 0; s < 5; s++) {
            get_cell(tmp, cell(i, j - 2 + s));
            y_to_x(w[s], tmp);
            a[s] = ay[cell(i, j - 2 + s)];
          }
          weno5_df_normal(ylR[face(i, j)], ylL[face(i, j + 1)], a, w[0], w[1], w[2], w[3], w[4], dy,
                          P.recon_variable, g);
        }
      }
#pragma omp parallel for
    for (int i = G - 1; i < nx - G + 1; i++)
      for (int j = G - 1; j < ny - G + 1; j++) {
        // x-face (i+1, j): stencil of faces (i+1, j-2..j+2)
        {
          int f = face(i + 1, j);
          R a1[5], a2[5];
          Pt<R> lineL[5], lineR[5];
          for (int s = 0; s < 5; s++) {
            a1[s] = ay[cell(i, j - 2 + s)];
            a2[s] = ay[cell(i + 1, j - 2 + s)];
            lineL[s] = xlL[face(i + 1, j - 2 + s)];
            lineR[s] = xlR[face(i + 1, j - 2 + s)];
          }
          weno5_df_tangent_side(&xgL[(size_t)f * gp], lineL, a1, dy, P.recon_variable, g);
          weno5_df_tangent_side(&xgR[(size_t)f * gp], lineR, a2, dy, P.recon_variable, g);
        }
        // y-face (i, j+1): stencil of faces (i+2..i-2, j+1), reversed
        {
          int f = face(i, j + 1);
          R a1[5], a2[5];
          Pt<R> lineL[5], lineR[5];
          for (int s = 0; s < 5; s++) {
            a1[s] = ax[cell(i + 2 - s, j)];
            a2[s] = ax[cell(i + 2 - s, j + 1)];
            lineL[s] = ylL[face(i + 2 - s, j + 1)];
            lineR[s] = ylR[face(i + 2 - s, j + 1)];
          }
          weno5_df_tangent_side(&ygL[(size_t)f * gp], lineL, a1, dx, P.recon_variable, g);
          weno5_df_tangent_side(&ygR[(size_t)f * gp], lineR, a2, dx, P.recon_variable, g);
        }
      }
  }

  // Reconstruction_within_cell with the WENO7/9_AO_with_df_normal and
  // _tangent leaves (reconstruction.cpp:1861-1892, 2030-2061, 2840-2865,
  // 3176-3207): the same sweeps as WENO5 with stencils of radius 3 / 4
  template <int ORDER> void reconstruct79() {
    constexpr int r = (ORDER - 1) / 2;
#pragma omp parallel for
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        if (i > G - 2 && i < nx - G + 1) {
          R w[2 * r + 1][4], a[2 * r + 1];
          const R* wv[2 * r + 1];
          for (int s = 0; s < 2 * r + 1; s++) {
            get_cell(w[s], cell(i - r + s, j));
            a[s] = ax[cell(i - r + s, j)];
            wv[s] = w[s];
          }
          weno79_normal<ORDER>(xlR[face(i, j)], xlL[face(i + 1, j)], a + r, wv + r, dx, P.recon_variable, g);
        }
        if (j > G - 2 && j < ny - G + 1) {
          R w[2 * r + 1][4], a[2 * r + 1], tmp[4];
          const R* wv[2 * r + 1];
          for (int s = 0; s < 2 * r + 1; s++) {
            get_cell(tmp, cell(i, j - r + s));
            y_to_x(w[s], tmp);
            a[s] = ay[cell(i, j - r + s)];
            wv[s] = w[s];
          }
          weno79_normal<ORDER>(ylR[face(i, j)], ylL[face(i, j + 1)], a + r, wv + r, dy, P.recon_variable, g);
        }
      }
#pragma omp parallel for
    for (int i = G - 1; i < nx - G + 1; i++)
      for (int j = G - 1; j < ny - G + 1; j++) {
        {  // x-face (i+1, j): faces (i+1, j-r..j+r)
          int f = face(i + 1, j);
          R a1[2 * r + 1], a2[2 * r + 1];
          Pt<R> lineL[2 * r + 1], lineR[2 * r + 1];
          for (int s = 0; s < 2 * r + 1; s++) {
            a1[s] = ay[cell(i, j - r + s)];
            a2[s] = ay[cell(i + 1, j - r + s)];
            lineL[s] = xlL[face(i + 1, j - r + s)];
            lineR[s] = xlR[face(i + 1, j - r + s)];
          }
          weno79_tangent_face<ORDER>(&xgL[(size_t)f * gp], &xgR[(size_t)f * gp], lineL, lineR, a1, a2, dy,
                                     P.recon_variable, g);
        }
        {  // y-face (i, j+1): faces (i+r..i-r, j+1), reversed
          int f = face(i, j + 1);
          R a1[2 * r + 1], a2[2 * r + 1];
          Pt<R> lineL[2 * r + 1], lineR[2 * r + 1];
          for (int s = 0; s < 2 * r + 1; s++) {
            a1[s] = ax[cell(i + r - s, j)];
            a2[s] = ax[cell(i + r - s, j + 1)];
            lineL[s] = ylL[face(i + r - s, j + 1)];
            lineR[s] = ylR[face(i + r - s, j + 1)];
          }
          weno79_tangent_face<ORDER>(&ygL[(size_t)f * gp], &ygR[(size_t)f * gp], lineL, lineR, a1, a2, dx,
                                     P.recon_variable, g);
        }
      }
  }

  // Reconstruction_forg0 + Calculate_flux: reconstruction.cpp:3612-3633,
  // flux_function.cpp:317-343; centre state computed once per point.
  void flux(int stage, double dt) {
#pragma omp parallel for
    for (int i = G; i < P.nodex + G + 1; i++)
      for (int j = G; j < P.nodey + G; j++)
        for (int q = 0; q < gp; q++) {
          size_t k = (size_t)face(i, j) * gp + q;
          Pt<R> c;
          MM2<R> ml, mr;
          center_collision(c, ml, mr, xgL[k], xgR[k], g);
          gks2d_flux(&xfl[stage][k * 8], &xfl[stage][k * 8 + 4], xgL[k], xgR[k], c, ml, mr, dt, g);
        }
#pragma omp parallel for
    for (int i = G; i < P.nodex + G; i++)
      for (int j = G; j < P.nodey + G + 1; j++)
        for (int q = 0; q < gp; q++) {
          size_t k = (size_t)face(i, j) * gp + q;
          Pt<R> c;
          MM2<R> ml, mr;
          center_collision(c, ml, mr, ygL[k], ygR[k], g);
          gks2d_flux(&yfl[stage][k * 8], &yfl[stage][k * 8 + 4], ygL[k], ygR[k], c, ml, mr, dt, g);
        }
  }

  // S2O4_2D coefficients, time_advance.cpp:221-230
  static double coe(int s, int k, int d) {
    static const double c[2][2][2] = {{{0.5, 1.0 / 8.0}, {0.0, 0.0}}, {{1.0, 1.0 / 6.0}, {0.0, 1.0 / 3.0}}};
    return c[s][k][d];
  }
  // stage-combined Gauss-point flux x[4] (time_advance.cpp:380-426)
  void combine(R* x, const std::vector<R>* fl, size_t k, int stage, int q) const {
    for (int v = 0; v < 4; v++) {
      R F = 0.0;
      for (int s = 0; s < stage + 1; s++)
        F = F + gw[q] * (coe(stage, s, 0) * fl[s][k * 8 + v] + coe(stage, s, 1) * fl[s][k * 8 + 4 + v]);
      x[v] = F;
    }
  }
  // Update + Update_with_gauss: time_advance.cpp:307-474
  void update(int stage) {
    const double nxn[2] = {1.0, 0.0}, nyn[2] = {0.0, 1.0};
#pragma omp parallel for
    for (int i = G; i < P.nodex + G; i++)
      for (int j = G; j < P.nodey + G; j++) {
        int c = cell(i, j);
        R xs[4][4][4];  // [face: bottom, top, left, right][q][v]
        int fb = face(i, j), ft = face(i, j + 1), fl_ = face(i, j), fr = face(i + 1, j);
        for (int q = 0; q < gp; q++) {
          combine(xs[0][q], yfl, (size_t)fb * gp + q, stage, q);
          local_to_global(xs[0][q], nyn);
          combine(xs[1][q], yfl, (size_t)ft * gp + q, stage, q);
          local_to_global(xs[1][q], nyn);
          combine(xs[2][q], xfl, (size_t)fl_ * gp + q, stage, q);
          local_to_global(xs[2][q], nxn);
          combine(xs[3][q], xfl, (size_t)fr * gp + q, stage, q);
          local_to_global(xs[3][q], nxn);
        }
        for (int v = 0; v < 4; v++) {
          R u = Uold[v][c];
          R total = 0.0;
          for (int q = 0; q < gp; q++) {
            total += dx * xs[0][q][v];
            total += -dx * xs[1][q][v];
            total += dy * xs[2][q][v];
            total += -dy * xs[3][q][v];
          }
          u += total / area;
          U[v][c] = u;
        }
      }
  }

  // Calculate_alpha for one face (average over Gauss points, threshold):
  // reconstruction.cpp:1466-1516.  y faces go back to the global frame first.
  R face_alpha(const std::vector<Pt<R>>& L, const std::vector<Pt<R>>& Rt, int f, bool yface) const {
    R a[4];
    for (int q = 0; q < gp; q++) {
      R pl[4], pr[4];
      convar_to_primvar_2d(pl, L[(size_t)f * gp + q].convar, g);
      convar_to_primvar_2d(pr, Rt[(size_t)f * gp + q].convar, g);
      if (yface) {
        R tl[4], tr[4];
        x_to_y(tl, pl);
        x_to_y(tr, pr);
        a[q] = alpha_k(tl, tr, g);
      } else {
        a[q] = alpha_k(pl, pr, g);
      }
    }
    R s;
    if (gp == 2) s = 0.5 * (a[0] + a[1]);
    else if (gp == 3) s = (a[0] + a[1] + a[2]) / 3.0;
    else s = 0.25 * (a[0] + a[1] + a[2] + a[3]);
    return s < 2.0 ? R(0.0) : s;
  }
  // Update_alpha: reconstruction.cpp:1523-1534.  The reference bounds j by
  // nodex (:1528); that is kept for nodey >= nodex, and clipped to nodey
  // otherwise (the reference then reads faces no sweep wrote).
  void update_alpha() {
    int jend = G + (P.nodey < P.nodex ? P.nodey : P.nodex);
#pragma omp parallel for
    for (int i = G; i < P.nodex + G; i++)
      for (int j = G; j < jend; j++) {
        R al = face_alpha(xgL, xgR, face(i, j), false);
        R ar = face_alpha(xgL, xgR, face(i + 1, j), false);
        R ad = face_alpha(ygL, ygR, face(i, j), true);
        R au = face_alpha(ygL, ygR, face(i, j + 1), true);
        ax[cell(i, j)] = al + ar;
        ay[cell(i, j)] = ad + au;
      }
  }

  // one stage of the driver loop (2Dproblem.cpp:137-158)
  int stage_run(int stage, double dt, const int* bc) {
    for (int s = 0; s < 4; s++) boundary(s, bc[s]);
    primvar();
    if (blowup) return GKS_ERR_BLOWUP;
    reconstruct();
    flux(stage, dt);
    update(stage);
    if (P.use_df && !P.df_frozen) update_alpha();
    return GKS_OK;
  }

  // time loop (2Dproblem.cpp:109-161) without the stdin prompt
  int run(const int* bc, double tstop, int max_steps, double* dt_hist) {
    int n = 0;
    while (t < tstop && n < max_steps) {
      copy_old();
      R dt = cfl(tstop);
      double d = to_double(dt);
      if (dt_hist) dt_hist[n] = d;
      for (int s = 0; s < 2; s++) {
        int rc = stage_run(s, d, bc);
        if (rc) return rc;
      }
      step++;
      t = t + d;
      n++;
    }
    return GKS_OK;
  }
  static double to_double(const R& r);
};

}  // namespace orc
