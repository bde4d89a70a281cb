// gks_oracle.cpp — C API over the CPU restatement (oracle).  TEST
// INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; the product never links it.
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "solver1d.hpp"
#include "solver2d.hpp"
#include "solver3d.hpp"
#include "../include/gks1d.h"
#include "../include/gks3d.h"

namespace orc {
template <> double Solver2d<double>::to_double(const double& r) { return r; }
}  // namespace orc

using S2 = orc::Solver2d<double>;

static bool params_ok(const gks2d_params* p) {
  if (!p) return false;
  if (p->nodex < 5 || p->nodey < 5 || p->ghost != 5) return false;
  const bool ok = (p->weno_order == 5 && p->gausspoint == 2) || (p->weno_order == 7 && p->gausspoint == 3) ||
                  (p->weno_order == 9 && p->gausspoint == 4);
  return ok;
}

extern "C" {

// sizeof(gks2d_params) and the byte offset of each field, in declaration
// order (checks the ctypes mirror against the C compiler)
int64_t orc2d_params_layout(int64_t* off) {
  const gks2d_params* z = nullptr;
  const char* b = reinterpret_cast<const char*>(z);
  int k = 0;
#define OFF(f) off[k++] = reinterpret_cast<const char*>(&z->f) - b
  OFF(nodex); OFF(nodey); OFF(ghost); OFF(gausspoint); OFF(weno_order); OFF(recon_variable);
  OFF(use_df); OFF(tau_type); OFF(K); OFF(prandtl_fix); OFF(gamma); OFF(mu); OFF(nu); OFF(pr);
  OFF(c1_euler); OFF(c2_euler); OFF(left); OFF(right); OFF(down); OFF(up); OFF(cfl);
  OFF(fast_math); OFF(df_frozen);
#undef OFF
  return (int64_t)sizeof(gks2d_params);
}

void* orc2d_create(const gks2d_params* p) {
  if (!params_ok(p)) return nullptr;
  return new S2(*p);
}
void orc2d_destroy(void* h) { delete static_cast<S2*>(h); }

// Sets primvar of all nx*ny cells (SoA [4][nx*ny]) and convar =
// Primvar_to_convar_2D(primvar) for every cell, as IC_for_riemann_2d does
// (2Dproblem.cpp:345-385).
int orc2d_set_prim(void* h, const double* prim) {
  S2* s = static_cast<S2*>(h);
  size_t nc = (size_t)s->nx * s->ny;
  for (size_t c = 0; c < nc; c++) {
    double p[4], u[4];
    for (int v = 0; v < 4; v++) p[v] = prim[v * nc + c];
    orc::primvar_to_convar_2d(u, p, s->g);
    for (int v = 0; v < 4; v++) {
      s->prim[v][c] = p[v];
      s->U[v][c] = u[v];
    }
  }
  return GKS_OK;
}
int orc2d_set_convar(void* h, const double* conv) {
  S2* s = static_cast<S2*>(h);
  size_t nc = (size_t)s->nx * s->ny;
  for (int v = 0; v < 4; v++) std::memcpy(s->U[v].data(), conv + v * nc, nc * 8);
  return GKS_OK;
}
// primvar of every cell from the current convar (Convar_to_Primvar,
// basic_function.cpp:752-773), for starts from conservative cell averages
int orc2d_primvar(void* h) {
  static_cast<S2*>(h)->primvar();
  return GKS_OK;
}
int orc2d_set_alpha(void* h, const double* ax, const double* ay) {
  S2* s = static_cast<S2*>(h);
  size_t nc = (size_t)s->nx * s->ny;
  std::memcpy(s->ax.data(), ax, nc * 8);
  std::memcpy(s->ay.data(), ay, nc * 8);
  return GKS_OK;
}
int orc2d_get(void* h, double* conv, double* prim, double* ax, double* ay, double* t,
              int32_t* step) {
  S2* s = static_cast<S2*>(h);
  size_t nc = (size_t)s->nx * s->ny;
  for (int v = 0; v < 4; v++) {
    if (conv) std::memcpy(conv + v * nc, s->U[v].data(), nc * 8);
    if (prim) std::memcpy(prim + v * nc, s->prim[v].data(), nc * 8);
  }
  if (ax) std::memcpy(ax, s->ax.data(), nc * 8);
  if (ay) std::memcpy(ay, s->ay.data(), nc * 8);
  if (t) *t = s->t;
  if (step) *step = s->step;
  return GKS_OK;
}
int orc2d_set_time(void* h, double t, int32_t step) {
  S2* s = static_cast<S2*>(h);
  s->t = t;
  s->step = step;
  return GKS_OK;
}
int orc2d_run(void* h, const int32_t* bc, double tstop, int32_t max_steps, double* dt_hist) {
  S2* s = static_cast<S2*>(h);
  int b[4] = {bc[0], bc[1], bc[2], bc[3]};
  return s->run(b, tstop, max_steps, dt_hist);
}
int orc2d_copy_old(void* h) {
  static_cast<S2*>(h)->copy_old();
  return GKS_OK;
}
int orc2d_cfl(void* h, double tstop, double* dt) {
  *dt = static_cast<S2*>(h)->cfl(tstop);
  return GKS_OK;
}
int orc2d_stage(void* h, int32_t stage, double dt, const int32_t* bc) {
  S2* s = static_cast<S2*>(h);
  int b[4] = {bc[0], bc[1], bc[2], bc[3]};
  return s->stage_run(stage, dt, b);
}
// fluxes of `stage`: xf (nx+1)*ny faces, yf nx*(ny+1) faces, gp*8 doubles each
int orc2d_get_face_fluxes(void* h, int32_t stage, double* xf, double* yf) {
  S2* s = static_cast<S2*>(h);
  int gp = s->gp;
  for (int i = 0; i <= s->nx; i++)
    for (int j = 0; j < s->ny; j++)
      std::memcpy(xf + ((size_t)(i * s->ny + j) * gp) * 8,
                  &s->xfl[stage][(size_t)s->face(i, j) * gp * 8], gp * 8 * 8);
  for (int i = 0; i < s->nx; i++)
    for (int j = 0; j <= s->ny; j++)
      std::memcpy(yf + ((size_t)(i * (s->ny + 1) + j) * gp) * 8,
                  &s->yfl[stage][(size_t)s->face(i, j) * gp * 8], gp * 8 * 8);
  return GKS_OK;
}
// Gauss-point states: xg (nx+1)*ny faces, yg nx*(ny+1) faces, each
// gp x [left 12, right 12]
int orc2d_get_gauss(void* h, double* xg, double* yg) {
  S2* s = static_cast<S2*>(h);
  int gp = s->gp;
  auto put = [&](double* dst, const orc::Pt<double>& a) {
    std::memcpy(dst, a.convar, 32);
    std::memcpy(dst + 4, a.der1x, 32);
    std::memcpy(dst + 8, a.der1y, 32);
  };
  for (int i = 0; i <= s->nx; i++)
    for (int j = 0; j < s->ny; j++)
      for (int q = 0; q < gp; q++) {
        size_t k = (size_t)s->face(i, j) * gp + q;
        double* d = xg + (((size_t)(i * s->ny + j) * gp + q) * 24);
        put(d, s->xgL[k]);
        put(d + 12, s->xgR[k]);
      }
  for (int i = 0; i < s->nx; i++)
    for (int j = 0; j <= s->ny; j++)
      for (int q = 0; q < gp; q++) {
        size_t k = (size_t)s->face(i, j) * gp + q;
        double* d = yg + (((size_t)(i * (s->ny + 1) + j) * gp + q) * 24);
        put(d, s->ygL[k]);
        put(d + 12, s->ygR[k]);
      }
  return GKS_OK;
}

// ------------------------------------------------------------ leaf entries
static void load_pt(orc::Pt<double>& p, const double* s) {
  std::memcpy(p.convar, s, 32);
  std::memcpy(p.der1x, s + 4, 32);
  std::memcpy(p.der1y, s + 8, 32);
}
static void store_pt(double* d, const orc::Pt<double>& p) {
  std::memcpy(d, p.convar, 32);
  std::memcpy(d + 4, p.der1x, 32);
  std::memcpy(d + 8, p.der1y, 32);
}
// centre state + GKS2D for n points (left/right/center n x 12, f/derf n x 4)
int orc2d_flux_points(const gks2d_params* p, int64_t n, const double* left, const double* right,
                      double dt, double* center, double* f, double* derf) {
  orc::Gas g = orc::gas_from(*p);
#pragma omp parallel for
  for (int64_t k = 0; k < n; k++) {
    orc::Pt<double> L, R, C;
    orc::MM2<double> ml, mr;
    load_pt(L, left + k * 12);
    load_pt(R, right + k * 12);
    orc::center_collision(C, ml, mr, L, R, g);
    if (center) store_pt(center + k * 12, C);
    orc::gks2d_flux(f + k * 4, derf + k * 4, L, R, C, ml, mr, dt, g);
  }
  return GKS_OK;
}
// WENO5_AO_with_df for one cell: w5 = 5 x 4 convar (x-like frame), alpha5,
// out: left8 / right8 = convar[4], der1x[4] at the cell's left/right face
int orc2d_weno5_normal(const gks2d_params* p, int64_t n, const double* w5, const double* alpha5,
                       double h, double* left8, double* right8) {
  orc::Gas g = orc::gas_from(*p);
  for (int64_t k = 0; k < n; k++) {
    orc::Pt<double> L, R;
    const double* w = w5 + k * 20;
    orc::weno5_df_normal(L, R, alpha5 + k * 5, w, w + 4, w + 8, w + 12, w + 16, h, p->recon_variable,
                         g);
    std::memcpy(left8 + k * 8, L.convar, 32);
    std::memcpy(left8 + k * 8 + 4, L.der1x, 32);
    std::memcpy(right8 + k * 8, R.convar, 32);
    std::memcpy(right8 + k * 8 + 4, R.der1x, 32);
  }
  return GKS_OK;
}
// weno_5th_ao_with_df_tangential, one side: line5 = 5 x 8 (convar, der1x),
// alpha5; out gp2 = 2 x 12 Gauss-point states
int orc2d_weno5_tangent(const gks2d_params* p, int64_t n, const double* line5,
                        const double* alpha5, double h, double* gp2) {
  orc::Gas g = orc::gas_from(*p);
  for (int64_t k = 0; k < n; k++) {
    orc::Pt<double> line[5], out[2];
    for (int s = 0; s < 5; s++) {
      std::memcpy(line[s].convar, line5 + k * 40 + s * 8, 32);
      std::memcpy(line[s].der1x, line5 + k * 40 + s * 8 + 4, 32);
      for (int v = 0; v < 4; v++) line[s].der1y[v] = 0.0;
    }
    orc::weno5_df_tangent_side(out, line, alpha5 + k * 5, h, p->recon_variable, g);
    store_pt(gp2 + k * 24, out[0]);
    store_pt(gp2 + k * 24 + 12, out[1]);
  }
  return GKS_OK;
}

// WENO7/9_AO_with_df for n cells: w (n x (order) x 4, x-like frame), alpha
// (n x order) -> left/right line states (n x 8: convar, der1x)
int orc2d_weno_normal(const gks2d_params* p, int64_t n, int32_t order, const double* w, const double* alpha,
                      double h, double* left8, double* right8) {
  if (order != 7 && order != 9) return GKS_ERR_ARG;
  orc::Gas g = orc::gas_from(*p);
  for (int64_t k = 0; k < n; k++) {
    const double* wk = w + k * order * 4;
    const double* wv[9];
    for (int s = 0; s < order; s++) wv[s] = wk + s * 4;
    orc::Pt<double> L, R;
    const int r = (order - 1) / 2;
    if (order == 7) orc::weno79_normal<7>(L, R, alpha + k * order + r, wv + r, h, p->recon_variable, g);
    else orc::weno79_normal<9>(L, R, alpha + k * order + r, wv + r, h, p->recon_variable, g);
    std::memcpy(left8 + k * 8, L.convar, 32);
    std::memcpy(left8 + k * 8 + 4, L.der1x, 32);
    std::memcpy(right8 + k * 8, R.convar, 32);
    std::memcpy(right8 + k * 8 + 4, R.der1x, 32);
  }
  return GKS_OK;
}
// weno_7th/9th_ao_with_df_tangential for n faces: lineL/lineR (n x order x
// 8: convar, der1x), a1/a2 (n x order) -> gpL/gpR (n x NG x 12)
int orc2d_weno_tangent(const gks2d_params* p, int64_t n, int32_t order, const double* lineL, const double* lineR,
                       const double* a1, const double* a2, double h, double* gpL, double* gpR) {
  if (order != 7 && order != 9) return GKS_ERR_ARG;
  orc::Gas g = orc::gas_from(*p);
  const int ng = order == 7 ? 3 : 4;
  for (int64_t k = 0; k < n; k++) {
    orc::Pt<double> L[9], R[9], oL[4], oR[4];
    for (int s = 0; s < order; s++) {
      std::memcpy(L[s].convar, lineL + (k * order + s) * 8, 32);
      std::memcpy(L[s].der1x, lineL + (k * order + s) * 8 + 4, 32);
      std::memcpy(R[s].convar, lineR + (k * order + s) * 8, 32);
      std::memcpy(R[s].der1x, lineR + (k * order + s) * 8 + 4, 32);
      for (int v = 0; v < 4; v++) L[s].der1y[v] = R[s].der1y[v] = 0.0;
    }
    if (order == 7)
      orc::weno79_tangent_face<7>(oL, oR, L, R, a1 + k * order, a2 + k * order, h, p->recon_variable, g);
    else
      orc::weno79_tangent_face<9>(oL, oR, L, R, a1 + k * order, a2 + k * order, h, p->recon_variable, g);
    for (int q = 0; q < ng; q++) {
      store_pt(gpL + (k * ng + q) * 12, oL[q]);
      store_pt(gpR + (k * ng + q) * 12, oR[q]);
    }
  }
  return GKS_OK;
}

}  // extern "C"

// =================================================================== 1D
using S1 = orc::Solver1d<double>;

static orc::Gas gas1_from(const gks1d_params& p) {
  orc::Gas g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.pr = p.pr;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  g.prandtl_fix = 0;
  return g;
}

extern "C" {

int64_t orc1d_params_layout(int64_t* off) {
  const gks1d_params* z = nullptr;
  const char* b = reinterpret_cast<const char*>(z);
  int k = 0;
#define OFF(f) off[k++] = reinterpret_cast<const char*>(&z->f) - b
  OFF(nodex); OFF(ghost); OFF(recon_variable); OFF(use_df); OFF(tau_type); OFF(K); OFF(fast_math);
  OFF(df_from_start); OFF(gamma); OFF(mu); OFF(nu); OFF(pr); OFF(c1_euler); OFF(c2_euler); OFF(left);
  OFF(right); OFF(cfl);
#undef OFF
  return (int64_t)sizeof(gks1d_params);
}

void* orc1d_create(const gks1d_params* p) {
  if (!p || p->ghost != 5 || p->nodex < 5) return nullptr;
  return new S1(p->nodex, p->ghost, p->left, p->right, p->cfl, p->recon_variable, p->use_df,
                p->df_from_start, gas1_from(*p));
}
void orc1d_destroy(void* h) { delete static_cast<S1*>(h); }
// ICforSodTube-style set-up: primvar of all nx cells ([3][nx]), convar =
// Primvar_to_convar_1D for every cell
int orc1d_set_prim(void* h, const double* prim) {
  S1* s = static_cast<S1*>(h);
  for (int i = 0; i < s->nx; i++) {
    double p[3] = {prim[i], prim[s->nx + i], prim[2 * s->nx + i]}, u[3];
    orc::primvar_to_convar_1d(u, p, s->g);
    for (int v = 0; v < 3; v++) s->U[v][i] = u[v];
  }
  return GKS_OK;
}
int orc1d_get(void* h, double* conv, double* alpha, double* t, int32_t* step) {
  S1* s = static_cast<S1*>(h);
  for (int v = 0; v < 3; v++)
    if (conv) std::memcpy(conv + v * s->nx, s->U[v].data(), s->nx * 8);
  if (alpha) std::memcpy(alpha, s->alpha.data(), s->nx * 8);
  if (t) *t = s->t;
  if (step) *step = s->step;
  return GKS_OK;
}
int orc1d_run(void* h, const int32_t* bc, double tstop, int32_t max_steps, double* dt_hist) {
  int b[2] = {bc[0], bc[1]};
  return static_cast<S1*>(h)->run(b, tstop, max_steps, dt_hist);
}
// face fluxes of `stage`: (nx+1) x 6 (f[3], derf[3])
int orc1d_get_face_fluxes(void* h, int32_t stage, double* out) {
  S1* s = static_cast<S1*>(h);
  std::memcpy(out, s->fl[stage].data(), (size_t)(s->nx + 1) * 6 * 8);
  return GKS_OK;
}
int orc1d_flux_points(const gks1d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  orc::Gas g = gas1_from(*p);
  for (int64_t k = 0; k < n; k++) {
    orc::Pt1<double> L, R, C;
    std::memcpy(L.convar, left + k * 6, 24);
    std::memcpy(L.der1, left + k * 6 + 3, 24);
    std::memcpy(R.convar, right + k * 6, 24);
    std::memcpy(R.der1, right + k * 6 + 3, 24);
    orc::center_collision_1d(C, L, R, g);
    if (center) {
      std::memcpy(center + k * 6, C.convar, 24);
      std::memcpy(center + k * 6 + 3, C.der1, 24);
    }
    orc::gks1d_flux(f + k * 3, derf + k * 3, L, R, C, dt, g);
  }
  return GKS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ 3D
namespace {
orc3d::Gas3 gas3_from(const gks3d_params& p) {
  orc3d::Gas3 g;
  g.K = p.K;
  g.gamma = p.gamma;
  g.mu = p.mu;
  g.nu = p.nu;
  g.c1 = p.c1_euler;
  g.c2 = p.c2_euler;
  g.tau_type = p.tau_type;
  return g;
}
}  // namespace

extern "C" {

int64_t orc3d_params_layout(int64_t* off) {
  int k = 0;
#define OFF3(f) off[k++] = (int64_t)offsetof(gks3d_params, f)
  OFF3(nodex); OFF3(nodey); OFF3(nodez); OFF3(ghost); OFF3(recon_variable); OFF3(tau_type); OFF3(K);
  OFF3(fast_math); OFF3(gamma); OFF3(mu); OFF3(nu); OFF3(c1_euler); OFF3(c2_euler); OFF3(left); OFF3(right);
  OFF3(down); OFF3(up); OFF3(back); OFF3(front); OFF3(cfl);
#undef OFF3
  return (int64_t)sizeof(gks3d_params);
}

void* orc3d_create(const gks3d_params* p) {
  if (!p || p->ghost < 3 || p->nodex < 3 || p->nodey < 3 || p->nodez < 3) return nullptr;
  orc3d::Params3 P;
  P.NX = p->nodex;
  P.NY = p->nodey;
  P.NZ = p->nodez;
  P.G = p->ghost;
  P.recon_var = p->recon_variable;
  P.dx = (p->right - p->left) / p->nodex;
  P.dy = (p->up - p->down) / p->nodey;
  P.dz = (p->front - p->back) / p->nodez;
  P.cfl = p->cfl;
  P.gas = gas3_from(*p);
  return new orc3d::Solver3d(P);
}
void orc3d_destroy(void* h) { delete static_cast<orc3d::Solver3d*>(h); }
// SoA [5][nc] conservative state, all cells
int orc3d_set_convar(void* h, const double* conv) {
  auto* s = static_cast<orc3d::Solver3d*>(h);
  std::memcpy(s->U.data(), conv, s->U.size() * 8);
  return GKS_OK;
}
int orc3d_set_cfl_primvar(void* h, const double* prim) {
  static_cast<orc3d::Solver3d*>(h)->cfl_from_prim(prim);
  return GKS_OK;
}
int orc3d_get(void* h, double* conv, double* t, int32_t* step) {
  auto* s = static_cast<orc3d::Solver3d*>(h);
  if (conv) std::memcpy(conv, s->U.data(), s->U.size() * 8);
  if (t) *t = s->t;
  if (step) *step = s->step;
  return GKS_OK;
}
int orc3d_run(void* h, const int32_t* bc, double tstop, int32_t max_steps, double* dt_hist) {
  int b[6];
  for (int k = 0; k < 6; k++) b[k] = bc[k];
  return static_cast<orc3d::Solver3d*>(h)->run(b, tstop, max_steps, dt_hist);
}
int orc3d_flux_points(const gks3d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  const orc3d::Gas3 g = gas3_from(*p);
  for (int64_t k = 0; k < n; k++)
    orc3d::gks3d_point(f + 5 * k, derf + 5 * k, center ? center + 20 * k : nullptr, left + 20 * k, right + 20 * k,
                       dt, g);
  return GKS_OK;
}
// leaves: face values of n 5-cell stencils (n x 5 x 5, face frame) ->
// lo, ro (n x 10 each); tangential side of n stencils of records with nc
// components (10: face records, 15: line records) -> n x 2 x (nc + 5)
int orc3d_weno_face(const gks3d_params* p, int64_t n, const double* w, double h, double* lo, double* ro) {
  const orc3d::Gas3 g = gas3_from(*p);
  for (int64_t k = 0; k < n; k++)
    orc3d::weno5_face(lo + 10 * k, ro + 10 * k, reinterpret_cast<const double(*)[5]>(w + 25 * k), h,
                      p->recon_variable, g);
  return GKS_OK;
}
int orc3d_weno_tangential(const gks3d_params* p, int64_t n, int32_t nc, const double* st, double h, double* out) {
  const orc3d::Gas3 g = gas3_from(*p);
  for (int64_t k = 0; k < n; k++) {
    if (nc == 10)
      orc3d::weno5_tangential_side<10>(reinterpret_cast<double(*)[15]>(out + 30 * k),
                                      reinterpret_cast<const double(*)[10]>(st + 50 * k), h, p->recon_variable, g);
    else
      orc3d::weno5_tangential_side<15>(reinterpret_cast<double(*)[20]>(out + 40 * k),
                                      reinterpret_cast<const double(*)[15]>(st + 75 * k), h, p->recon_variable, g);
  }
  return GKS_OK;
}

}  // extern "C"
