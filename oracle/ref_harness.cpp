// ref_harness.cpp — drives the compiled reference solver (its own sources,
// built by oracle/Makefile into oracle/_ref/) through the same C API as the
// oracle (orc2d_* <-> ref2d_*), so the tests can check the oracle against the
// reference on identical inputs.  TEST INFRASTRUCTURE ONLY.
//
// It sets the reference's globals the way PlanarShock does
// (2Dproblem.cpp:8-53), allocates with the reference's Set* functions and
// calls the reference's sweeps in the driver's order (2Dproblem.cpp:109-161),
// without the stdin prompt and with std::cout silenced.
#include <cstdint>
#include <cstring>
#include <iostream>

#include "2Dproblem.h"
#include "../include/gks2d.h"
#include "../include/gks1d.h"

namespace {

struct RefCtx {
  gks2d_params p;
  Block2d block;
  Fluid2d* fluids = nullptr;
  Interface2d* xi = nullptr;
  Interface2d* yi = nullptr;
  Flux2d_gauss** xf = nullptr;
  Flux2d_gauss** yf = nullptr;
  Fluid2d bcvalue[4];
};

void set_globals(const gks2d_params* p) {
  K = p->K;
  Gamma = p->gamma;
  Mu = p->mu;
  Nu = p->nu;
  Pr = p->pr;
  c1_euler = p->c1_euler;
  c2_euler = p->c2_euler;
  tau_type = p->tau_type == GKS_TAU_NS ? NS : Euler;
  is_Prandtl_fix = p->prandtl_fix != 0;
  reconstruction_variable = p->recon_variable == GKS_VAR_CHARACTERISTIC ? characteristic : conservative;
  wenotype = wenoz;
  is_reduce_order_warning = false;
  is_using_df_factor = p->use_df != 0;
  g0reconstruction_2D_normal = Center_do_nothing_normal;
  g0reconstruction_2D_tangent = Center_all_collision_multi;
  gks2dsolver = gks2nd_2d;
  flux_function_2d = GKS2D;
  timecoe_list_2d = S2O4_2D;
  if (p->weno_order == 5) {
    cellreconstruction_2D_normal = p->use_df ? WENO5_AO_with_df_normal : WENO5_AO_normal;
    cellreconstruction_2D_tangent = p->use_df ? WENO5_AO_with_df_tangent : WENO5_AO_tangent;
  } else if (p->weno_order == 7) {
    cellreconstruction_2D_normal = WENO7_AO_with_df_normal;
    cellreconstruction_2D_tangent = WENO7_AO_with_df_tangent;
  } else {
    cellreconstruction_2D_normal = WENO9_AO_with_df_normal;
    cellreconstruction_2D_tangent = WENO9_AO_with_df_tangent;
  }
}

BoundaryCondition2d bc_fn(int side, int kind) {
  static const BoundaryCondition2d tab[3][4] = {
      {free_boundary_left, free_boundary_right, free_boundary_down, free_boundary_up},
      {periodic_boundary_left, periodic_boundary_right, periodic_boundary_down, periodic_boundary_up},
      {reflection_boundary_left, reflection_boundary_right, reflection_boundary_down,
       reflection_boundary_up}};
  return tab[kind][side];
}

void silence() {
  static bool done = false;
  if (!done) {
    std::cout.rdbuf(nullptr);
    done = true;
  }
}

}  // namespace

extern "C" {

void* ref2d_create(const gks2d_params* p) {
  silence();
  RefCtx* c = new RefCtx();
  c->p = *p;
  set_globals(p);
  gausspoint = p->gausspoint;
  SetGuassPoint();
  Block2d& b = c->block;
  b.uniform = true;
  b.nodex = p->nodex;
  b.nodey = p->nodey;
  b.ghost = p->ghost;
  b.CFL = p->cfl;
  Initial_stages(b);
  b.nx = b.nodex + 2 * b.ghost;
  b.ny = b.nodey + 2 * b.ghost;
  c->fluids = Setfluid(b);
  c->xi = Setinterface_array(b);
  c->yi = Setinterface_array(b);
  c->xf = Setflux_gauss_array(b);
  c->yf = Setflux_gauss_array(b);
  b.left = p->left;
  b.right = p->right;
  b.down = p->down;
  b.up = p->up;
  b.dx = (b.right - b.left) / b.nodex;
  b.dy = (b.up - b.down) / b.nodey;
  b.overdx = 1 / b.dx;
  b.overdy = 1 / b.dy;
  SetUniformMesh(b, c->fluids, c->xi, c->yi, c->xf, c->yf);
  b.t = 0;
  b.step = 0;
  return c;
}

// the reference never frees its arrays either (2Dproblem.cpp); the harness
// releases the top-level blocks only
void ref2d_destroy(void* h) {
  RefCtx* c = static_cast<RefCtx*>(h);
  delete[] c->fluids;
  delete c;
}

int ref2d_set_prim(void* h, const double* prim) {
  RefCtx* c = static_cast<RefCtx*>(h);
  size_t nc = (size_t)c->block.nx * c->block.ny;
  for (size_t k = 0; k < nc; k++) {
    for (int v = 0; v < 4; v++) c->fluids[k].primvar[v] = prim[v * nc + k];
    Primvar_to_convar_2D(c->fluids[k].convar, c->fluids[k].primvar);
  }
  return GKS_OK;
}
int ref2d_set_convar(void* h, const double* conv) {
  RefCtx* c = static_cast<RefCtx*>(h);
  size_t nc = (size_t)c->block.nx * c->block.ny;
  for (size_t k = 0; k < nc; k++)
    for (int v = 0; v < 4; v++) c->fluids[k].convar[v] = conv[v * nc + k];
  return GKS_OK;
}
int ref2d_primvar(void* h) {
  RefCtx* c = static_cast<RefCtx*>(h);
  set_globals(&c->p);
  Convar_to_Primvar(c->fluids, c->block);
  return GKS_OK;
}
int ref2d_set_alpha(void* h, const double* ax, const double* ay) {
  RefCtx* c = static_cast<RefCtx*>(h);
  size_t nc = (size_t)c->block.nx * c->block.ny;
  for (size_t k = 0; k < nc; k++) {
    c->fluids[k].alpha_x = ax[k];
    c->fluids[k].alpha_y = ay[k];
  }
  return GKS_OK;
}
int ref2d_get(void* h, double* conv, double* prim, double* ax, double* ay, double* t,
              int32_t* step) {
  RefCtx* c = static_cast<RefCtx*>(h);
  size_t nc = (size_t)c->block.nx * c->block.ny;
  for (size_t k = 0; k < nc; k++) {
    for (int v = 0; v < 4; v++) {
      if (conv) conv[v * nc + k] = c->fluids[k].convar[v];
      if (prim) prim[v * nc + k] = c->fluids[k].primvar[v];
    }
    if (ax) ax[k] = c->fluids[k].alpha_x;
    if (ay) ay[k] = c->fluids[k].alpha_y;
  }
  if (t) *t = c->block.t;
  if (step) *step = c->block.step;
  return GKS_OK;
}
int ref2d_set_time(void* h, double t, int32_t step) {
  RefCtx* c = static_cast<RefCtx*>(h);
  c->block.t = t;
  c->block.step = step;
  return GKS_OK;
}

int ref2d_copy_old(void* h) {
  RefCtx* c = static_cast<RefCtx*>(h);
  CopyFluid_new_to_old(c->fluids, c->block);
  return GKS_OK;
}
int ref2d_cfl(void* h, double tstop, double* dt) {
  RefCtx* c = static_cast<RefCtx*>(h);
  *dt = Get_CFL(c->block, c->fluids, tstop);
  return GKS_OK;
}
int ref2d_stage(void* h, int32_t stage, double dt, const int32_t* bc) {
  RefCtx* c = static_cast<RefCtx*>(h);
  set_globals(&c->p);
  Block2d& b = c->block;
  b.dt = dt;
  for (int s = 0; s < 4; s++) bc_fn(s, bc[s])(c->fluids, b, c->bcvalue[s]);
  Convar_to_Primvar(c->fluids, b);
  Reconstruction_within_cell(c->xi, c->yi, c->fluids, b);
  Reconstruction_forg0(c->xi, c->yi, c->fluids, b);
  Calculate_flux(c->xf, c->yf, c->xi, c->yi, b, stage);
  Update(c->fluids, c->xf, c->yf, b, stage);
  // df_frozen: the sensor is never updated (sinwave_2d comments the call
  // out, 2Dproblem.cpp:1514-1519)
  if (is_using_df_factor && !c->p.df_frozen) Update_alpha(c->xi, c->yi, c->fluids, b);
  return GKS_OK;
}
int ref2d_run(void* h, const int32_t* bc, double tstop, int32_t max_steps, double* dt_hist) {
  RefCtx* c = static_cast<RefCtx*>(h);
  Block2d& b = c->block;
  int n = 0;
  while (b.t < tstop && n < max_steps) {
    CopyFluid_new_to_old(c->fluids, b);
    b.dt = Get_CFL(b, c->fluids, tstop);
    if (dt_hist) dt_hist[n] = b.dt;
    for (int s = 0; s < b.stages; s++) ref2d_stage(h, s, b.dt, bc);
    b.step++;
    b.t = b.t + b.dt;
    n++;
  }
  return GKS_OK;
}
int ref2d_get_face_fluxes(void* h, int32_t stage, double* xfo, double* yfo) {
  RefCtx* c = static_cast<RefCtx*>(h);
  int nx = c->block.nx, ny = c->block.ny, gp = gausspoint;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j < ny; j++)
      for (int q = 0; q < gp; q++) {
        Flux2d& f = c->xf[i * (ny + 1) + j][stage].gauss[q];
        double* d = xfo + ((size_t)(i * ny + j) * gp + q) * 8;
        std::memcpy(d, f.f, 32);
        std::memcpy(d + 4, f.derf, 32);
      }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int q = 0; q < gp; q++) {
        Flux2d& f = c->yf[i * (ny + 1) + j][stage].gauss[q];
        double* d = yfo + ((size_t)(i * (ny + 1) + j) * gp + q) * 8;
        std::memcpy(d, f.f, 32);
        std::memcpy(d + 4, f.derf, 32);
      }
  return GKS_OK;
}
static void put_point(double* d, const Point2d& p) {
  std::memcpy(d, p.convar, 32);
  std::memcpy(d + 4, p.der1x, 32);
  std::memcpy(d + 8, p.der1y, 32);
}
int ref2d_get_gauss(void* h, double* xg, double* yg) {
  RefCtx* c = static_cast<RefCtx*>(h);
  int nx = c->block.nx, ny = c->block.ny, gp = gausspoint;
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j < ny; j++)
      for (int q = 0; q < gp; q++) {
        Recon2d& r = c->xi[i * (ny + 1) + j].gauss[q];
        double* d = xg + ((size_t)(i * ny + j) * gp + q) * 24;
        put_point(d, r.left);
        put_point(d + 12, r.right);
      }
  for (int i = 0; i < nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int q = 0; q < gp; q++) {
        Recon2d& r = c->yi[i * (ny + 1) + j].gauss[q];
        double* d = yg + ((size_t)(i * (ny + 1) + j) * gp + q) * 24;
        put_point(d, r.left);
        put_point(d + 12, r.right);
      }
  return GKS_OK;
}

static void get_point(Point2d& p, const double* s) {
  std::memcpy(p.convar, s, 32);
  std::memcpy(p.der1x, s + 4, 32);
  std::memcpy(p.der1y, s + 8, 32);
}
int ref2d_flux_points(const gks2d_params* p, int64_t n, const double* left, const double* right,
                      double dt, double* center, double* f, double* derf) {
  silence();
  set_globals(p);
  for (int64_t k = 0; k < n; k++) {
    Recon2d r;
    get_point(r.left, left + k * 12);
    get_point(r.right, right + k * 12);
    Center_all_collision_2d_multi(r);
    Flux2d fl;
    GKS2D(fl, r, dt);
    if (center) put_point(center + k * 12, r.center);
    std::memcpy(f + k * 4, fl.f, 32);
    std::memcpy(derf + k * 4, fl.derf, 32);
  }
  return GKS_OK;
}
int ref2d_weno5_normal(const gks2d_params* p, int64_t n, const double* w5, const double* alpha5,
                       double h, double* left8, double* right8) {
  silence();
  set_globals(p);
  for (int64_t k = 0; k < n; k++) {
    Point2d L, R;
    double w[5][4], a[5];
    std::memcpy(w, w5 + k * 20, sizeof(w));
    std::memcpy(a, alpha5 + k * 5, sizeof(a));
    WENO5_AO_with_df(L, R, a, w[0], w[1], w[2], w[3], w[4], h);
    std::memcpy(left8 + k * 8, L.convar, 32);
    std::memcpy(left8 + k * 8 + 4, L.der1x, 32);
    std::memcpy(right8 + k * 8, R.convar, 32);
    std::memcpy(right8 + k * 8 + 4, R.der1x, 32);
  }
  return GKS_OK;
}
int ref2d_weno5_tangent(const gks2d_params* p, int64_t n, const double* line5,
                        const double* alpha5, double h, double* gp2) {
  silence();
  set_globals(p);
  gausspoint = 2;
  for (int64_t k = 0; k < n; k++) {
    Recon2d line[5], re[2];
    for (int s = 0; s < 5; s++) {
      std::memcpy(line[s].left.convar, line5 + k * 40 + s * 8, 32);
      std::memcpy(line[s].left.der1x, line5 + k * 40 + s * 8 + 4, 32);
      line[s].right = line[s].left;
    }
    double a1[5], a2[5];
    std::memcpy(a1, alpha5 + k * 5, sizeof(a1));
    std::memcpy(a2, alpha5 + k * 5, sizeof(a2));
    weno_5th_ao_with_df_tangential(re, line[0], line[1], line[2], line[3], line[4], a1, a2, h);
    put_point(gp2 + k * 24, re[0].left);
    put_point(gp2 + k * 24 + 12, re[1].left);
  }
  return GKS_OK;
}

// WENO7_AO_with_df / WENO9_AO_with_df for n cells (layout as orc2d_weno_normal)
int ref2d_weno_normal(const gks2d_params* p, int64_t n, int32_t order, const double* w, const double* alpha,
                      double h, double* left8, double* right8) {
  if (order != 7 && order != 9) return GKS_ERR_ARG;
  silence();
  set_globals(p);
  for (int64_t k = 0; k < n; k++) {
    Point2d L, R;
    double wk[9][4], a[9];
    std::memcpy(wk, w + k * order * 4, order * 32);
    std::memcpy(a, alpha + k * order, order * 8);
    if (order == 7) WENO7_AO_with_df(L, R, a, wk[0], wk[1], wk[2], wk[3], wk[4], wk[5], wk[6], h);
    else WENO9_AO_with_df(L, R, a, wk[0], wk[1], wk[2], wk[3], wk[4], wk[5], wk[6], wk[7], wk[8], h);
    std::memcpy(left8 + k * 8, L.convar, 32);
    std::memcpy(left8 + k * 8 + 4, L.der1x, 32);
    std::memcpy(right8 + k * 8, R.convar, 32);
    std::memcpy(right8 + k * 8 + 4, R.der1x, 32);
  }
  return GKS_OK;
}
// weno_7th/9th_ao_with_df_tangential for n faces (layout as orc2d_weno_tangent)
int ref2d_weno_tangent(const gks2d_params* p, int64_t n, int32_t order, const double* lineL, const double* lineR,
                       const double* a1, const double* a2, double h, double* gpL, double* gpR) {
  if (order != 7 && order != 9) return GKS_ERR_ARG;
  silence();
  set_globals(p);
  const int ng = order == 7 ? 3 : 4;
  gausspoint = ng;
  for (int64_t k = 0; k < n; k++) {
    Recon2d line[9], re[4];
    for (int s = 0; s < order; s++) {
      std::memcpy(line[s].left.convar, lineL + (k * order + s) * 8, 32);
      std::memcpy(line[s].left.der1x, lineL + (k * order + s) * 8 + 4, 32);
      std::memcpy(line[s].right.convar, lineR + (k * order + s) * 8, 32);
      std::memcpy(line[s].right.der1x, lineR + (k * order + s) * 8 + 4, 32);
    }
    double b1[9], b2[9];
    std::memcpy(b1, a1 + k * order, order * 8);
    std::memcpy(b2, a2 + k * order, order * 8);
    if (order == 7)
      weno_7th_ao_with_df_tangential(re, line[0], line[1], line[2], line[3], line[4], line[5], line[6], b1, b2, h);
    else
      weno_9th_ao_with_df_tangential(re, line[0], line[1], line[2], line[3], line[4], line[5], line[6], line[7],
                                     line[8], b1, b2, h);
    for (int q = 0; q < ng; q++) {
      put_point(gpL + (k * ng + q) * 12, re[q].left);
      put_point(gpR + (k * ng + q) * 12, re[q].right);
    }
  }
  return GKS_OK;
}

}  // extern "C"

// =================================================================== 1D
// SodTubeProblem's set-up and loop (1Dproblem.cpp:3-149) with the globals of
// the parameters, boundary kinds as function pointers
namespace {
struct RefCtx1 {
  gks1d_params p;
  Block1d block;
  Fluid1d* fluids = nullptr;
  Interface1d* interfaces = nullptr;
  Flux1d** fluxes = nullptr;
  Fluid1d bcvalue[2];
};
void set_globals1(const gks1d_params* p) {
  K = p->K;
  Gamma = p->gamma;
  Mu = p->mu;
  Nu = p->nu;
  Pr = p->pr;
  c1_euler = p->c1_euler;
  c2_euler = p->c2_euler;
  tau_type = p->tau_type == GKS_TAU_NS ? NS : Euler;
  gks1dsolver = gks2nd;
  wenotype = wenoz;
  reconstruction_variable = p->recon_variable == GKS_VAR_CHARACTERISTIC ? characteristic : conservative;
  g0reconstruction = Center_collision;
  is_reduce_order_warning = false;
  is_using_df_factor = p->use_df != 0;
  flux_function = GKS;
  timecoe_list = S2O4;
}
BoundaryCondition bc1_fn(int side, int kind) {
  static const BoundaryCondition tab[3][2] = {{free_boundary_left, free_boundary_right},
                                             {periodic_boundary_left, periodic_boundary_right},
                                             {reflection_boundary_left, reflection_boundary_right}};
  return tab[kind][side];
}
}  // namespace

extern "C" {

void* ref1d_create(const gks1d_params* p) {
  silence();
  RefCtx1* c = new RefCtx1();
  c->p = *p;
  set_globals1(p);
  Block1d& b = c->block;
  b.nodex = p->nodex;
  b.ghost = p->ghost;
  b.CFL = p->cfl;
  Initial_stages(b);
  b.nx = b.nodex + 2 * b.ghost;
  b.nodex_begin = b.ghost;
  b.nodex_end = b.nodex + b.ghost - 1;
  c->fluids = new Fluid1d[b.nx];
  c->interfaces = new Interface1d[b.nx + 1];
  c->fluxes = Setflux_array(b);
  b.left = p->left;
  b.right = p->right;
  b.dx = (b.right - b.left) / b.nodex;
  SetUniformMesh(b, c->fluids, c->interfaces, c->fluxes);
  b.t = 0;
  b.step = 0;
  return c;
}
void ref1d_destroy(void* h) {
  RefCtx1* c = static_cast<RefCtx1*>(h);
  delete[] c->fluids;
  delete c;
}
int ref1d_set_prim(void* h, const double* prim) {
  RefCtx1* c = static_cast<RefCtx1*>(h);
  int nx = c->block.nx;
  for (int i = 0; i < nx; i++) {
    for (int v = 0; v < 3; v++) c->fluids[i].primvar[v] = prim[v * nx + i];
    Primvar_to_convar_1D(c->fluids[i].convar, c->fluids[i].primvar);
  }
  return GKS_OK;
}
int ref1d_get(void* h, double* conv, double* alpha, double* t, int32_t* step) {
  RefCtx1* c = static_cast<RefCtx1*>(h);
  int nx = c->block.nx;
  for (int i = 0; i < nx; i++) {
    for (int v = 0; v < 3; v++)
      if (conv) conv[v * nx + i] = c->fluids[i].convar[v];
    if (alpha) alpha[i] = c->fluids[i].alpha;
  }
  if (t) *t = c->block.t;
  if (step) *step = c->block.step;
  return GKS_OK;
}
int ref1d_run(void* h, const int32_t* bc, double tstop, int32_t max_steps, double* dt_hist) {
  RefCtx1* c = static_cast<RefCtx1*>(h);
  set_globals1(&c->p);
  Block1d& b = c->block;
  int n = 0;
  while (b.t < tstop && n < max_steps) {
    CopyFluid_new_to_old(c->fluids, b);
    b.dt = Get_CFL(b, c->fluids, tstop);
    if (dt_hist) dt_hist[n] = b.dt;
    if (c->p.df_from_start || (b.step > 0 && is_using_df_factor)) cellreconstruction = WENO5_AO_with_DF;
    else cellreconstruction = WENO5_AO;
    for (int s = 0; s < b.stages; s++) {
      bc1_fn(0, bc[0])(c->fluids, b, c->bcvalue[0]);
      bc1_fn(1, bc[1])(c->fluids, b, c->bcvalue[1]);
      Reconstruction_within_cell(c->interfaces, c->fluids, b);
      Reconstruction_forg0(c->interfaces, c->fluids, b);
      Calculate_flux(c->fluxes, c->interfaces, b, s);
      Update(c->fluids, c->fluxes, b, s);
      if (is_using_df_factor) Update_alpha(c->interfaces, c->fluids, b);
    }
    b.step++;
    b.t = b.t + b.dt;
    n++;
  }
  return GKS_OK;
}
int ref1d_get_face_fluxes(void* h, int32_t stage, double* out) {
  RefCtx1* c = static_cast<RefCtx1*>(h);
  for (int i = 0; i <= c->block.nx; i++) {
    std::memcpy(out + i * 6, c->fluxes[i][stage].f, 24);
    std::memcpy(out + i * 6 + 3, c->fluxes[i][stage].derf, 24);
  }
  return GKS_OK;
}
int ref1d_flux_points(const gks1d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf) {
  silence();
  set_globals1(p);
  for (int64_t k = 0; k < n; k++) {
    Interface1d in;
    std::memcpy(in.left.convar, left + k * 6, 24);
    std::memcpy(in.left.der1, left + k * 6 + 3, 24);
    std::memcpy(in.right.convar, right + k * 6, 24);
    std::memcpy(in.right.der1, right + k * 6 + 3, 24);
    Center_collision(in, nullptr);
    Flux1d fl;
    GKS(fl, in, dt);
    if (center) {
      std::memcpy(center + k * 6, in.center.convar, 24);
      std::memcpy(center + k * 6 + 3, in.center.der1, 24);
    }
    std::memcpy(f + k * 3, fl.f, 24);
    std::memcpy(derf + k * 3, fl.derf, 24);
  }
  return GKS_OK;
}

}  // extern "C"

// =================================================================== 3D
// The reference's 3D leaves, and its 3D driver loop (CubicTube,
// 3Dproblem.cpp:3-166) made runnable by closing the two gaps of the shipped
// path in the harness: the Gauss-point records Interface3d::gauss are
// allocated here (Allocate_reconstruction_variable_along_interface,
// basic_function.cpp:1896-1904, allocates the lines only), and
// flux_function_3d points at a GKS3D that takes lambda from the
// conservative state (the oracle's gks3d_point; GKS3D,
// flux_function.cpp:813-1163, hands the pressure to MMDF3d_left/right, A
// and Get_Tau where lambda belongs).  The boundary functions,
// Convar_to_primvar, the face / line / point reconstructions,
// Calculate_flux, Update_gauss, Get_CFL and the S2O4 coefficients are the
// reference's own code.
#include "3Dproblem.h"
#include "../include/gks3d.h"
#include "solver3d.hpp"

namespace {
orc3d::Gas3 g_gas3;

void set_globals3(const gks3d_params* p) {
  K = p->K;
  Gamma = p->gamma;
  Mu = p->mu;
  Nu = p->nu;
  c1_euler = p->c1_euler;
  c2_euler = p->c2_euler;
  tau_type = p->tau_type == GKS_TAU_NS ? NS : Euler;
  reconstruction_variable = p->recon_variable == GKS_VAR_CHARACTERISTIC ? characteristic : conservative;
  wenotype = wenoz;
  is_reduce_order_warning = false;
  is_multi_dimensional = true;
  gks3dsolver = gks2nd_3d;
  g0type = all_collisionn;
  timecoe_list_3d = S2O4_3D;
  cellreconstruction_3D_of_face_value = WENO5_AO_splitting_3d;
  cellreconstruction_3D_of_line_value = WENO5_AO_of_line_value;
  cellreconstruction_3D_of_point_value = WENO5_AO_of_point_value;
  g0reconstruction_3D_of_face_value = Do_nothing_splitting_3d;
  g0reconstruction_3D_of_line_value = Do_nothing_reconstruction_of_line_value;
  g0reconstruction_3D_of_point_value = Do_nothing_reconstruction_of_point_value;
  g_gas3 = orc3d::Gas3{p->K, p->gamma, p->mu, p->nu, p->c1_euler, p->c2_euler, p->tau_type};
}

void pack_point(double* o, const Point3d& pt) {
  std::memcpy(o, pt.convar, 40);
  std::memcpy(o + 5, pt.der1x, 40);
  std::memcpy(o + 10, pt.der1y, 40);
  std::memcpy(o + 15, pt.der1z, 40);
}
void unpack_point(Point3d& pt, const double* s) {
  std::memcpy(pt.convar, s, 40);
  std::memcpy(pt.der1x, s + 5, 40);
  std::memcpy(pt.der1y, s + 10, 40);
  std::memcpy(pt.der1z, s + 15, 40);
}

// the flux the harness gives flux_function_3d
void gks3d_lambda_fixed(Flux3d& flux, Recon3d& re, double area, double dt) {
  double L[20], R[20];
  pack_point(L, re.left);
  pack_point(R, re.right);
  orc3d::gks3d_point(flux.f, flux.derf, static_cast<double*>(nullptr), L, R, dt, g_gas3);
}

struct RefCtx3 {
  gks3d_params p;
  Block3d block;
  Fluid3d* fluids = nullptr;
  Interface3d* xi = nullptr;
  Interface3d* yi = nullptr;
  Interface3d* zi = nullptr;
  Flux3d_gauss** xf = nullptr;
  Flux3d_gauss** yf = nullptr;
  Flux3d_gauss** zf = nullptr;
  Fluid3d bcvalue[6];
};
}  // namespace

extern "C" {

void* ref3d_create(const gks3d_params* p) {
  silence();
  RefCtx3* c = new RefCtx3();
  c->p = *p;
  set_globals3(p);
  flux_function_3d = gks3d_lambda_fixed;
  gausspoint = 4;
  SetGuassPoint_3D();
  Block3d& b = c->block;
  b.uniform = true;
  b.nodex = p->nodex;
  b.nodey = p->nodey;
  b.nodez = p->nodez;
  b.ghost = p->ghost;
  b.CFL = p->cfl;
  Initial_stages(b);
  b.nx = b.nodex + 2 * b.ghost;
  b.ny = b.nodey + 2 * b.ghost;
  b.nz = b.nodez + 2 * b.ghost;
  c->fluids = Setfluid_array(b);
  c->xi = Setinterface_array(b);
  c->yi = Setinterface_array(b);
  c->zi = Setinterface_array(b);
  c->xf = Setflux_gauss_array(b);
  c->yf = Setflux_gauss_array(b);
  c->zf = Setflux_gauss_array(b);
  b.left = p->left;
  b.right = p->right;
  b.down = p->down;
  b.up = p->up;
  b.back = p->back;
  b.front = p->front;
  b.dx = (b.right - b.left) / b.nodex;
  b.dy = (b.up - b.down) / b.nodey;
  b.dz = (b.front - b.back) / b.nodez;
  b.xarea = b.dy * b.dz;
  b.yarea = b.dx * b.dz;
  b.zarea = b.dx * b.dy;
  b.overdx = 1 / b.dx;
  b.overdy = 1 / b.dy;
  b.overdz = 1 / b.dz;
  SetUniformMesh(b, c->fluids, c->xi, c->yi, c->zi);
  const size_t nf = (size_t)(b.nx + 1) * (b.ny + 1) * (b.nz + 1);
  for (Interface3d* arr : {c->xi, c->yi, c->zi})
    for (size_t f = 0; f < nf; f++) {
      arr[f].gauss = new Recon3d[gausspoint];
      for (int q = 0; q < gausspoint; q++) arr[f].gauss[q].order_reduce = false;
    }
  b.t = 0;
  b.step = 0;
  return c;
}
// the top-level blocks only, as ref2d_destroy
void ref3d_destroy(void* h) {
  RefCtx3* c = static_cast<RefCtx3*>(h);
  delete[] c->fluids;
  delete c;
}
// convar and the primitive variables Get_CFL reads first, [5][nc] each
int ref3d_set_state(void* h, const double* conv, const double* prim) {
  RefCtx3* c = static_cast<RefCtx3*>(h);
  const Block3d& b = c->block;
  size_t nc = (size_t)b.nx * b.ny * b.nz;
  for (size_t k = 0; k < nc; k++)
    for (int v = 0; v < 5; v++) {
      c->fluids[k].convar[v] = conv[v * nc + k];
      c->fluids[k].primvar[v] = prim[v * nc + k];
    }
  return GKS_OK;
}
int ref3d_get(void* h, double* conv, double* t, int32_t* step) {
  RefCtx3* c = static_cast<RefCtx3*>(h);
  const Block3d& b = c->block;
  size_t nc = (size_t)b.nx * b.ny * b.nz;
  for (size_t k = 0; k < nc; k++)
    for (int v = 0; v < 5; v++) conv[v * nc + k] = c->fluids[k].convar[v];
  if (t) *t = b.t;
  if (step) *step = b.step;
  return GKS_OK;
}
// the driver loop (3Dproblem.cpp:109-166) with the free boundaries, which
// are the only 3D boundary functions the reference has
int ref3d_run(void* h, double tstop, int32_t max_steps, double* dt_hist) {
  RefCtx3* c = static_cast<RefCtx3*>(h);
  set_globals3(&c->p);
  flux_function_3d = gks3d_lambda_fixed;
  gausspoint = 4;
  Block3d& b = c->block;
  int n = 0;
  while (b.t < tstop && n < max_steps) {
    CopyFluid_new_to_old(c->fluids, b);
    b.dt = Get_CFL(b, c->fluids, tstop);
    if (dt_hist) dt_hist[n] = b.dt;
    for (int s = 0; s < b.stages; s++) {
      free_boundary_xleft(c->fluids, b, c->bcvalue[0]);
      free_boundary_xright(c->fluids, b, c->bcvalue[1]);
      free_boundary_yleft(c->fluids, b, c->bcvalue[2]);
      free_boundary_yright(c->fluids, b, c->bcvalue[3]);
      free_boundary_zleft(c->fluids, b, c->bcvalue[4]);
      free_boundary_zright(c->fluids, b, c->bcvalue[5]);
      Convar_to_primvar(c->fluids, b);
      Reconstruction_within_cell(c->xi, c->yi, c->zi, c->fluids, b);
      Reconstruction_forg0(c->xi, c->yi, c->zi, c->fluids, b);
      Calculate_flux(c->xf, c->yf, c->zf, c->xi, c->yi, c->zi, b, s);
      Update_gauss(c->fluids, c->xf, c->yf, c->zf, b, s);
    }
    b.step++;
    b.t = b.t + b.dt;
    n++;
  }
  return GKS_OK;
}

// WENO5_AO (reconstruction.cpp:3803-3923) on n stencils w[n][5][5]; lo, ro:
// n x 10 (convar, der1x) at the cell's left / right face
int ref3d_weno_face(const gks3d_params* p, int64_t n, const double* w, double h, double* lo, double* ro) {
  silence();
  set_globals3(p);
  for (int64_t k = 0; k < n; k++) {
    double s[5][5];
    std::memcpy(s, w + k * 25, sizeof(s));
    Point3d l, r;
    WENO5_AO(l, r, s[0], s[1], s[2], s[3], s[4], h);
    std::memcpy(lo + k * 10, l.convar, 40);
    std::memcpy(lo + k * 10 + 5, l.der1x, 40);
    std::memcpy(ro + k * 10, r.convar, 40);
    std::memcpy(ro + k * 10 + 5, r.der1x, 40);
  }
  return GKS_OK;
}
// WENO5_AO_for_line_value (nc 10: convar, der1x -> + der1y) or
// WENO5_AO_for_point_value (nc 15: convar, der1x, der1y -> + der1z)
// (reconstruction.cpp:3955-4145, 4163-4399) on n stencils st[n][5][nc],
// given to both sides; out[n][2][nc + 5] the left side's two Gauss points
int ref3d_weno_tangential(const gks3d_params* p, int64_t n, int32_t nc, const double* st, double h, double* out) {
  silence();
  set_globals3(p);
  gausspoint = 4;
  if (nc != 10 && nc != 15) return GKS_ERR_ARG;
  for (int64_t k = 0; k < n; k++) {
    Recon3d s[5], res[2];
    for (int q = 0; q < 5; q++) {
      double rec[20] = {0.0};
      std::memcpy(rec, st + (k * 5 + q) * nc, sizeof(double) * nc);
      unpack_point(s[q].left, rec);
      unpack_point(s[q].right, rec);
    }
    res[0].order_reduce = false;
    res[1].order_reduce = false;
    if (nc == 10) WENO5_AO_for_line_value(res, s[0], s[1], s[2], s[3], s[4], h);
    else WENO5_AO_for_point_value(res, s[0], s[1], s[2], s[3], s[4], h);
    for (int g = 0; g < 2; g++) {
      double rec[20];
      pack_point(rec, res[g].left);
      double* o = out + (k * 2 + g) * (nc + 5);
      std::memcpy(o, rec, sizeof(double) * (nc + 5));
    }
  }
  return GKS_OK;
}
// the reference's own GKS3D (flux_function.cpp:813-1163) at n Gauss points;
// it is right only where lambda equals the pressure in the left, right and
// collision states (p = sqrt(rho (K+3) (gamma-1) / 4))
int ref3d_gks3d(const gks3d_params* p, int64_t n, const double* left, const double* right, double dt, double* f,
                double* derf) {
  silence();
  set_globals3(p);
  for (int64_t k = 0; k < n; k++) {
    Recon3d re;
    unpack_point(re.left, left + k * 20);
    unpack_point(re.right, right + k * 20);
    Flux3d fl;
    GKS3D(fl, re, 1.0, dt);
    std::memcpy(f + k * 5, fl.f, 40);
    std::memcpy(derf + k * 5, fl.derf, 40);
  }
  return GKS_OK;
}

}  // extern "C"
