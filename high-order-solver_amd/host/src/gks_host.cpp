// gks_host.cpp — the reference's 2D call surface over the gfx950 C ABI.
// Each sweep function binds the cell array it is given to a device context
// (created at first use from a snapshot of the scheme globals) and enqueues
// the matching device work; see gks_host.h for the residency rules.
#include "gks_host.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <vector>

#include "../../../include/gks2d.h"

// ------------------------------------------------------------- the globals
// defaults as the reference initialises them (basic_function.cpp:3-13,
// fluid_mesh.cpp:3-5, reconstruction.cpp:4-11, 1408-1412,
// flux_function.cpp:313-314, time_advance.cpp:198)
int K = 0;
double Gamma = 2.0 / (K + 2) + 1.0;
double T_inf = 285;
double R_gas = 8.31441 / (28.959e-3);
double Mu = -1.0;
double Nu = -1.0;
double c1_euler = -1.0;
double c2_euler = -1.0;
bool is_Prandtl_fix = false;
double Pr = 1.0;
TAU_TYPE tau_type = Euler;

int gausspoint = 0;
double* gauss_loc = nullptr;
double* gauss_weight = nullptr;

Reconstruction_variable reconstruction_variable = conservative;
WENOtype wenotype = wenojs;
bool is_reduce_order_warning = false;
bool is_using_df_factor = false;
Reconstruction_within_Cell_2D_normal cellreconstruction_2D_normal = WENO5_AO_normal;
Reconstruction_within_Cell_2D_tangent cellreconstruction_2D_tangent = WENO5_AO_tangent;
Reconstruction_forG0_2D_normal g0reconstruction_2D_normal = Center_do_nothing_normal;
Reconstruction_forG0_2D_tangent g0reconstruction_2D_tangent = Center_all_collision_multi;

GKS2d_type gks2dsolver = kfvs1st_2d;
Flux_function_2d flux_function_2d = GKS2D;
TimeMarchingCoefficient_2d timecoe_list_2d = nullptr;

namespace {

struct Bound {
  gks2d_ctx* ctx = nullptr;
};
std::map<const void*, Fluid2d*> g_owner;  // interface / flux arrays -> cell array
std::map<const Fluid2d*, Bound> g_bound;

[[noreturn]] void die(const std::string& where, const std::string& what) {
  cout << "gks: " << where << ": " << what << endl;
  exit(0);
}
void check(int rc, const char* where) {
  if (rc != GKS_OK) die(where, gks2d_last_error());
}

gks2d_params snapshot(const Block2d& b) {
  gks2d_params p;
  std::memset(&p, 0, sizeof(p));
  p.nodex = b.nodex;
  p.nodey = b.nodey;
  p.ghost = b.ghost;
  p.gausspoint = gausspoint;
  p.weno_order = 5;
  p.recon_variable = reconstruction_variable == characteristic ? GKS_VAR_CHARACTERISTIC : GKS_VAR_CONSERVATIVE;
  p.use_df = is_using_df_factor ? 1 : 0;
  p.tau_type = tau_type == NS ? GKS_TAU_NS : GKS_TAU_EULER;
  p.K = K;
  p.prandtl_fix = is_Prandtl_fix ? 1 : 0;
  p.gamma = Gamma;
  p.mu = Mu;
  p.nu = Nu;
  p.pr = Pr;
  p.c1_euler = c1_euler;
  p.c2_euler = c2_euler;
  p.left = b.left;
  p.right = b.right;
  p.down = b.down;
  p.up = b.up;
  p.cfl = b.CFL;
  const char* fm = getenv("GKS_CONTRACTED");
  p.fast_math = (fm && fm[0] == '1') ? 1 : 0;
  return p;
}

void upload(gks2d_ctx* ctx, Fluid2d* f) {
  const int64_t cs = sizeof(Fluid2d);
  check(gks2d_set_convar(ctx, &f[0].convar[0], cs, sizeof(double)), "upload convar");
  check(gks2d_set_alpha(ctx, &f[0].alpha_x, &f[0].alpha_y, cs), "upload alpha");
  check(gks2d_set_cfl_primvar(ctx, &f[0].primvar[0], cs, sizeof(double)), "upload primvar");
}

// the device context of a cell array, created on first use
gks2d_ctx* bind(Fluid2d* fluids, const Block2d& block) {
  auto it = g_bound.find(fluids);
  if (it != g_bound.end()) return it->second.ctx;
  if (timecoe_list_2d != S2O4_2D) die("time scheme", "the device path implements S2O4_2D");
  if (flux_function_2d != GKS2D || gks2dsolver != gks2nd_2d)
    die("flux", "the device path implements GKS2D with gks2nd_2d");
  gks2d_params p = snapshot(block);
  gks2d_ctx* ctx = nullptr;
  check(gks2d_create(&p, &ctx), "gks2d_create");
  g_bound[fluids].ctx = ctx;
  upload(ctx, fluids);
  return ctx;
}

gks2d_ctx* bind_via(const void* arr) {
  auto it = g_owner.find(arr);
  if (it == g_owner.end()) die("sweep", "face array not set up by SetUniformMesh");
  auto b = g_bound.find(it->second);
  if (b == g_bound.end()) die("sweep", "cell array not bound yet (call a cell sweep first)");
  return b->second.ctx;
}

void boundary(Fluid2d* fluids, const Block2d& block, int side, int kind) {
  check(gks2d_boundary(bind(fluids, block), side, kind), "boundary");
}

}  // namespace

// ---------------------------------------------------------------- fluid_mesh
// SetGuassPoint: Gauss-Legendre (1-3 points) and Gauss-Lobatto (4 points)
// on [-1, 1] (fluid_mesh.cpp:9-45)
void SetGuassPoint() {
  if (gausspoint == 0) {
    cout << "no gausspoint specify" << endl;
    exit(0);
  }
  delete[] gauss_loc;
  delete[] gauss_weight;
  gauss_loc = new double[gausspoint];
  gauss_weight = new double[gausspoint];
  if (gausspoint == 1) {
    gauss_loc[0] = 0.0;
    gauss_weight[0] = 1.0;
  } else if (gausspoint == 2) {
    gauss_loc[0] = -sqrt(1.0 / 3.0);
    gauss_loc[1] = -gauss_loc[0];
    gauss_weight[0] = gauss_weight[1] = 0.5;
  } else if (gausspoint == 3) {
    gauss_loc[0] = 0.0;
    gauss_loc[1] = sqrt(3.0 / 5.0);
    gauss_loc[2] = -gauss_loc[1];
    gauss_weight[0] = 4.0 / 9.0;
    gauss_weight[1] = gauss_weight[2] = 0.5 - 0.5 * gauss_weight[0];
  } else if (gausspoint == 4) {
    gauss_loc[0] = -1.0;
    gauss_loc[1] = 1.0;
    gauss_loc[2] = -sqrt(5) / 5.0;
    gauss_loc[3] = -gauss_loc[2];
    gauss_weight[0] = gauss_weight[1] = 1.0 / 12.0;
    gauss_weight[2] = gauss_weight[3] = 5.0 / 12.0;
  }
  cout << gausspoint << " gausspoint(s) specify" << endl;
}

// ------------------------------------------------------------ basic_function
Fluid2d* Setfluid(Block2d& block) {
  Fluid2d* var = new Fluid2d[(size_t)block.nx * block.ny]();
  for (int i = 0; i < block.nx; i++)
    for (int j = 0; j < block.ny; j++) {
      var[i * block.ny + j].xindex = i;
      var[i * block.ny + j].yindex = j;
    }
  cout << "fluid variable allocate done..." << endl;
  return var;
}
Interface2d* Setinterface_array(Block2d block) {
  Interface2d* var = new Interface2d[(size_t)(block.nx + 1) * (block.ny + 1)];
  cout << "interface variable allocate done..." << endl;
  return var;
}
Flux2d_gauss** Setflux_gauss_array(Block2d block) {
  Flux2d_gauss** var = new Flux2d_gauss*[(size_t)(block.nx + 1) * (block.ny + 1)]();
  cout << "flux with gausspoint variable allocate done..." << endl;
  return var;
}

// geometry of SetUniformMesh (basic_function.cpp:1411-1491); also records
// which cell array the face arrays belong to
void SetUniformMesh(Block2d& block, Fluid2d* fluids, Interface2d* xinterfaces, Interface2d* yinterfaces,
                    Flux2d_gauss** xfluxes, Flux2d_gauss** yfluxes) {
  block.dx = (block.right - block.left) / block.nodex;
  block.dy = (block.up - block.down) / block.nodey;
  block.overdx = 1 / block.dx;
  block.overdy = 1 / block.dy;
  block.xcell_begin = block.ghost;
  block.xcell_end = block.ghost + block.nodex - 1;
  block.ycell_begin = block.ghost;
  block.ycell_end = block.ghost + block.nodey - 1;
  block.xinterface_begin_n = block.ghost;
  block.xinterface_end_n = block.ghost + block.nodex;
  block.xinterface_begin_t = block.ghost;
  block.xinterface_end_t = block.ghost + block.nodex - 1;
  block.yinterface_begin_n = block.ghost;
  block.yinterface_end_n = block.ghost + block.nodey;
  block.yinterface_begin_t = block.ghost;
  block.yinterface_end_t = block.ghost + block.nodey - 1;
  for (int i = 0; i < block.nx; i++)
    for (int j = 0; j < block.ny; j++) {
      Fluid2d& f = fluids[i * block.ny + j];
      f.dx = block.dx;
      f.dy = block.dy;
      f.coordx = block.left + (i + 0.5 - block.ghost) * block.dx;
      f.coordy = block.down + (j + 0.5 - block.ghost) * block.dy;
      f.area = block.dx * block.dy;
      const double x0 = block.left + (i - block.ghost) * block.dx, x1 = block.left + (i + 1 - block.ghost) * block.dx;
      const double y0 = block.down + (j - block.ghost) * block.dy, y1 = block.down + (j + 1 - block.ghost) * block.dy;
      const double nd[8] = {x0, y0, x1, y0, x1, y1, x0, y1};
      std::memcpy(f.node, nd, sizeof(nd));
    }
  for (int i = 0; i <= block.nx; i++)
    for (int j = 0; j <= block.ny; j++) {
      Interface2d& xf = xinterfaces[i * (block.ny + 1) + j];
      xf.x = block.left + (i - block.ghost) * block.dx;
      xf.y = block.down + (j - block.ghost + 0.5) * block.dy;
      xf.length = block.dy;
      xf.normal[0] = 1.0;
      xf.normal[1] = 0.0;
      Interface2d& yf = yinterfaces[i * (block.ny + 1) + j];
      yf.y = block.down + (j - block.ghost) * block.dy;
      yf.x = block.left + (i - block.ghost + 0.5) * block.dx;
      yf.length = block.dx;
      yf.normal[0] = 0.0;
      yf.normal[1] = 1.0;
    }
  g_owner[xinterfaces] = fluids;
  g_owner[yinterfaces] = fluids;
  g_owner[xfluxes] = fluids;
  g_owner[yfluxes] = fluids;
  cout << "set uniform information done..." << endl;
}

void CopyFluid_new_to_old(Fluid2d* fluids, Block2d block) { check(gks2d_copy_old(bind(fluids, block)), "copy_old"); }
void Convar_to_Primvar(Fluid2d* fluids, Block2d block) { check(gks2d_primvar(bind(fluids, block)), "primvar"); }

void Copy_Array(double* target, double* origin, int dim) {
  for (int i = 0; i < dim; i++) target[i] = origin[i];
}
// conversions restated from basic_function.cpp:811-817, 860-866
void Convar_to_primvar_2D(double* primvar, double convar[4]) {
  primvar[0] = convar[0];
  primvar[1] = convar[1] / convar[0];
  primvar[2] = convar[2] / convar[0];
  primvar[3] = (Gamma - 1.0) * (convar[3] - 0.5 * convar[1] * convar[1] / convar[0] -
                                0.5 * convar[2] * convar[2] / convar[0]);
}
void Primvar_to_convar_2D(double* convar, double primvar[4]) {
  convar[0] = primvar[0];
  convar[1] = primvar[0] * primvar[1];
  convar[2] = primvar[0] * primvar[2];
  convar[3] = primvar[0] * (primvar[3] / primvar[0] / (Gamma - 1) +
                            0.5 * (primvar[1] * primvar[1] + primvar[2] * primvar[2]));
}

// -------------------------------------------------------- boundary_condition
void free_boundary_left(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_LEFT, GKS_BC_FREE); }
void free_boundary_right(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_RIGHT, GKS_BC_FREE); }
void free_boundary_down(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_DOWN, GKS_BC_FREE); }
void free_boundary_up(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_UP, GKS_BC_FREE); }
void periodic_boundary_left(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_LEFT, GKS_BC_PERIODIC); }
void periodic_boundary_right(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_RIGHT, GKS_BC_PERIODIC); }
void periodic_boundary_down(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_DOWN, GKS_BC_PERIODIC); }
void periodic_boundary_up(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_UP, GKS_BC_PERIODIC); }
void reflection_boundary_left(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_LEFT, GKS_BC_REFLECT); }
void reflection_boundary_right(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_RIGHT, GKS_BC_REFLECT); }
void reflection_boundary_down(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_DOWN, GKS_BC_REFLECT); }
void reflection_boundary_up(Fluid2d* f, Block2d b, Fluid2d) { boundary(f, b, GKS_SIDE_UP, GKS_BC_REFLECT); }

// ------------------------------------------------------------ reconstruction
static void leaf_called(const char* name) {
  die(name, "is a plugin leaf of the device sweep; assign it to the scheme pointer instead of calling it");
}
void WENO5_AO_normal(Interface2d&, Interface2d&, Interface2d&, Interface2d&, Fluid2d*, Block2d) {
  leaf_called("WENO5_AO_normal");
}
void WENO5_AO_with_df_normal(Interface2d&, Interface2d&, Interface2d&, Interface2d&, Fluid2d*, Block2d) {
  leaf_called("WENO5_AO_with_df_normal");
}
void WENO5_AO_tangent(Interface2d*, Interface2d*, Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  leaf_called("WENO5_AO_tangent");
}
void WENO5_AO_with_df_tangent(Interface2d*, Interface2d*, Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  leaf_called("WENO5_AO_with_df_tangent");
}
void Center_do_nothing_normal(Interface2d*, Interface2d*, Fluid2d*, Block2d) {}
void Center_all_collision_multi(Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  leaf_called("Center_all_collision_multi");
}

void Reconstruction_within_cell(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids,
                                Block2d block) {
  if (cellreconstruction_2D_normal != WENO5_AO_with_df_normal ||
      cellreconstruction_2D_tangent != WENO5_AO_with_df_tangent)
    die("Reconstruction_within_cell", "the device path implements WENO5_AO_with_df_normal/_tangent");
  if (!is_using_df_factor) die("Reconstruction_within_cell", "the DF variants need is_using_df_factor = true");
  if (gausspoint != 2) die("Reconstruction_within_cell", "WENO5 tangential reconstruction needs gausspoint = 2");
  (void)xinterfaces;
  (void)yinterfaces;
  check(gks2d_reconstruct(bind(fluids, block)), "reconstruct");
}
// the collision centre state is computed inside the device flux kernel, once
// per Gauss point (the reference computes it here and again in GKS2D)
void Reconstruction_forg0(Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  if (g0reconstruction_2D_normal != Center_do_nothing_normal ||
      g0reconstruction_2D_tangent != Center_all_collision_multi)
    die("Reconstruction_forg0", "the device path implements Center_all_collision_multi");
}
void Update_alpha(Interface2d* xinterfaces, Interface2d*, Fluid2d* fluids, Block2d block) {
  (void)xinterfaces;
  check(gks2d_update_alpha(bind(fluids, block)), "update_alpha");
}

// ------------------------------------------------------------- flux_function
void Calculate_flux(Flux2d_gauss** xfluxes, Flux2d_gauss**, Interface2d* xinterfaces, Interface2d*,
                    Block2d block, int stage) {
  gks2d_ctx* ctx = xinterfaces ? bind_via(xinterfaces) : bind_via(xfluxes);
  check(gks2d_flux(ctx, stage, block.dt), "flux");
}
void GKS2D(Flux2d& flux, Recon2d& re, double dt) {
  Block2d b;
  b.nodex = b.nodey = 8;
  b.right = b.up = 1.0;
  gks2d_params p = snapshot(b);
  double L[12], R[12], C[12];
  std::memcpy(L, re.left.convar, 32);
  std::memcpy(L + 4, re.left.der1x, 32);
  std::memcpy(L + 8, re.left.der1y, 32);
  std::memcpy(R, re.right.convar, 32);
  std::memcpy(R + 4, re.right.der1x, 32);
  std::memcpy(R + 8, re.right.der1y, 32);
  check(gks2d_flux_points(&p, 1, L, R, dt, C, flux.f, flux.derf), "GKS2D");
  std::memcpy(re.center.convar, C, 32);
  std::memcpy(re.center.der1x, C + 4, 32);
  std::memcpy(re.center.der1y, C + 8, 32);
}

// -------------------------------------------------------------- time_advance
// S2O4 coefficients (time_advance.cpp:221-230)
void S2O4_2D(Block2d& block) {
  block.stages = 2;
  block.timecoefficient[0][0][0] = 0.5;
  block.timecoefficient[0][0][1] = 1.0 / 8.0;
  block.timecoefficient[1][0][0] = 1.0;
  block.timecoefficient[1][1][0] = 0.0;
  block.timecoefficient[1][0][1] = 1.0 / 6.0;
  block.timecoefficient[1][1][1] = 1.0 / 3.0;
}
void Initial_stages(Block2d& block) {
  std::memset(block.timecoefficient, 0, sizeof(block.timecoefficient));
  if (!timecoe_list_2d) die("Initial_stages", "no time scheme assigned to timecoe_list_2d");
  timecoe_list_2d(block);
}
// Get_CFL (time_advance.cpp:264-290): dt from the device's stale primvar;
// prints like the reference and exits on a reported blow-up
double Get_CFL(Block2d& block, Fluid2d* fluids, double tstop) {
  gks2d_ctx* ctx = bind(fluids, block);
  double dt = 0.0;
  int rc = gks2d_cfl(ctx, block.t, std::numeric_limits<double>::infinity(), &dt);
  if (rc == GKS_ERR_BLOWUP) {
    cout << "the program blows up at t=" << block.t << "!" << endl;
    exit(0);
  }
  check(rc, "Get_CFL");
  if (block.step % 10 == 0) cout << "step = " << block.step << " time size is " << dt << " time = " << block.t << endl;
  if (block.t + dt > tstop) {
    dt = tstop - block.t + 1e-16;
    cout << "last step here" << endl;
  }
  return dt;
}
void Update(Fluid2d* fluids, Flux2d_gauss**, Flux2d_gauss**, Block2d block, int stage) {
  if (stage > block.stages) die("Update", "wrong middle stage,pls check the time marching setting");
  check(gks2d_update(bind(fluids, block), stage), "update");
}

// -------------------------------------------------------------------- output
void Download_fluid(Fluid2d* fluids, Block2d block) {
  auto it = g_bound.find(fluids);
  if (it == g_bound.end()) return;  // never left the host
  const int64_t cs = sizeof(Fluid2d);
  check(gks2d_get_convar(it->second.ctx, &fluids[0].convar[0], cs, sizeof(double)), "download convar");
  check(gks2d_get_alpha(it->second.ctx, &fluids[0].alpha_x, &fluids[0].alpha_y, cs), "download alpha");
  (void)block;
}
void Upload_fluid(Fluid2d* fluids, Block2d block) { upload(bind(fluids, block), fluids); }
void Release_fluid(Fluid2d* fluids) {
  auto it = g_bound.find(fluids);
  if (it == g_bound.end()) return;
  gks2d_destroy(it->second.ctx);
  g_bound.erase(it);
  for (auto o = g_owner.begin(); o != g_owner.end();) {
    if (o->second == fluids) o = g_owner.erase(o);
    else ++o;
  }
}

void output2d(Fluid2d* fluids, Block2d block) {
  Download_fluid(fluids, block);
  std::stringstream name;
  name << "../data/R-" << block.step << ".txt";
  std::ofstream ofs(name.str(), std::ios::trunc);
  for (int j = block.ghost; j < block.ghost + block.nodey; j++)
    for (int i = block.ghost; i < block.ghost + block.nodex; i++) ofs << fluids[i * block.ny + j].convar[0] << endl;
}

// ----------------------------------------------------------- problem set-up
// IC_for_riemann_2d (2Dproblem.cpp:345-385): quadrant primitive states in the
// interior (both axes tested against discon/dx), convar for every cell
void IC_for_riemann_2d(Fluid2d* fluid, double* zone1, double* zone2, double* zone3, double* zone4, Block2d block,
                       double* discon) {
  for (int i = block.ghost; i < block.nx - block.ghost; i++)
    for (int j = block.ghost; j < block.ny - block.ghost; j++) {
      const bool xl = i - block.ghost + 1 < discon[0] / block.dx;
      const bool yl = j - block.ghost + 1 < discon[1] / block.dx;
      double* z = (xl && yl) ? zone1 : (!xl && yl) ? zone2 : (!xl && !yl) ? zone3 : zone4;
      Copy_Array(fluid[i * block.ny + j].primvar, z, 4);
    }
  for (int i = 0; i < block.nx; i++)
    for (int j = 0; j < block.ny; j++)
      Primvar_to_convar_2D(fluid[i * block.ny + j].convar, fluid[i * block.ny + j].primvar);
}
