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
bool gks_fast_kernels = false;
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
  // the scheme from the identity of the reconstruction pointers
  p.weno_order = cellreconstruction_2D_normal == WENO9_AO_with_df_normal   ? 9
                 : cellreconstruction_2D_normal == WENO7_AO_with_df_normal ? 7
                                                                           : 5;
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
  p.fast_math = gks_fast_kernels ? 1 : 0;
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
// the face arrays with their per-Gauss-point records, as the reference
// allocates them (basic_function.cpp:1322-1345); the records stay zero
Interface2d* Setinterface_array(Block2d block) {
  const size_t nf = (size_t)(block.nx + 1) * (block.ny + 1);
  Interface2d* var = new Interface2d[nf];
  const int gp = gausspoint > 0 ? gausspoint : 1;
  Recon2d* rec = new Recon2d[nf * gp]();
  for (size_t f = 0; f < nf; f++) var[f].gauss = rec + f * gp;
  cout << "interface variable allocate done..." << endl;
  return var;
}
Flux2d_gauss** Setflux_gauss_array(Block2d block) {
  const size_t nf = (size_t)(block.nx + 1) * (block.ny + 1);
  const int stages = block.stages > 0 ? block.stages : 1;  // set by Initial_stages
  const int gp = gausspoint > 0 ? gausspoint : 1;
  Flux2d_gauss** var = new Flux2d_gauss*[nf];
  Flux2d_gauss* rows = new Flux2d_gauss[nf * stages];
  Flux2d* rec = new Flux2d[nf * stages * gp]();
  for (size_t f = 0; f < nf; f++) {
    var[f] = rows + f * stages;
    for (int s = 0; s < stages; s++) var[f][s].gauss = rec + (f * stages + s) * gp;
  }
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

void WENO7_AO_with_df_normal(Interface2d&, Interface2d&, Interface2d&, Interface2d&, Fluid2d*, Block2d) {
  leaf_called("WENO7_AO_with_df_normal");
}
void WENO7_AO_with_df_tangent(Interface2d*, Interface2d*, Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  leaf_called("WENO7_AO_with_df_tangent");
}
void WENO9_AO_with_df_normal(Interface2d&, Interface2d&, Interface2d&, Interface2d&, Fluid2d*, Block2d) {
  leaf_called("WENO9_AO_with_df_normal");
}
void WENO9_AO_with_df_tangent(Interface2d*, Interface2d*, Interface2d*, Interface2d*, Fluid2d*, Block2d) {
  leaf_called("WENO9_AO_with_df_tangent");
}

void Reconstruction_within_cell(Interface2d* xinterfaces, Interface2d* yinterfaces, Fluid2d* fluids,
                                Block2d block) {
  // the normal / tangential pair of one order, with its Gauss points
  struct Pair {
    Reconstruction_within_Cell_2D_normal n;
    Reconstruction_within_Cell_2D_tangent t;
    int gp;
    const char* name;
  };
  static const Pair pairs[3] = {{WENO5_AO_with_df_normal, WENO5_AO_with_df_tangent, 2, "WENO5"},
                                {WENO7_AO_with_df_normal, WENO7_AO_with_df_tangent, 3, "WENO7"},
                                {WENO9_AO_with_df_normal, WENO9_AO_with_df_tangent, 4, "WENO9"}};
  const Pair* pr = nullptr;
  for (const Pair& q : pairs)
    if (cellreconstruction_2D_normal == q.n && cellreconstruction_2D_tangent == q.t) pr = &q;
  if (!pr)
    die("Reconstruction_within_cell",
        "the device path implements WENO5/7/9_AO_with_df_normal with the _tangent of the same order");
  if (!is_using_df_factor) die("Reconstruction_within_cell", "the DF variants need is_using_df_factor = true");
  if (gausspoint != pr->gp)
    die("Reconstruction_within_cell", std::string(pr->name) + " tangential reconstruction needs gausspoint = " +
                                          std::to_string(pr->gp));
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

// ======================================================================= 3D
// The 3D call surface over include/gks3d.h.  Defaults of the scheme globals
// as the reference initialises them (fluid_mesh.cpp:6-7,
// reconstruction.cpp:3690-3696, flux_function.cpp:757-758); the reference's
// default time scheme (S1O1_3D) is not implemented, so timecoe_list_3d
// starts unset.
#include "../../../include/gks3d.h"

double** gauss_loc_3d = nullptr;
bool is_multi_dimensional = true;
G0_construct_type g0type = all_collisionn;
Reconstruction_within_Cell_3D_of_face_value cellreconstruction_3D_of_face_value = WENO5_AO_splitting_3d;
Reconstruction_forG0_3D_of_face_value g0reconstruction_3D_of_face_value = Do_nothing_splitting_3d;
Reconstruction_within_Cell_3D_of_line_value cellreconstruction_3D_of_line_value = WENO5_AO_of_line_value;
Reconstruction_within_Cell_3D_of_point_value cellreconstruction_3D_of_point_value = WENO5_AO_of_point_value;
Reconstruction_forG0_3D_of_line_value g0reconstruction_3D_of_line_value = Do_nothing_reconstruction_of_line_value;
Reconstruction_forG0_3D_of_point_value g0reconstruction_3D_of_point_value = Do_nothing_reconstruction_of_point_value;
GKS3d_type gks3dsolver = gks2nd_3d;
Flux_function_3d flux_function_3d = GKS3D;
TimeMarchingCoefficient_3d timecoe_list_3d = nullptr;

namespace {

std::map<const void*, Fluid3d*> g_owner3;  // interface / flux arrays -> cell array
std::map<const Fluid3d*, gks3d_ctx*> g_bound3;

void check3(int rc, const char* where) {
  if (rc != GKS_OK) die(where, gks3d_last_error());
}

gks3d_params snapshot3(const Block3d& b) {
  gks3d_params p;
  std::memset(&p, 0, sizeof(p));
  p.nodex = b.nodex;
  p.nodey = b.nodey;
  p.nodez = b.nodez;
  p.ghost = b.ghost;
  p.recon_variable = reconstruction_variable == characteristic ? GKS_VAR_CHARACTERISTIC : GKS_VAR_CONSERVATIVE;
  p.tau_type = tau_type == NS ? GKS_TAU_NS : GKS_TAU_EULER;
  p.K = K;
  p.fast_math = gks_fast_kernels ? 1 : 0;
  p.gamma = Gamma;
  p.mu = Mu;
  p.nu = Nu;
  p.c1_euler = c1_euler;
  p.c2_euler = c2_euler;
  p.left = b.left;
  p.right = b.right;
  p.down = b.down;
  p.up = b.up;
  p.back = b.back;
  p.front = b.front;
  p.cfl = b.CFL;
  return p;
}

void upload3(gks3d_ctx* ctx, Fluid3d* f) {
  const int64_t cs = sizeof(Fluid3d);
  check3(gks3d_set_convar(ctx, &f[0].convar[0], cs, sizeof(double)), "upload convar");
  check3(gks3d_set_cfl_primvar(ctx, &f[0].primvar[0], cs, sizeof(double)), "upload primvar");
}

gks3d_ctx* bind3(Fluid3d* fluids, const Block3d& block) {
  auto it = g_bound3.find(fluids);
  if (it != g_bound3.end()) return it->second;
  if (timecoe_list_3d != S2O4_3D) die("time scheme", "the device path implements S2O4_3D");
  if (flux_function_3d != GKS3D || gks3dsolver != gks2nd_3d)
    die("flux", "the device path implements GKS3D with gks2nd_3d");
  gks3d_params p = snapshot3(block);
  gks3d_ctx* ctx = nullptr;
  check3(gks3d_create(&p, &ctx), "gks3d_create");
  g_bound3[fluids] = ctx;
  upload3(ctx, fluids);
  return ctx;
}

gks3d_ctx* bind3_via(const void* arr) {
  auto it = g_owner3.find(arr);
  if (it == g_owner3.end()) die("sweep", "face array not set up by SetUniformMesh");
  auto b = g_bound3.find(it->second);
  if (b == g_bound3.end()) die("sweep", "cell array not bound yet (call a cell sweep first)");
  return b->second;
}

void boundary3(Fluid3d* fluids, const Block3d& block, int side, int kind) {
  check3(gks3d_boundary(bind3(fluids, block), side, kind), "boundary");
}

}  // namespace

// SetGuassPoint_3D (fluid_mesh.cpp:47-95): 1 or 2 x 2 Gauss points on a face
void SetGuassPoint_3D() {
  if (!is_multi_dimensional) gausspoint = 0;
  if (gausspoint != 1 && gausspoint != 4) die("SetGuassPoint_3D", "the device path uses gausspoint = 4 (2 x 2)");
  gauss_loc_3d = new double*[gausspoint];
  for (int i = 0; i < gausspoint; i++) gauss_loc_3d[i] = new double[2];
  delete[] gauss_weight;
  gauss_weight = new double[gausspoint];
  if (gausspoint == 1) {
    gauss_loc_3d[0][0] = gauss_loc_3d[0][1] = 0.0;
    gauss_weight[0] = 1.0;
  } else {
    const double a = -sqrt(1.0 / 3.0);
    const double loc[4][2] = {{a, a}, {a, -a}, {-a, a}, {-a, -a}};
    for (int q = 0; q < 4; q++) {
      gauss_loc_3d[q][0] = loc[q][0];
      gauss_loc_3d[q][1] = loc[q][1];
      gauss_weight[q] = 0.25;
    }
  }
}

Fluid3d* Setfluid_array(Block3d& block) {
  Fluid3d* var = new Fluid3d[(size_t)block.nx * block.ny * block.nz]();
  for (int i = 0; i < block.nx; i++)
    for (int j = 0; j < block.ny; j++)
      for (int k = 0; k < block.nz; k++) {
        Fluid3d& f = var[((size_t)i * block.ny + j) * block.nz + k];
        f.index[0] = i;
        f.index[1] = j;
        f.index[2] = k;
      }
  cout << "fluid memory allocate done" << endl;
  return var;
}
Interface3d* Setinterface_array(Block3d& block) {
  Interface3d* var = new Interface3d[(size_t)(block.nx + 1) * (block.ny + 1) * (block.nz + 1)];
  for (int i = 0; i <= block.nx; i++)
    for (int j = 0; j <= block.ny; j++)
      for (int k = 0; k <= block.nz; k++) {
        Interface3d& f = var[((size_t)i * (block.ny + 1) + j) * (block.nz + 1) + k];
        f.index[0] = i;
        f.index[1] = j;
        f.index[2] = k;
      }
  cout << "interface variable allocate done..." << endl;
  return var;
}
Flux3d_gauss** Setflux_gauss_array(Block3d& block) {
  Flux3d_gauss** var = new Flux3d_gauss*[(size_t)(block.nx + 1) * (block.ny + 1) * (block.nz + 1)]();
  cout << "flux with gausspoint variable allocate done..." << endl;
  return var;
}

// the geometry of SetUniformMesh (3D): cell centres, the eight nodes, volume
// and R_c (the smallest cell size, Get_CFL's length); face Gauss-point
// locations, weights, normals and areas; also records which cell array the
// face arrays belong to
void SetUniformMesh(Block3d block, Fluid3d* fluids, Interface3d* xinterfaces, Interface3d* yinterfaces,
                    Interface3d* zinterfaces) {
  if (gausspoint == 0) die("SetUniformMesh", "the number of gauss points must be > 0");
  const size_t ny = block.ny, nz = block.nz;
  for (int i = 0; i < block.nx; i++)
    for (int j = 0; j < block.ny; j++)
      for (int k = 0; k < block.nz; k++) {
        Fluid3d& f = fluids[((size_t)i * ny + j) * nz + k];
        f.loc[0] = block.left + (i + 0.5 - block.ghost) * block.dx;
        f.loc[1] = block.down + (j + 0.5 - block.ghost) * block.dy;
        f.loc[2] = block.back + (k + 0.5 - block.ghost) * block.dz;
        for (int n = 0; n < 8; n++) {  // nodes 0-3 at z-, 4-7 at z+, counter-clockwise from (x-, y-)
          const int sx = (n & 3) == 1 || (n & 3) == 2, sy = (n & 3) >= 2, sz = n >= 4;
          f.node[3 * n + 0] = f.loc[0] + (sx ? 0.5 : -0.5) * block.dx;
          f.node[3 * n + 1] = f.loc[1] + (sy ? 0.5 : -0.5) * block.dy;
          f.node[3 * n + 2] = f.loc[2] + (sz ? 0.5 : -0.5) * block.dz;
        }
        f.vol = block.dx * block.dy * block.dz;
        f.R_c = block.dx;
        if (block.dy < f.R_c) f.R_c = block.dy;
        if (block.dz < f.R_c) f.R_c = block.dz;
      }
  for (int i = 0; i <= block.nx; i++)
    for (int j = 0; j <= block.ny; j++)
      for (int k = 0; k <= block.nz; k++) {
        const size_t idx = ((size_t)i * (ny + 1) + j) * (nz + 1) + k;
        const double x0 = block.left + (i - block.ghost) * block.dx;
        const double y0 = block.down + (j - block.ghost) * block.dy;
        const double z0 = block.back + (k - block.ghost) * block.dz;
        Interface3d* faces[3] = {&xinterfaces[idx], &yinterfaces[idx], &zinterfaces[idx]};
        const double areas[3] = {block.xarea, block.yarea, block.zarea};
        for (int d = 0; d < 3; d++) {
          Interface3d& f = *faces[d];
          // the face's centre and its two in-plane edge directions
          double c[3] = {x0 + (d == 0 ? 0.0 : 0.5 * block.dx), y0 + (d == 1 ? 0.0 : 0.5 * block.dy),
                         z0 + (d == 2 ? 0.0 : 0.5 * block.dz)};
          double e0[3] = {0, 0, 0}, e1[3] = {0, 0, 0};
          if (d == 0) { e0[1] = block.dy; e1[2] = block.dz; }
          else if (d == 1) { e0[0] = block.dx; e1[2] = block.dz; }
          else { e0[0] = block.dx; e1[1] = block.dy; }
          for (int q = 0; q < gausspoint; q++) {
            for (int a = 0; a < 3; a++)
              f.loc[3 * q + a] = c[a] + 0.5 * e0[a] * gauss_loc_3d[q][0] + 0.5 * e1[a] * gauss_loc_3d[q][1];
            for (int a = 0; a < 3; a++) f.normal[3 * q + a] = a == d ? 1.0 : 0.0;
            f.weight[q] = gauss_weight[q];
          }
          f.area = areas[d];
        }
      }
  g_owner3[xinterfaces] = fluids;
  g_owner3[yinterfaces] = fluids;
  g_owner3[zinterfaces] = fluids;
  cout << "set uniform information done..." << endl;
}

void CopyFluid_new_to_old(Fluid3d* fluids, Block3d block) { check3(gks3d_copy_old(bind3(fluids, block)), "copy_old"); }
void Convar_to_primvar(Fluid3d* fluids, Block3d& block) { check3(gks3d_primvar(bind3(fluids, block)), "primvar"); }
// Convar_to_primvar_3D / Primvar_to_convar_3D (basic_function.cpp:819-826,
// 859-866)
void Convar_to_primvar_3D(double* primvar, double convar[5]) {
  primvar[0] = convar[0];
  primvar[1] = convar[1] / convar[0];
  primvar[2] = convar[2] / convar[0];
  primvar[3] = convar[3] / convar[0];
  primvar[4] = (Gamma - 1.0) * (convar[4] - 0.5 * convar[1] * convar[1] / convar[0] -
                                0.5 * convar[2] * convar[2] / convar[0] - 0.5 * convar[3] * convar[3] / convar[0]);
}
void Primvar_to_convar_3D(double* convar, double* primvar) {
  convar[0] = primvar[0];
  convar[1] = primvar[0] * primvar[1];
  convar[2] = primvar[0] * primvar[2];
  convar[3] = primvar[0] * primvar[3];
  convar[4] = primvar[0] * (primvar[4] / primvar[0] / (Gamma - 1) +
                            0.5 * (primvar[1] * primvar[1] + primvar[2] * primvar[2] + primvar[3] * primvar[3]));
}

void free_boundary_xleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_XLEFT, GKS_BC_FREE); }
void free_boundary_xright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_XRIGHT, GKS_BC_FREE); }
void free_boundary_yleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_YLEFT, GKS_BC_FREE); }
void free_boundary_yright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_YRIGHT, GKS_BC_FREE); }
void free_boundary_zleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_ZLEFT, GKS_BC_FREE); }
void free_boundary_zright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_ZRIGHT, GKS_BC_FREE); }
void periodic_boundary_xleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_XLEFT, GKS_BC_PERIODIC); }
void periodic_boundary_xright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_XRIGHT, GKS_BC_PERIODIC); }
void periodic_boundary_yleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_YLEFT, GKS_BC_PERIODIC); }
void periodic_boundary_yright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_YRIGHT, GKS_BC_PERIODIC); }
void periodic_boundary_zleft(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_ZLEFT, GKS_BC_PERIODIC); }
void periodic_boundary_zright(Fluid3d* f, Block3d b, Fluid3d) { boundary3(f, b, GKS_SIDE3_ZRIGHT, GKS_BC_PERIODIC); }

void WENO5_AO_splitting_3d(Interface3d&, Interface3d&, Interface3d&, Interface3d&, Interface3d&, Interface3d&,
                           Fluid3d*, Block3d) {
  leaf_called("WENO5_AO_splitting_3d");
}
void WENO5_AO_of_line_value(Interface3d*, Interface3d*, Interface3d*, Block3d) { leaf_called("WENO5_AO_of_line_value"); }
void WENO5_AO_of_point_value(Interface3d*, Interface3d*, Interface3d*, Block3d) {
  leaf_called("WENO5_AO_of_point_value");
}
void Do_nothing_splitting_3d(Interface3d*, Interface3d*, Interface3d*, Fluid3d*, Block3d) {}
void Do_nothing_reconstruction_of_line_value(Interface3d*, Interface3d*, Interface3d*, Block3d) {}
void Do_nothing_reconstruction_of_point_value(Interface3d*, Interface3d*, Interface3d*, Block3d) {}

// the face, line and point reconstructions run inside the device flux
// sweep, one direction at a time (their buffers are reused by the next
// direction); this checks the scheme pointers
void Reconstruction_within_cell(Interface3d*, Interface3d*, Interface3d*, Fluid3d*, Block3d) {
  if (cellreconstruction_3D_of_face_value != WENO5_AO_splitting_3d ||
      cellreconstruction_3D_of_line_value != WENO5_AO_of_line_value ||
      cellreconstruction_3D_of_point_value != WENO5_AO_of_point_value)
    die("Reconstruction_within_cell", "the device path implements WENO5_AO_splitting_3d / _of_line_value / _of_point_value");
  if (gausspoint != 4) die("Reconstruction_within_cell", "the 3D reconstruction needs gausspoint = 4");
}
void Reconstruction_forg0(Interface3d*, Interface3d*, Interface3d*, Fluid3d*, Block3d) {
  if (g0reconstruction_3D_of_face_value != Do_nothing_splitting_3d ||
      g0reconstruction_3D_of_line_value != Do_nothing_reconstruction_of_line_value ||
      g0reconstruction_3D_of_point_value != Do_nothing_reconstruction_of_point_value || g0type != all_collisionn)
    die("Reconstruction_forg0", "the device path computes the collision state inside GKS3D (all_collisionn)");
}

void Calculate_flux(Flux3d_gauss** xfluxes, Flux3d_gauss**, Flux3d_gauss**, Interface3d* xinterfaces, Interface3d*,
                    Interface3d*, Block3d block, int stage) {
  gks3d_ctx* ctx = xinterfaces ? bind3_via(xinterfaces) : bind3_via(xfluxes);
  check3(gks3d_flux(ctx, stage, block.dt), "flux");
}
void GKS3D(Flux3d& flux, Recon3d& re, double area, double dt) {
  (void)area;
  Block3d b;
  b.nodex = b.nodey = b.nodez = 8;
  b.right = b.up = b.front = 1.0;
  b.CFL = 0.5;
  gks3d_params p = snapshot3(b);
  double L[20], R[20], C[20];
  const Point3d* sides[2] = {&re.left, &re.right};
  double* dst[2] = {L, R};
  for (int s = 0; s < 2; s++) {
    std::memcpy(dst[s], sides[s]->convar, 40);
    std::memcpy(dst[s] + 5, sides[s]->der1x, 40);
    std::memcpy(dst[s] + 10, sides[s]->der1y, 40);
    std::memcpy(dst[s] + 15, sides[s]->der1z, 40);
  }
  check3(gks3d_flux_points(&p, 1, L, R, dt, C, flux.f, flux.derf), "GKS3D");
  std::memcpy(re.center.convar, C, 40);
  std::memcpy(re.center.der1x, C + 5, 40);
  std::memcpy(re.center.der1y, C + 10, 40);
  std::memcpy(re.center.der1z, C + 15, 40);
}

// S2O4 coefficients (time_advance.cpp:732-745)
void S2O4_3D(Block3d& block) {
  block.stages = 2;
  block.timecoefficient[0][0][0] = 0.5;
  block.timecoefficient[0][0][1] = 1.0 / 8.0;
  block.timecoefficient[1][0][0] = 1.0;
  block.timecoefficient[1][1][0] = 0.0;
  block.timecoefficient[1][0][1] = 1.0 / 6.0;
  block.timecoefficient[1][1][1] = 1.0 / 3.0;
  block.timecoefficient_hweno[0][0][0] = 0.5;
  block.timecoefficient_hweno[1][0][0] = 0.0;
  block.timecoefficient_hweno[1][1][0] = 1.0;
}
void Initial_stages(Block3d& block) {
  std::memset(block.timecoefficient, 0, sizeof(block.timecoefficient));
  std::memset(block.timecoefficient_hweno, 0, sizeof(block.timecoefficient_hweno));
  if (!timecoe_list_3d) die("Initial_stages", "no time scheme assigned to timecoe_list_3d");
  timecoe_list_3d(block);
}
// Get_CFL (time_advance.cpp:793-831): dt from the last primitive sweep (or
// the initial condition's primvar), clipped to tstop; prints like the
// reference and exits on a reported blow-up
double Get_CFL(Block3d& block, Fluid3d* fluids, double tstop) {
  gks3d_ctx* ctx = bind3(fluids, block);
  double dt = 0.0;
  const int rc = gks3d_cfl(ctx, block.t, tstop, &dt);
  if (rc == GKS_ERR_BLOWUP) {
    cout << "the program blows up at t=" << block.t << "!" << endl;
    exit(0);
  }
  check3(rc, "Get_CFL");
  cout << "step = " << block.step << "time size is " << dt << " time = " << block.t << endl;
  return dt;
}
void Update_gauss(Fluid3d* fluids, Flux3d_gauss**, Flux3d_gauss**, Flux3d_gauss**, Block3d block, int stage) {
  if (stage > block.stages) die("Update_gauss", "wrong middle stage ,pls check the time marching setting");
  check3(gks3d_update(bind3(fluids, block), stage), "update");
}

void Download_fluid(Fluid3d* fluids, Block3d block) {
  auto it = g_bound3.find(fluids);
  if (it == g_bound3.end()) return;
  check3(gks3d_get_convar(it->second, &fluids[0].convar[0], sizeof(Fluid3d), sizeof(double)), "download convar");
  (void)block;
}
void Upload_fluid(Fluid3d* fluids, Block3d block) { upload3(bind3(fluids, block), fluids); }
void Release_fluid(Fluid3d* fluids) {
  auto it = g_bound3.find(fluids);
  if (it == g_bound3.end()) return;
  gks3d_destroy(it->second);
  g_bound3.erase(it);
  for (auto o = g_owner3.begin(); o != g_owner3.end();) {
    if (o->second == fluids) o = g_owner3.erase(o);
    else ++o;
  }
}

void output3d(Fluid3d* fluids, Block3d block) {
  Download_fluid(fluids, block);
  std::ofstream ofs("../data/R3d.txt", std::ios::trunc);
  for (int k = block.ghost; k < block.ghost + block.nodez; k++)
    for (int j = block.ghost; j < block.ghost + block.nodey; j++)
      for (int i = block.ghost; i < block.ghost + block.nodex; i++)
        ofs << fluids[((size_t)i * block.ny + j) * block.nz + k].convar[0] << endl;
}

// ICfor_cubic_sod (3Dproblem.cpp:174-223): zone1 in the cells whose centres
// lie within 0.2 of the domain centre in every direction, zone2 elsewhere;
// interior cells only
void ICfor_cubic_sod(Fluid3d* fluids, double* zone1, double* zone2, Block3d block) {
  for (int i = block.ghost; i < block.nx - block.ghost; i++)
    for (int j = block.ghost; j < block.ny - block.ghost; j++)
      for (int k = block.ghost; k < block.nz - block.ghost; k++) {
        Fluid3d& f = fluids[((size_t)i * block.ny + j) * block.nz + k];
        const double xr = std::abs(f.loc[0] - 0.5 * (block.right + block.left));
        const double yr = std::abs(f.loc[1] - 0.5 * (block.up + block.down));
        const double zr = std::abs(f.loc[2] - 0.5 * (block.front + block.back));
        const bool inner = xr < 0.2 * (block.right - block.left) && yr < 0.2 * (block.up - block.down) &&
                           zr < 0.2 * (block.front - block.back);
        Copy_Array(f.primvar, inner ? zone1 : zone2, 5);
        Primvar_to_convar_3D(f.convar, f.primvar);
      }
}

// ======================================================================= 1D
// The 1D call surface over include/gks1d.h.  Defaults as the reference
// initialises them (reconstruction.cpp:12-14, flux_function.cpp:11-14,
// time_advance.cpp:5).
#include "../../../include/gks1d.h"

Reconstruction_within_Cell cellreconstruction = WENO5_AO;
Reconstruction_forG0 g0reconstruction = Center_collision;
GKS1d_type gks1dsolver = nothing;
Flux_function flux_function = GKS;
TimeMarchingCoefficient timecoe_list = nullptr;

namespace {

std::map<const void*, Fluid1d*> g_owner1;  // interface / flux arrays -> cell array
std::map<const Fluid1d*, gks1d_ctx*> g_bound1;

void check1(int rc, const char* where) {
  if (rc != GKS_OK) die(where, gks1d_last_error());
}

gks1d_params snapshot1(const Block1d& b) {
  gks1d_params p;
  std::memset(&p, 0, sizeof(p));
  p.nodex = b.nodex;
  p.ghost = b.ghost;
  p.recon_variable = reconstruction_variable == characteristic ? GKS_VAR_CHARACTERISTIC : GKS_VAR_CONSERVATIVE;
  p.use_df = is_using_df_factor ? 1 : 0;
  p.df_from_start = 0;  // the driver's pointer assignment decides, per call
  p.tau_type = tau_type == NS ? GKS_TAU_NS : GKS_TAU_EULER;
  p.K = K;
  p.fast_math = gks_fast_kernels ? 1 : 0;
  p.gamma = Gamma;
  p.mu = Mu;
  p.nu = Nu;
  p.pr = Pr;
  p.c1_euler = c1_euler;
  p.c2_euler = c2_euler;
  p.left = b.left;
  p.right = b.right;
  p.cfl = b.CFL;
  return p;
}

gks1d_ctx* bind1(Fluid1d* fluids, const Block1d& block) {
  auto it = g_bound1.find(fluids);
  if (it != g_bound1.end()) return it->second;
  if (timecoe_list != S2O4) die("time scheme", "the device path implements S2O4");
  if (flux_function != GKS || gks1dsolver != gks2nd) die("flux", "the device path implements GKS with gks2nd");
  gks1d_params p = snapshot1(block);
  gks1d_ctx* ctx = nullptr;
  check1(gks1d_create(&p, &ctx), "gks1d_create");
  g_bound1[fluids] = ctx;
  check1(gks1d_set_convar(ctx, &fluids[0].convar[0], sizeof(Fluid1d), sizeof(double)), "upload convar");
  return ctx;
}

gks1d_ctx* bind1_via(const void* arr) {
  auto it = g_owner1.find(arr);
  if (it == g_owner1.end()) die("sweep", "interface array not set up by SetUniformMesh");
  auto b = g_bound1.find(it->second);
  if (b == g_bound1.end()) die("sweep", "cell array not bound yet (call a cell sweep first)");
  return b->second;
}

void boundary1(Fluid1d* fluids, const Block1d& block, int side, int kind) {
  check1(gks1d_boundary(bind1(fluids, block), side, kind), "boundary");
}

}  // namespace

Flux1d** Setflux_array(Block1d block) {
  Flux1d** var = new Flux1d*[block.nx + 1]();
  return var;
}
// the geometry of SetUniformMesh (1D): cell centres and sizes, interface
// coordinates; also records which cell array the face arrays belong to
void SetUniformMesh(Block1d block, Fluid1d* fluids, Interface1d* interfaces, Flux1d** fluxes) {
  for (int i = 0; i < block.nx; i++) {
    fluids[i].dx = block.dx;
    fluids[i].cx = block.left + (i + 0.5 - block.ghost) * block.dx;
  }
  for (int i = 0; i <= block.nx; i++) {
    interfaces[i].x = block.left + (i - block.ghost) * block.dx;
    interfaces[i].left.x = interfaces[i].x;
    interfaces[i].center.x = interfaces[i].x;
    interfaces[i].right.x = interfaces[i].x;
  }
  g_owner1[interfaces] = fluids;
  g_owner1[fluxes] = fluids;
}
void CopyFluid_new_to_old(Fluid1d* fluids, Block1d block) { check1(gks1d_copy_old(bind1(fluids, block)), "copy_old"); }
// Convar_to_primvar_1D / Primvar_to_convar_1D (basic_function.cpp:805-809, 853-858)
void Convar_to_primvar_1D(double* primvar, double convar[3]) {
  primvar[0] = convar[0];
  primvar[1] = convar[1] / convar[0];
  primvar[2] = (Gamma - 1.0) * (convar[2] - 0.5 * convar[1] * convar[1] / convar[0]);
}
void Primvar_to_convar_1D(double* convar, double primvar[3]) {
  convar[0] = primvar[0];
  convar[1] = primvar[0] * primvar[1];
  convar[2] = primvar[0] * (primvar[2] / primvar[0] / (Gamma - 1) + 0.5 * (primvar[1] * primvar[1]));
}

void free_boundary_left(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 0, GKS_BC_FREE); }
void free_boundary_right(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 1, GKS_BC_FREE); }
void reflection_boundary_left(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 0, GKS_BC_REFLECT); }
void reflection_boundary_right(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 1, GKS_BC_REFLECT); }
void periodic_boundary_left(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 0, GKS_BC_PERIODIC); }
void periodic_boundary_right(Fluid1d* f, Block1d b, Fluid1d) { boundary1(f, b, 1, GKS_BC_PERIODIC); }

void WENO5_AO(Point1d&, Point1d&, Fluid1d*, Block1d) { leaf_called("WENO5_AO"); }
void WENO5_AO_with_DF(Point1d&, Point1d&, Fluid1d*, Block1d) { leaf_called("WENO5_AO_with_DF"); }
void Center_collision(Interface1d&, Fluid1d*) { leaf_called("Center_collision"); }

// Reconstruction_within_cell (reconstruction.cpp:46-53) with the leaf the
// driver assigned: WENO5_AO, or WENO5_AO_with_DF
void Reconstruction_within_cell(Interface1d* interfaces, Fluid1d*, Block1d) {
  gks1d_ctx* ctx = bind1_via(interfaces);
  if (cellreconstruction != WENO5_AO && cellreconstruction != WENO5_AO_with_DF)
    die("Reconstruction_within_cell", "the device path implements WENO5_AO and WENO5_AO_with_DF");
  check1(gks1d_reconstruct(ctx, cellreconstruction == WENO5_AO_with_DF ? 1 : 0), "reconstruct");
}
// the collision centre state is formed inside the device flux sweep
void Reconstruction_forg0(Interface1d*, Fluid1d*, Block1d) {
  if (g0reconstruction != Center_collision) die("Reconstruction_forg0", "the device path implements Center_collision");
}
void Update_alpha(Interface1d* interfaces, Fluid1d*, Block1d) {
  check1(gks1d_update_alpha(bind1_via(interfaces)), "update_alpha");
}

void Calculate_flux(Flux1d** fluxes, Interface1d* interfaces, Block1d& block, int stage) {
  gks1d_ctx* ctx = interfaces ? bind1_via(interfaces) : bind1_via(fluxes);
  check1(gks1d_flux(ctx, stage, block.dt), "flux");
}
void GKS(Flux1d& flux, Interface1d& in, double dt) {
  Block1d b;
  b.nodex = 8;
  b.right = 1.0;
  b.CFL = 0.5;
  gks1d_params p = snapshot1(b);
  double L[6], R[6], C[6];
  std::memcpy(L, in.left.convar, 24);
  std::memcpy(L + 3, in.left.der1, 24);
  std::memcpy(R, in.right.convar, 24);
  std::memcpy(R + 3, in.right.der1, 24);
  check1(gks1d_flux_points(&p, 1, L, R, dt, C, flux.f, flux.derf), "GKS");
  std::memcpy(in.center.convar, C, 24);
  std::memcpy(in.center.der1, C + 3, 24);
}

// S2O4 coefficients (time_advance.cpp:20-30)
void S2O4(Block1d& block) {
  block.stages = 2;
  block.timecoefficient[0][0][0] = 0.5;
  block.timecoefficient[0][0][1] = 1.0 / 8.0;
  block.timecoefficient[1][0][0] = 1.0;
  block.timecoefficient[1][1][0] = 0.0;
  block.timecoefficient[1][0][1] = 1.0 / 6.0;
  block.timecoefficient[1][1][1] = 1.0 / 3.0;
}
void Initial_stages(Block1d& block) {
  std::memset(block.timecoefficient, 0, sizeof(block.timecoefficient));
  if (!timecoe_list) die("Initial_stages", "no time scheme assigned to timecoe_list");
  timecoe_list(block);
}
// Get_CFL (time_advance.cpp:75-91): the reference's sequential scan over the
// current state, on the device
double Get_CFL(Block1d& block, Fluid1d* fluids, double tstop) {
  double dt = 0.0;
  check1(gks1d_cfl(bind1(fluids, block), block.t, tstop, &dt), "Get_CFL");
  cout << "step = " << block.step << " time size is " << dt << " time = " << block.t << endl;
  return dt;
}
void Update(Fluid1d* fluids, Flux1d**, Block1d block, int stage) {
  if (stage > block.stages) die("Update", "wrong middle stage,pls check the time marching setting");
  check1(gks1d_update(bind1(fluids, block), stage), "update");
}

void Download_fluid(Fluid1d* fluids, Block1d block) {
  auto it = g_bound1.find(fluids);
  if (it == g_bound1.end()) return;
  check1(gks1d_get_convar(it->second, &fluids[0].convar[0], sizeof(Fluid1d), sizeof(double)), "download convar");
  check1(gks1d_get_alpha(it->second, &fluids[0].alpha, sizeof(Fluid1d)), "download alpha");
  (void)block;
}
void Upload_fluid(Fluid1d* fluids, Block1d block) {
  check1(gks1d_set_convar(bind1(fluids, block), &fluids[0].convar[0], sizeof(Fluid1d), sizeof(double)),
         "upload convar");
}
void Release_fluid(Fluid1d* fluids) {
  auto it = g_bound1.find(fluids);
  if (it == g_bound1.end()) return;
  gks1d_destroy(it->second);
  g_bound1.erase(it);
  for (auto o = g_owner1.begin(); o != g_owner1.end();) {
    if (o->second == fluids) o = g_owner1.erase(o);
    else ++o;
  }
}

void output1d(Fluid1d* fluids, Block1d block) {
  Download_fluid(fluids, block);
  std::ofstream ofs("../data/R.txt");
  for (int i = block.ghost; i < block.nodex + block.ghost; i++) ofs << fluids[i].convar[0] << endl;
}

// ICforSodTube (1Dproblem.cpp:150-170): (1, 0, 1) where (i - ghost + 1) dx
// <= 0.5, (0.125, 0, 0.1) beyond, from the first interior cell to the last
// ghost cell; convar of every cell
void ICforSodTube(Fluid1d* fluids, Block1d block) {
  for (int i = block.ghost; i <= block.nodex + 2 * block.ghost - 1; i++) {
    const bool left = (double)(i - block.ghost + 1) * block.dx <= 0.5;
    fluids[i].primvar[0] = left ? 1.0 : 0.125;
    fluids[i].primvar[1] = 0.0;
    fluids[i].primvar[2] = left ? 1.0 : 0.1;
  }
  for (int i = 0; i < block.nx; i++) Primvar_to_convar_1D(fluids[i].convar, fluids[i].primvar);
}
