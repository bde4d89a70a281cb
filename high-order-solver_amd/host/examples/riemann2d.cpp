// riemann2d.cpp — a 2D problem driver written against the reference's call
// surface (the structure of PlanarShock, reference src/2Dproblem.cpp:3-168),
// built against the device shim: the same globals, pointer assignments and
// sweep order, with the stdin prompt replaced by arguments.
//
//   riemann2d N seed tstop max_steps out.bin
//
// Four-quadrant Riemann problem on [0,1]^2 split at (0.5, 0.5); quadrant
// states drawn from std::mt19937_64(seed), zone 1..4, each (rho, u, v, p) in
// [0.1,1.5] x [-0.8,0.8]^2 x [0.02,1.5].  Writes t, step and the interior
// convar (SoA [4][N][N], i outer) to out.bin.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "2Dproblem.h"

int main(int argc, char** argv) {
  const int N = argc > 1 ? atoi(argv[1]) : 64;
  const unsigned long long seed = argc > 2 ? strtoull(argv[2], nullptr, 10) : 0;
  const double tstop = argc > 3 ? atof(argv[3]) : 0.25;
  const int max_steps = argc > 4 ? atoi(argv[4]) : 1 << 30;
  const char* out = argc > 5 ? argv[5] : "riemann2d.bin";

  Runtime runtime;
  runtime.start_initial = clock();
  Block2d block;
  block.uniform = true;
  block.nodex = N;
  block.nodey = N;
  block.ghost = 5;
  block.CFL = 0.5;
  Fluid2d* bcvalue = new Fluid2d[4];

  K = 3;
  Gamma = 1.4;
  is_using_df_factor = true;

  BoundaryCondition2d leftboundary = free_boundary_left;
  BoundaryCondition2d rightboundary = free_boundary_right;
  BoundaryCondition2d downboundary = free_boundary_down;
  BoundaryCondition2d upboundary = free_boundary_up;

  gausspoint = 2;
  SetGuassPoint();
  reconstruction_variable = characteristic;
  wenotype = wenoz;
  cellreconstruction_2D_normal = WENO5_AO_normal;
  cellreconstruction_2D_tangent = WENO5_AO_tangent;
  g0reconstruction_2D_normal = Center_do_nothing_normal;
  g0reconstruction_2D_tangent = Center_all_collision_multi;
  is_reduce_order_warning = true;

  gks2dsolver = gks2nd_2d;
  tau_type = Euler;
  c1_euler = 0.05;
  c2_euler = 1.0;
  flux_function_2d = GKS2D;

  timecoe_list_2d = S2O4_2D;
  Initial_stages(block);

  block.nx = block.nodex + 2 * block.ghost;
  block.ny = block.nodey + 2 * block.ghost;
  Fluid2d* fluids = Setfluid(block);
  Interface2d* xinterfaces = Setinterface_array(block);
  Interface2d* yinterfaces = Setinterface_array(block);
  Flux2d_gauss** xfluxes = Setflux_gauss_array(block);
  Flux2d_gauss** yfluxes = Setflux_gauss_array(block);

  block.left = 0.0;
  block.right = 1.0;
  block.down = 0.0;
  block.up = 1.0;
  block.dx = (block.right - block.left) / block.nodex;
  block.dy = (block.up - block.down) / block.nodey;
  block.overdx = 1 / block.dx;
  block.overdy = 1 / block.dy;
  SetUniformMesh(block, fluids, xinterfaces, yinterfaces, xfluxes, yfluxes);

  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> drho(0.1, 1.5), dvel(-0.8, 0.8), dp(0.02, 1.5);
  double zone[4][4];
  for (int z = 0; z < 4; z++) {
    zone[z][0] = drho(gen);
    zone[z][1] = dvel(gen);
    zone[z][2] = dvel(gen);
    zone[z][3] = dp(gen);
  }
  double discon[]{0.5, 0.5};
  IC_for_riemann_2d(fluids, zone[0], zone[1], zone[2], zone[3], block, discon);

  runtime.finish_initial = clock();
  block.t = 0;
  block.step = 0;
  runtime.start_compute = clock();
  while (block.t < tstop && block.step < max_steps) {
    CopyFluid_new_to_old(fluids, block);
    block.dt = Get_CFL(block, fluids, tstop);
    if (is_using_df_factor) {
      cellreconstruction_2D_normal = WENO5_AO_with_df_normal;
      cellreconstruction_2D_tangent = WENO5_AO_with_df_tangent;
    }
    for (int i = 0; i < block.stages; i++) {
      leftboundary(fluids, block, bcvalue[0]);
      rightboundary(fluids, block, bcvalue[1]);
      downboundary(fluids, block, bcvalue[2]);
      upboundary(fluids, block, bcvalue[3]);
      Convar_to_Primvar(fluids, block);
      Reconstruction_within_cell(xinterfaces, yinterfaces, fluids, block);
      Reconstruction_forg0(xinterfaces, yinterfaces, fluids, block);
      Calculate_flux(xfluxes, yfluxes, xinterfaces, yinterfaces, block, i);
      Update(fluids, xfluxes, yfluxes, block, i);
      if (is_using_df_factor) Update_alpha(xinterfaces, yinterfaces, fluids, block);
    }
    block.step++;
    block.t = block.t + block.dt;
  }
  Download_fluid(fluids, block);
  runtime.finish_compute = clock();

  FILE* f = fopen(out, "wb");
  if (!f) {
    cout << "cannot write " << out << endl;
    return 1;
  }
  const double hdr[2] = {block.t, (double)block.step};
  fwrite(hdr, sizeof(double), 2, f);
  for (int v = 0; v < 4; v++)
    for (int i = block.ghost; i < block.ghost + N; i++)
      for (int j = block.ghost; j < block.ghost + N; j++) fwrite(&fluids[i * block.ny + j].convar[v], 8, 1, f);
  fclose(f);
  cout << "t = " << block.t << " steps = " << block.step << endl;
  return 0;
}
