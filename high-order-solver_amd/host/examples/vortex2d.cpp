// vortex2d.cpp — a smooth-flow 2D driver written against the reference's
// call surface (the loop of PlanarShock / sinwave_2d, reference
// src/2Dproblem.cpp:109-161, 1462-1524), built against the device shim:
// WENO7-AO or WENO9-AO with the DF factor (3 or 4 Gauss points),
// characteristic variables, Euler tau with c1 = c2 = 0, periodic
// boundaries, S2O4, CFL 0.5 on [0, 10]^2 (the isentropic vortex workload).
//
//   vortex2d N order tstop max_steps in.bin out.bin [fast]
//
// in.bin: the initial interior cell averages, SoA [4][N][N] doubles, i
// outer (the cell averages of the vortex come from the caller).  Writes t,
// step and the interior convar in the same layout to out.bin.  "fast"
// selects the contracted kernel build (gks_fast_kernels).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "2Dproblem.h"

int main(int argc, char** argv) {
  if (argc < 7) {
    cout << "usage: vortex2d N order tstop max_steps in.bin out.bin [fast]" << endl;
    return 2;
  }
  const int N = atoi(argv[1]);
  const int order = atoi(argv[2]);
  const double tstop = atof(argv[3]);
  const int max_steps = atoi(argv[4]);
  const char* in = argv[5];
  const char* out = argv[6];
  gks_fast_kernels = argc > 7 && std::string(argv[7]) == "fast";

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
  tau_type = Euler;
  c1_euler = 0.0;
  c2_euler = 0.0;
  Mu = 0.0;

  BoundaryCondition2d leftboundary = periodic_boundary_left;
  BoundaryCondition2d rightboundary = periodic_boundary_right;
  BoundaryCondition2d downboundary = periodic_boundary_down;
  BoundaryCondition2d upboundary = periodic_boundary_up;

  reconstruction_variable = characteristic;
  wenotype = wenoz;
  if (order == 9) {
    gausspoint = 4;
    cellreconstruction_2D_normal = WENO9_AO_with_df_normal;
    cellreconstruction_2D_tangent = WENO9_AO_with_df_tangent;
  } else {
    gausspoint = 3;
    cellreconstruction_2D_normal = WENO7_AO_with_df_normal;
    cellreconstruction_2D_tangent = WENO7_AO_with_df_tangent;
  }
  SetGuassPoint();
  g0reconstruction_2D_normal = Center_do_nothing_normal;
  g0reconstruction_2D_tangent = Center_all_collision_multi;

  gks2dsolver = gks2nd_2d;
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
  block.right = 10.0;
  block.down = 0.0;
  block.up = 10.0;
  block.dx = (block.right - block.left) / block.nodex;
  block.dy = (block.up - block.down) / block.nodey;
  block.overdx = 1 / block.dx;
  block.overdy = 1 / block.dy;
  SetUniformMesh(block, fluids, xinterfaces, yinterfaces, xfluxes, yfluxes);

  // initial condition: the given cell averages, and their primitive
  // variables for the first Get_CFL
  std::vector<double> u0(4 * (size_t)N * N);
  FILE* fi = fopen(in, "rb");
  if (!fi || fread(u0.data(), sizeof(double), u0.size(), fi) != u0.size()) {
    cout << "cannot read " << in << endl;
    return 1;
  }
  fclose(fi);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      Fluid2d& c = fluids[(i + block.ghost) * block.ny + j + block.ghost];
      for (int v = 0; v < 4; v++) c.convar[v] = u0[((size_t)v * N + i) * N + j];
      Convar_to_primvar_2D(c.primvar, c.convar);
    }

  block.t = 0;
  block.step = 0;
  while (block.t < tstop && block.step < max_steps) {
    CopyFluid_new_to_old(fluids, block);
    block.dt = Get_CFL(block, fluids, tstop);
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
