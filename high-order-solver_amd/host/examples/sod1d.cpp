// sod1d.cpp — a 1D problem driver written against the reference's call
// surface (the structure of SodTubeProblem, reference src/1Dproblem.cpp:3-148),
// built against the device shim: the same globals, pointer assignments and
// sweep order, with the stdin prompt replaced by arguments.
//
//   sod1d N tstop max_steps out.bin [df]
//
// Sod tube on [0,1]; with "df" the DF variant of the loop (is_using_df_factor,
// WENO5_AO_with_DF from step 1 on).  Writes t, step and the interior convar
// and alpha (SoA [4][N]) to out.bin.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "1Dproblem.h"

int main(int argc, char** argv) {
  const int N = argc > 1 ? atoi(argv[1]) : 100;
  const double tstop = argc > 2 ? atof(argv[2]) : 0.2;
  const int max_steps = argc > 3 ? atoi(argv[3]) : 1 << 30;
  const char* out = argc > 4 ? argv[4] : "sod1d.bin";
  is_using_df_factor = argc > 5 && strcmp(argv[5], "df") == 0;

  Runtime runtime;
  runtime.start_initial = clock();
  Block1d block;
  block.nodex = N;
  block.ghost = 5;
  block.CFL = 0.5;
  Fluid1d* bcvalue = new Fluid1d[2];
  K = 4;
  Gamma = 1.4;
  R_gas = 1.0;
  gks1dsolver = gks2nd;
  tau_type = Euler;
  c1_euler = 0.05;
  c2_euler = 1.0;
  BoundaryCondition leftboundary = free_boundary_left;
  BoundaryCondition rightboundary = free_boundary_right;
  cellreconstruction = WENO5_AO;
  wenotype = wenoz;
  reconstruction_variable = characteristic;
  g0reconstruction = Center_collision;
  is_reduce_order_warning = true;
  flux_function = GKS;
  timecoe_list = S2O4;
  Initial_stages(block);

  block.nx = block.nodex + 2 * block.ghost;
  block.nodex_begin = block.ghost;
  block.nodex_end = block.nodex + block.ghost - 1;
  Fluid1d* fluids = new Fluid1d[block.nx]();
  Interface1d* interfaces = new Interface1d[block.nx + 1];
  Flux1d** fluxes = Setflux_array(block);
  block.left = 0.0;
  block.right = 1.0;
  block.dx = (block.right - block.left) / block.nodex;
  SetUniformMesh(block, fluids, interfaces, fluxes);
  ICforSodTube(fluids, block);

  block.t = 0;
  block.step = 0;
  runtime.finish_initial = clock();
  while (block.t < tstop && block.step < max_steps) {
    CopyFluid_new_to_old(fluids, block);
    block.dt = Get_CFL(block, fluids, tstop);
    if (block.step > 0 && is_using_df_factor) cellreconstruction = WENO5_AO_with_DF;
    for (int i = 0; i < block.stages; i++) {
      leftboundary(fluids, block, bcvalue[0]);
      rightboundary(fluids, block, bcvalue[1]);
      Reconstruction_within_cell(interfaces, fluids, block);
      Reconstruction_forg0(interfaces, fluids, block);
      Calculate_flux(fluxes, interfaces, block, i);
      Update(fluids, fluxes, block, i);
      if (is_using_df_factor) Update_alpha(interfaces, fluids, block);
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
  for (int v = 0; v < 3; v++)
    for (int i = block.ghost; i < block.ghost + N; i++) fwrite(&fluids[i].convar[v], 8, 1, f);
  for (int i = block.ghost; i < block.ghost + N; i++) fwrite(&fluids[i].alpha, 8, 1, f);
  fclose(f);
  cout << "t = " << block.t << " steps = " << block.step << endl;
  return 0;
}
