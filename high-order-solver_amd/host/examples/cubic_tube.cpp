// cubic_tube.cpp — a 3D problem driver written against the reference's call
// surface (the structure of CubicTube, reference src/3Dproblem.cpp:3-172),
// built against the device shim: the same globals, pointer assignments and
// sweep order, with the stdin prompt replaced by arguments.  The reference's
// own loop cannot run (its Gauss-point records are never allocated and its
// GKS3D reads the pressure as lambda); through the shim the same loop runs
// the 3D scheme of include/gks3d.h on the GPU.
//
//   cubic_tube N tstop max_steps out.bin [periodic]
//
// Cubic Sod problem on [0,1]^3: (rho, u, v, w, p) = (1, 0, 0, 0, 1) in the
// central cube of half-width 0.2, (0.125, 0, 0, 0, 0.1) outside; free
// boundaries (periodic with the fifth argument).  Writes t, step and the
// interior convar (SoA [5][N][N][N], i outer) to out.bin.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "3Dproblem.h"

int main(int argc, char** argv) {
  const int N = argc > 1 ? atoi(argv[1]) : 20;
  const double tstop = argc > 2 ? atof(argv[2]) : 0.2;
  const int max_steps = argc > 3 ? atoi(argv[3]) : 1 << 30;
  const char* out = argc > 4 ? argv[4] : "cubic_tube.bin";
  const bool periodic = argc > 5 && strcmp(argv[5], "periodic") == 0;

  Runtime runtime;
  runtime.start_initial = clock();
  Block3d block;
  block.uniform = true;
  block.nodex = N;
  block.nodey = N;
  block.nodez = N;
  block.ghost = 3;
  block.CFL = 0.5;
  Fluid3d* bcvalue = new Fluid3d[6];
  K = 2;
  Gamma = 1.4;
  gks3dsolver = gks2nd_3d;
  tau_type = Euler;
  c1_euler = 0.05;
  c2_euler = 1;

  BoundaryCondition3d leftboundary = periodic ? periodic_boundary_xleft : free_boundary_xleft;
  BoundaryCondition3d rightboundary = periodic ? periodic_boundary_xright : free_boundary_xright;
  BoundaryCondition3d downboundary = periodic ? periodic_boundary_yleft : free_boundary_yleft;
  BoundaryCondition3d upboundary = periodic ? periodic_boundary_yright : free_boundary_yright;
  BoundaryCondition3d backboundary = periodic ? periodic_boundary_zleft : free_boundary_zleft;
  BoundaryCondition3d frontboundary = periodic ? periodic_boundary_zright : free_boundary_zright;

  gausspoint = 4;
  SetGuassPoint_3D();
  reconstruction_variable = characteristic;
  wenotype = wenoz;
  cellreconstruction_3D_of_face_value = WENO5_AO_splitting_3d;
  cellreconstruction_3D_of_line_value = WENO5_AO_of_line_value;
  cellreconstruction_3D_of_point_value = WENO5_AO_of_point_value;
  g0reconstruction_3D_of_face_value = Do_nothing_splitting_3d;
  g0reconstruction_3D_of_line_value = Do_nothing_reconstruction_of_line_value;
  g0reconstruction_3D_of_point_value = Do_nothing_reconstruction_of_point_value;
  gks3dsolver = gks2nd_3d;
  g0type = all_collisionn;
  flux_function_3d = GKS3D;
  timecoe_list_3d = S2O4_3D;
  Initial_stages(block);

  block.nx = block.nodex + 2 * block.ghost;
  block.ny = block.nodey + 2 * block.ghost;
  block.nz = block.nodez + 2 * block.ghost;
  Fluid3d* fluids = Setfluid_array(block);
  Interface3d* xinterfaces = Setinterface_array(block);
  Interface3d* yinterfaces = Setinterface_array(block);
  Interface3d* zinterfaces = Setinterface_array(block);
  Flux3d_gauss** xfluxes = Setflux_gauss_array(block);
  Flux3d_gauss** yfluxes = Setflux_gauss_array(block);
  Flux3d_gauss** zfluxes = Setflux_gauss_array(block);

  block.left = 0.0;
  block.right = 1.0;
  block.down = 0.0;
  block.up = 1.0;
  block.back = 0.0;
  block.front = 1.0;
  block.dx = (block.right - block.left) / block.nodex;
  block.dy = (block.up - block.down) / block.nodey;
  block.dz = (block.front - block.back) / block.nodez;
  block.xarea = block.dy * block.dz;
  block.yarea = block.dx * block.dz;
  block.zarea = block.dx * block.dy;
  block.overdx = 1 / block.dx;
  block.overdy = 1 / block.dy;
  block.overdz = 1 / block.dz;
  SetUniformMesh(block, fluids, xinterfaces, yinterfaces, zinterfaces);

  double zone1[5] = {1.0, 0, 0, 0, 1.0};
  double zone2[5] = {0.125, 0, 0, 0, 0.1};
  ICfor_cubic_sod(fluids, zone1, zone2, block);
  runtime.finish_initial = clock();

  block.t = 0;
  block.step = 0;
  runtime.start_compute = clock();
  while (block.t < tstop && block.step < max_steps) {
    CopyFluid_new_to_old(fluids, block);
    block.dt = Get_CFL(block, fluids, tstop);
    for (int i = 0; i < block.stages; i++) {
      leftboundary(fluids, block, bcvalue[0]);
      rightboundary(fluids, block, bcvalue[1]);
      downboundary(fluids, block, bcvalue[2]);
      upboundary(fluids, block, bcvalue[3]);
      backboundary(fluids, block, bcvalue[4]);
      frontboundary(fluids, block, bcvalue[5]);
      Convar_to_primvar(fluids, block);
      Reconstruction_within_cell(xinterfaces, yinterfaces, zinterfaces, fluids, block);
      Reconstruction_forg0(xinterfaces, yinterfaces, zinterfaces, fluids, block);
      Calculate_flux(xfluxes, yfluxes, zfluxes, xinterfaces, yinterfaces, zinterfaces, block, i);
      Update_gauss(fluids, xfluxes, yfluxes, zfluxes, block, i);
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
  for (int v = 0; v < 5; v++)
    for (int i = block.ghost; i < block.ghost + N; i++)
      for (int j = block.ghost; j < block.ghost + N; j++)
        for (int k = block.ghost; k < block.ghost + N; k++)
          fwrite(&fluids[((size_t)i * block.ny + j) * block.nz + k].convar[v], 8, 1, f);
  fclose(f);
  cout << "t = " << block.t << " steps = " << block.step << endl;
  return 0;
}
