// gks3d_launch.hpp — device-state descriptor of the 3D path shared by the two
// kernel builds and the C ABI layer, and the launchers each build exports.
#pragma once
#include <hip/hip_runtime.h>

#include "gks3d_device_types.hpp"

// blocks of the primitive-variable sweep (each leaves one partial)
constexpr int kParts3 = 1024;

struct Dev3 {
  int NX, NY, NZ, G;   // interior cells, ghost layers
  int nx, ny, nz;      // with ghosts
  long long nc, nf;    // cells; faces of the (nx+1)(ny+1)(nz+1) grid
  int recon_var;       // 0 conservative, 1 characteristic
  double dx, dy, dz;
  gksd3::Gas3 gas;
  double* U;      // [5][nc] conservative variables, cell (i*ny + j)*nz + k
  double* Uold;   // [5][nc]
  double* fv;     // [2 side][10][nf] face values of the direction in work:
                  //   side 0 the left cell's, side 1 the right cell's (convar, der1x)
  double* lv;     // [2 line][2 side][15][nf] line values (convar, der1x, der1y)
  double* xs[3];  // [5][nf] per direction: the stage combination of the
                  //   face's four Gauss points, summed
  double* pc[3];  // [5][nf] per direction: the stage-0 share of stage 1
  double* part;   // [3][kParts3] partials of the primitive-variable sweep
  double* stat;   // [4]: maximum speed, minimum density, blow-up flag
  // learned path (gks3d_set_classifier; null mlp: plain WENO5-AO)
  const float* mlp;             // classifier parameters (gks_learn.h layout)
  const unsigned short* mlpw;   // its weights in the MFMA operand layout
  float mlp_thr;                // logit below which a stencil takes the linear weights
  unsigned char* lin3;          // [3][nc]: the cell's x / y / z stencil took them (k3_classify)
  unsigned char* lin_line;      // [2 side][nf]: a face side's line values took them (direction in work)
  unsigned long long* lstat3;   // [3 kinds][256 slots][linear, total]: face, line, point reconstructions
};
// slots of lstat3 per kind (the adds of different waves spread over them)
constexpr int kStat3Slots = 256;

struct Launch3dTable {
  hipError_t (*bc)(const Dev3&, const int* bc6, hipStream_t);
  hipError_t (*bc_side)(const Dev3&, int side, int kind, hipStream_t);  // one side's fill
  hipError_t (*primvar)(const Dev3&, hipStream_t);  // sweep + reduction into stat
  hipError_t (*copy_old)(const Dev3&, hipStream_t);
  // one direction's face values, line values, point values + GKS3D + stage sum
  hipError_t (*face)(const Dev3&, int dir, hipStream_t);
  hipError_t (*line)(const Dev3&, int dir, hipStream_t);
  hipError_t (*flux)(const Dev3&, int dir, int stage, double dt, hipStream_t);
  hipError_t (*update)(const Dev3&, hipStream_t);
  hipError_t (*flux_points)(const gksd3::Gas3&, long long n, const double* L, const double* R, double dt,
                            double* C, double* F, double* D, hipStream_t);
  // learned path: every cell's three stencils through the classifier
  hipError_t (*classify)(const Dev3&, hipStream_t);
};

namespace ieee { extern const Launch3dTable table3d; }
namespace fast { extern const Launch3dTable table3d; }
