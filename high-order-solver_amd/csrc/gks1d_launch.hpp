// gks1d_launch.hpp — device-state descriptor of the 1D path shared by the two
// kernel builds and the C ABI layer, and the launchers each build exports.
#pragma once
#include <hip/hip_runtime.h>

#include "gks2d_types.hpp"

struct Dev1 {
  int N, G, nx;             // interior cells, ghost layers, nx = N + 2G
  int recon_var, use_df, df_from_start;
  double dx, idx, cfl;      // idx = 1/dx
  double cfl_dx, visc_dt;   // CFL*dx and 0.25*CFL*dx*dx/Mu, as Dtx forms them
  gksd::Gas gas;
  double* U;      // [3][nx]
  double* Uold;   // [3][nx]
  double* alpha;  // [nx], ghost cells stay 0
  double* face;   // [2 side][6][nx+1]: side 0 = interfaces[i].left, 1 = .right
  double* fl;     // [2 stage][6][nx+1]: f[3], derf[3]
  double* st;     // [4]: t, step, steps run by the last launch, dt of the last CFL phase
};

// the phases k_phase1d runs one at a time
enum { PH1_BC = 0, PH1_COPY_OLD, PH1_CFL, PH1_RECON, PH1_FLUX, PH1_UPDATE, PH1_ALPHA };

// steps one launch of the persistent kernel may take before it returns to
// the host (bounds a launch to well under a second at any size)
constexpr int kStepsPerLaunch1d = 4096;

struct Launch1dTable {
  // runs up to `steps` steps of the SodTube loop from (t, step); dt of each
  // step goes to dt_out[k] (device, kStepsPerLaunch1d entries)
  hipError_t (*run)(const Dev1&, int bc_left, int bc_right, double tstop, double t, int step, int steps,
                    double* dt_out, hipStream_t);
  hipError_t (*flux_points)(const gksd::Gas&, long long n, const double* L, const double* R, double dt,
                            double* C, double* F, double* D, hipStream_t);
  // one phase (PH1_*): a, b = side and kind (BC) or with_df (RECON)
  hipError_t (*phase)(const Dev1&, int phase, int a, int b, int stage, double dt, double t, double tstop,
                      hipStream_t);
};

namespace ieee { extern const Launch1dTable table1d; }
namespace fast { extern const Launch1dTable table1d; }
