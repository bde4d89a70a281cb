// gks1d_kernels.hpp — the 1D GKS path as one persistent workgroup.
//
// A 1D tube of a few hundred cells holds a few hundred faces of work per
// phase: far too little to fill 256 CUs, and a launch per phase would cost
// more than the phase.  So one workgroup of kBlock1 threads runs the whole
// time loop of the reference driver (1Dproblem.cpp:79-148): copy old, CFL,
// then per S2O4 stage the boundary fills, the reconstruction, the collision
// state + GKS flux, the update and the DF sensor, phases separated by
// workgroup barriers.  State lives in global memory (a few tens of KB, held
// in the CU's L1/L2); the CFL speeds go through LDS so that one lane can
// replay the reference's sequential Dtx scan exactly.
//
// Device layout (FP64, SoA): U/Uold [3][nx], alpha [nx], face [2][6][nx+1],
// fl [2][6][nx+1] (see gks1d_launch.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "gks1d_device.hpp"
#include "gks1d_launch.hpp"

#ifndef GKS_VARIANT
#error "GKS_VARIANT must name the kernel namespace"
#endif

namespace GKS_VARIANT {

using namespace gksd;

constexpr int kBlock1 = 512;
constexpr int kSpdChunk = 2048;  // LDS slab of CFL speeds (16 KB)

// boundary_condition.cpp:4-64 for one ghost cell; sources are interior cells
// (nodex >= ghost), so the sequential fill of the reference reduces to these
// direct copies: free repeats the first/last interior cell, periodic wraps by
// nodex, reflective mirrors and negates the momentum.
__device__ __forceinline__ void bc_cell1(const Dev1& d, int k, int bcl, int bcr) {
  const int side = k < d.G ? 0 : 1;
  const int kind = side == 0 ? bcl : bcr;
  const int i = side == 0 ? k : d.N + k;  // k in [0,G) left, [G,2G) right
  int src;
  if (kind == 0) src = side == 0 ? d.G : d.N + d.G - 1;
  else if (kind == 1) src = side == 0 ? i + d.N : i - d.N;
  else src = side == 0 ? 2 * d.G - 1 - i : 2 * (d.nx - d.G) - 1 - i;
#pragma unroll
  for (int v = 0; v < 3; v++) {
    double val = d.U[v * d.nx + src];
    if (kind == 2 && v == 1) val = -val;
    d.U[v * d.nx + i] = val;
  }
}

// ---- the phases of one S2O4 step, each run by one workgroup (tid of nthr)

// the CFL speeds of the interior and the reference's sequential scan (Get_CFL /
// Dtx, time_advance.cpp:75-111); with copy_old, CopyFluid_new_to_old in the
// same pass.  The scan's result is valid in lane 0.
__device__ double cfl_scan1(const Dev1& d, bool copy_old, double* s_spd, int tid) {
  const int nx = d.nx;
  const Gas& g = d.gas;
  double dscan = d.dx, q = 0.0;
  if (tid == 0) q = d.cfl_dx / dscan;
  for (int base = 0; base < d.N; base += kSpdChunk) {
    const int m = d.N - base < kSpdChunk ? d.N - base : kSpdChunk;
    for (int k = tid; k < m; k += kBlock1) {
      const int i = d.G + base + k;
      double c[3], p[3];
#pragma unroll
      for (int v = 0; v < 3; v++) {
        c[v] = d.U[v * nx + i];
        if (copy_old) d.Uold[v * nx + i] = c[v];
      }
      prim1(p, c, g);
      s_spd[k] = fabs(p[1]) + sqrt(g.gamma * p[2] / p[0]);
    }
    __syncthreads();
    if (tid == 0) {
      // the reference's scan, cell by cell; CFL*dx/dtx is re-formed only
      // when dtx changes (same value as forming it every time)
      for (int k = 0; k < m; k++) {
        const double tmp = s_spd[k];
        bool changed = false;
        if (tmp > q) {
          dscan = d.cfl_dx / tmp;
          changed = true;
        }
        if (dscan > d.visc_dt && g.tau_type == 1 && g.mu > 0) {
          dscan = d.visc_dt;
          changed = true;
        }
        if (changed) q = d.cfl_dx / dscan;
      }
    }
    __syncthreads();
  }
  return dscan;
}
// Reconstruction_within_cell: cells G-1 .. nx-G (reconstruction.cpp:46-53)
__device__ void recon1(const Dev1& d, bool with_df, int tid) {
  const int nx = d.nx, nf = d.nx + 1;
  for (int k = tid; k < d.N + 2; k += kBlock1) {
    const int i = d.G - 1 + k;
    double w[5][3], a[5];
#pragma unroll
    for (int s = 0; s < 5; s++) {
#pragma unroll
      for (int v = 0; v < 3; v++) w[s][v] = d.U[v * nx + i - 2 + s];
      a[s] = d.alpha[i - 2 + s];
    }
    double lo[6], ro[6];
    weno5_cell1(lo, ro, w, a, with_df, d.dx, d.idx, d.recon_var, d.gas);
#pragma unroll
    for (int v = 0; v < 6; v++) {
      d.face[(6 + v) * nf + i] = lo[v];  // interfaces[i].right
      d.face[v * nf + i + 1] = ro[v];    // interfaces[i+1].left
    }
  }
}
// Reconstruction_forg0 (Center_collision) + Calculate_flux: faces G .. N+G
__device__ void flux1(const Dev1& d, int stage, double dt, int tid) {
  const int nf = d.nx + 1;
  for (int k = tid; k < d.N + 1; k += kBlock1) {
    const int i = d.G + k;
    double L[6], R[6], f[3], df[3];
#pragma unroll
    for (int v = 0; v < 6; v++) {
      L[v] = d.face[v * nf + i];
      R[v] = d.face[(6 + v) * nf + i];
    }
    gks1_point(f, df, nullptr, L, R, dt, d.gas);
    double* o = d.fl + (size_t)stage * 6 * nf;
#pragma unroll
    for (int v = 0; v < 3; v++) {
      o[v * nf + i] = f[v];
      o[(3 + v) * nf + i] = df[v];
    }
  }
}
// Update, S2O4 branch (time_advance.cpp:167-192), of interior cell i
__device__ void update_cell1(const Dev1& d, int i, int stage) {
  const int nx = d.nx, nf = d.nx + 1;
  double F[2][3];
#pragma unroll
  for (int e = 0; e < 2; e++) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int fi = i + e;
      double acc = 0.0;
      // S2O4 weights: stage 0 {0.5, 1/8}; stage 1 {1, 1/6} and {0, 1/3}
      if (stage == 0) {
        acc = acc + 0.5 * d.fl[j * nf + fi] + (1.0 / 8.0) * d.fl[(3 + j) * nf + fi];
      } else {
        acc = acc + 1.0 * d.fl[j * nf + fi] + (1.0 / 6.0) * d.fl[(3 + j) * nf + fi];
        acc = acc + 0.0 * d.fl[(6 + j) * nf + fi] + (1.0 / 3.0) * d.fl[(9 + j) * nf + fi];
      }
      F[e][j] = acc;
    }
  }
#pragma unroll
  for (int j = 0; j < 3; j++) d.U[j * nx + i] = d.Uold[j * nx + i] + d.idx * (F[0][j] - F[1][j]);
}
// Update_alpha (reconstruction.cpp:374-392) of interior cell i, from the
// face states of its two faces
__device__ void alpha_cell1(const Dev1& d, int i) {
  const int nf = d.nx + 1;
  const Gas& g = d.gas;
  double cl[3], cr[3], pll[3], plr[3], prl[3], prr[3];
#pragma unroll
  for (int v = 0; v < 3; v++) {
    cl[v] = d.face[v * nf + i];
    cr[v] = d.face[(6 + v) * nf + i];
  }
  prim1(pll, cl, g);
  prim1(plr, cr, g);
#pragma unroll
  for (int v = 0; v < 3; v++) {
    cl[v] = d.face[v * nf + i + 1];
    cr[v] = d.face[(6 + v) * nf + i + 1];
  }
  prim1(prl, cl, g);
  prim1(prr, cr, g);
  double al = alpha_k1(pll, plr, g), ar = alpha_k1(prl, prr, g);
  al = al < 2.0 ? 0.0 : al;
  ar = ar < 2.0 ? 0.0 : ar;
  d.alpha[i] = al + ar;
}

__global__ void __launch_bounds__(kBlock1) k_run1d(Dev1 d, int bcl, int bcr, double tstop, double t0, int step0,
                                                   int max_n, double* dt_out) {
  __shared__ double s_spd[kSpdChunk];
  __shared__ double s_dt;
  const int tid = threadIdx.x;
  double t = t0;
  int step = step0, n = 0;
  while (t < tstop && n < max_n) {
    // CopyFluid_new_to_old + the CFL of the current state
    double dscan = cfl_scan1(d, true, s_spd, tid);
    if (tid == 0) {
      if (t + dscan > tstop) dscan = tstop - t + 1e-15;
      s_dt = dscan;
      dt_out[n] = dscan;
    }
    __syncthreads();
    const double dt = s_dt;
    const bool with_df = d.df_from_start || (d.use_df && step > 0);
    for (int stage = 0; stage < 2; stage++) {
      // leftboundary / rightboundary
      if (tid < 2 * d.G) bc_cell1(d, tid, bcl, bcr);
      __syncthreads();
      recon1(d, with_df, tid);
      __syncthreads();
      flux1(d, stage, dt, tid);
      __syncthreads();
      // Update and Update_alpha (the alpha of cell i reads only face states)
      for (int k = tid; k < d.N; k += kBlock1) {
        update_cell1(d, d.G + k, stage);
        if (d.use_df) alpha_cell1(d, d.G + k);
      }
      __syncthreads();
    }
    step++;
    t = t + dt;
    n++;
  }
  if (tid == 0) {
    d.st[0] = t;
    d.st[1] = (double)step;
    d.st[2] = (double)n;
  }
}

// one phase of the step for a driver that keeps the reference's loop on the
// host (gks1d_boundary ... gks1d_update_alpha); one workgroup
__global__ void __launch_bounds__(kBlock1) k_phase1d(Dev1 d, int phase, int a, int b, int stage, double dt, double t,
                                                     double tstop) {
  __shared__ double s_spd[kSpdChunk];
  const int tid = threadIdx.x;
  if (phase == PH1_BC) {
    // one side (a: 0 left, 1 right) of kind b
    if (tid < d.G) bc_cell1(d, a == 0 ? tid : d.G + tid, b, b);
  } else if (phase == PH1_COPY_OLD) {
    for (int k = tid; k < d.N; k += kBlock1)
#pragma unroll
      for (int v = 0; v < 3; v++) d.Uold[v * d.nx + d.G + k] = d.U[v * d.nx + d.G + k];
  } else if (phase == PH1_CFL) {
    double dscan = cfl_scan1(d, false, s_spd, tid);
    if (tid == 0) {
      if (t + dscan > tstop) dscan = tstop - t + 1e-15;
      d.st[3] = dscan;
    }
  } else if (phase == PH1_RECON) {
    recon1(d, a != 0, tid);
  } else if (phase == PH1_FLUX) {
    flux1(d, stage, dt, tid);
  } else if (phase == PH1_UPDATE) {
    for (int k = tid; k < d.N; k += kBlock1) update_cell1(d, d.G + k, stage);
  } else if (phase == PH1_ALPHA) {
    for (int k = tid; k < d.N; k += kBlock1) alpha_cell1(d, d.G + k);
  }
}

// Center_collision + GKS for independent points (leaf parity)
__global__ void __launch_bounds__(256) k_flux_points1d(Gas g, long long n, const double* L, const double* R,
                                                       double dt, double* C, double* F, double* D) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double l[6], r[6], c[6], f[3], df[3];
#pragma unroll
  for (int v = 0; v < 6; v++) {
    l[v] = L[k * 6 + v];
    r[v] = R[k * 6 + v];
  }
  gks1_point(f, df, c, l, r, dt, g);
#pragma unroll
  for (int v = 0; v < 6; v++) C[k * 6 + v] = c[v];
#pragma unroll
  for (int v = 0; v < 3; v++) {
    F[k * 3 + v] = f[v];
    D[k * 3 + v] = df[v];
  }
}

#if !defined(__HIP_DEVICE_COMPILE__)
static hipError_t run1d(const Dev1& d, int bcl, int bcr, double tstop, double t, int step, int steps,
                        double* dt_out, hipStream_t s) {
  hipLaunchKernelGGL(k_run1d, dim3(1), dim3(kBlock1), 0, s, d, bcl, bcr, tstop, t, step, steps, dt_out);
  return hipGetLastError();
}
static hipError_t flux_points1d(const Gas& g, long long n, const double* L, const double* R, double dt, double* C,
                                double* F, double* D, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_points1d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, n, L, R, dt, C, F, D);
  return hipGetLastError();
}

static hipError_t phase1d(const Dev1& d, int phase, int a, int b, int stage, double dt, double t, double tstop,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_phase1d, dim3(1), dim3(kBlock1), 0, s, d, phase, a, b, stage, dt, t, tstop);
  return hipGetLastError();
}

const Launch1dTable table1d = {run1d, flux_points1d, phase1d};
#endif

}  // namespace GKS_VARIANT
