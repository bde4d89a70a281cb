// gks2d_kernels.hpp — HIP kernels of the 2D GKS hot path and their launchers.
//
// Included by kernels_ieee.hip (GKS_FAST 0, -ffp-contract=off) and
// kernels_fast.hip (GKS_FAST 1, -ffp-contract=fast-honor-pragmas); GKS_VARIANT names the
// namespace of each build.
//
// Device layout (all FP64, SoA, y fastest as in the reference):
//   U, Uold      [4][nx*ny]          conservative state, cell i*ny+j
//   ax, ay       [nx*ny]             DF shock sensor (ghost cells stay 0)
//   line_x       [2 side][8][nfx]    x-face line states (convar, der1x);
//                                    x face (i,j) = i*ny+j, i in [0,nx]
//   line_y       [2][8][nfy]         y face (i,j) = i*(ny+1)+j, j in [0,ny]
//   gst_x/_y     [2][gp][12][nf]     Gauss-point states (convar,der1x,der1y)
//   xcx/xcy      [gp][4][nf]         stage-combined flux of the current stage,
//                                    weighted and rotated to the global frame
//   pcx/pcy      [gp][4][nf]         stage-1 share of the stage-0 flux
//   flx/fly[s]   [gp][8][nf]         f[4], derf[4] of stage s (face frame),
//                                    stored only when keep_face_fluxes is set
//   akx/aky      [gp][nf]            Calculate_alpha_k per Gauss point
// side 0 = state left of the face ("line.left"), 1 = right ("line.right").
#pragma once
#include <hip/hip_runtime.h>

#include "gks2d_device.hpp"
#include "gks2d_launch.hpp"
#include "gks2d_weno79_device.hpp"
#include "gks2d_learn_device.hpp"

#ifndef GKS_VARIANT
#error "GKS_VARIANT must name the kernel namespace"
#endif

namespace GKS_VARIANT {

using namespace gksd;

constexpr int kBlock = kSweepBlock;
constexpr int kPartialBlocks = kSweepPartials;

__device__ __forceinline__ long long gtid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

// The tile a workgroup takes: workgroups go to the eight XCDs round robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  Giving XCD x the
// x-th contiguous eighth of the tiles keeps the stencil halos of
// neighbouring tiles in one L2, and spreads a cost pattern that repeats
// every 8k tiles (a troubled band across every grid column of 16 tiles, in
// the learned reconstructions) over all eight XCDs instead of one.
#ifndef GKS_XCD_REMAP
#define GKS_XCD_REMAP 1
#endif
__device__ __forceinline__ unsigned xcd_block() {
#if GKS_XCD_REMAP
  const unsigned nb = gridDim.x, b = blockIdx.x, q = nb / 8, r = nb % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
#else
  return blockIdx.x;
#endif
}

// ------------------------------------------------------------------ boundary
// free / periodic / reflective fills of one side, in the reference's
// semantics (boundary_condition.cpp:66-186, 367-442): left/right cover all
// rows, down/up all columns (including the ghost columns just filled).
__device__ __forceinline__ void bc_body(const Dev& d, unsigned blk, int side, int kind) {
  long long t = (long long)blk * kBlock + threadIdx.x;
  const int len = side < 2 ? d.ny : d.nx;
  if (t >= (long long)d.G * len) return;
  const int g = (int)(t / len), k = (int)(t % len);
  int dst, src, neg;
  if (side < 2) {
    const int i = side == 0 ? d.G - 1 - g : d.nx - d.G + g;
    int si;
    if (kind == 0) si = side == 0 ? d.G : d.nx - d.G - 1;
    else if (kind == 1) si = side == 0 ? i + d.NX : i - d.NX;
    else si = side == 0 ? 2 * d.G - 1 - i : 2 * d.nx - 2 * d.G - 1 - i;
    dst = i * d.ny + k;
    src = si * d.ny + k;
    neg = 1;
  } else {
    const int j = side == 2 ? d.G - 1 - g : d.ny - d.G + g;
    int sj;
    if (kind == 0) sj = side == 2 ? d.G : d.ny - d.G - 1;
    else if (kind == 1) sj = side == 2 ? j + d.NY : j - d.NY;
    else sj = side == 2 ? 2 * d.G - 1 - j : 2 * d.ny - 2 * d.G - 1 - j;
    dst = k * d.ny + j;
    src = k * d.ny + sj;
    neg = 2;
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double val = d.U[v * d.nc + src];
    if (kind == 2 && v == neg) val = -val;
    d.U[v * d.nc + dst] = val;
  }
}
__global__ void __launch_bounds__(kBlock) k_bc(Dev d, int side, int kind) { bc_body(d, blockIdx.x, side, kind); }
// the four fills above in one launch (kinds: left, right, down, up).  The
// left/right fills write the ghost rows from interior rows (all columns),
// the down/up fills then the ghost columns of every row from interior
// columns (the ghost rows' included), so every ghost cell is a copy of one
// interior cell with component 1 negated by a reflective row fill and
// component 2 by a reflective column fill: each cell is formed from its
// interior source directly (the same values; needs NX, NY >= G, so that
// the periodic sources are interior cells).  A thread per ghost cell: the
// 2G ghost rows, then the 2G ghost columns of the interior rows.
__global__ void __launch_bounds__(kBlock) k_bc_all(Dev d, int k0, int k1, int k2, int k3) {
  const long long t = gtid();
  const long long nrow = 2LL * d.G * d.ny;
  if (t >= nrow + 2LL * d.G * d.NX) return;
  int i, j;
  if (t < nrow) {
    const int g = (int)(t / d.ny);
    i = g < d.G ? g : d.nx - 2 * d.G + g;
    j = (int)(t % d.ny);
  } else {
    const long long u = t - nrow;
    const int g = (int)(u % (2 * d.G));
    i = d.G + (int)(u / (2 * d.G));
    j = g < d.G ? g : d.ny - 2 * d.G + g;
  }
  int si = i, sj = j;
  bool neg1 = false, neg2 = false;
  if (i < d.G || i >= d.nx - d.G) {
    const bool lo = i < d.G;
    const int kind = lo ? k0 : k1;
    if (kind == 0) si = lo ? d.G : d.nx - d.G - 1;
    else if (kind == 1) si = lo ? i + d.NX : i - d.NX;
    else si = lo ? 2 * d.G - 1 - i : 2 * d.nx - 2 * d.G - 1 - i;
    neg1 = kind == 2;
  }
  if (j < d.G || j >= d.ny - d.G) {
    const bool lo = j < d.G;
    const int kind = lo ? k2 : k3;
    if (kind == 0) sj = lo ? d.G : d.ny - d.G - 1;
    else if (kind == 1) sj = lo ? j + d.NY : j - d.NY;
    else sj = lo ? 2 * d.G - 1 - j : 2 * d.ny - 2 * d.G - 1 - j;
    neg2 = kind == 2;
  }
  const long long dst = (long long)i * d.ny + j, src = (long long)si * d.ny + sj;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double val = d.U[v * d.nc + src];
    if ((v == 1 && neg1) || (v == 2 && neg2)) val = -val;
    d.U[v * d.nc + dst] = val;
  }
}

// ------------------------------------------------------- primvar + CFL speed
// Convar_to_Primvar (basic_function.cpp:752-773): blow-up test on all cells;
// plus the Dtx speed sqrt(u^2+v^2)+c (time_advance.cpp:292-305) maxed over
// the interior, which is what the next Get_CFL reads.  from_prim: take the
// primitive variables from `prim` instead (step 0, initial condition).
__device__ __forceinline__ void primvar_body(const Dev& d, unsigned blk, unsigned nblk, const double* prim,
                                             double* partial) {
  __shared__ double smax[kBlock];
  __shared__ int sbad[kBlock];
  double tmax = 0.0;
  int bad = 0;
  for (long long c = (long long)blk * kBlock + threadIdx.x; c < d.nc; c += (long long)nblk * kBlock) {
    double p[4];
    if (prim) {
#pragma unroll
      for (int v = 0; v < 4; v++) p[v] = prim[v * d.nc + c];
    } else {
      double u[4];
#pragma unroll
      for (int v = 0; v < 4; v++) u[v] = d.U[v * d.nc + c];
      convar_to_primvar(p, u, d.gas);
      if (p[0] < 0 || p[3] < 0 || p[0] != p[0] || p[3] != p[3]) bad = 1;
    }
    const int i = (int)(c / d.ny), j = (int)(c % d.ny);
    if (i >= d.G && i < d.nx - d.G && j >= d.G && j < d.ny - d.G) {
      double tmp = sqrt(p[1] * p[1] + p[2] * p[2]) + sqrt(d.gas.gamma * p[3] / p[0]);
      if (tmp > tmax) tmax = tmp;  // a NaN speed is skipped, as Dtx's comparison does
    }
  }
  smax[threadIdx.x] = tmax;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = smax[threadIdx.x + s];
      if (o > smax[threadIdx.x]) smax[threadIdx.x] = o;
      sbad[threadIdx.x] |= sbad[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blk] = smax[0];
    partial[2 * blk + 1] = sbad[0] ? 1.0 : 0.0;
  }
}
__global__ void __launch_bounds__(kBlock) k_primvar(Dev d, const double* prim, double* partial) {
  primvar_body(d, blockIdx.x, gridDim.x, prim, partial);
}
// final max / or over the partials into stat[slot*2 .. slot*2+1]; the flag
// accumulates (cleared by the host when it reads it)
__device__ __forceinline__ void reduce_body(const double* partial, int n, double* stat, int slot, int flag = 4) {
  __shared__ double smax[kBlock];
  __shared__ double sbad[kBlock];
  // eight partials per thread and round, their loads issued together (the
  // update's reduction has a partial per 256 cells: 16384 at 2048^2)
  constexpr int U = 8;
  double mu[U], bu[U];
#pragma unroll
  for (int u = 0; u < U; u++) mu[u] = bu[u] = 0.0;
  int k = threadIdx.x;
  for (; k + (U - 1) * kBlock < n; k += U * kBlock) {
    double o[U], f[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      o[u] = partial[2 * (k + u * kBlock)];
      f[u] = partial[2 * (k + u * kBlock) + 1];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (o[u] > mu[u]) mu[u] = o[u];
      if (f[u] != 0.0) bu[u] = 1.0;
    }
  }
  for (; k < n; k += kBlock) {
    const double o = partial[2 * k];
    if (o > mu[0]) mu[0] = o;
    if (partial[2 * k + 1] != 0.0) bu[0] = 1.0;
  }
  double m = mu[0], b = bu[0];
#pragma unroll
  for (int u = 1; u < U; u++) {
    if (mu[u] > m) m = mu[u];
    if (bu[u] != 0.0) b = 1.0;
  }
  smax[threadIdx.x] = m;
  sbad[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      double o = smax[threadIdx.x + s];
      if (o > smax[threadIdx.x]) smax[threadIdx.x] = o;
      if (sbad[threadIdx.x + s] != 0.0) sbad[threadIdx.x] = 1.0;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (slot >= 0) stat[2 * slot] = smax[0];  // (slot < 0: the blow-up test alone)
    if (sbad[0] != 0.0) stat[flag] = 1.0;     // sticky blow-up flag
  }
}
__global__ void __launch_bounds__(kBlock) k_reduce(const double* partial, int n, double* stat, int slot, int flag) {
  reduce_body(partial, n, stat, slot, flag);
}

// CopyFluid_new_to_old (basic_function.cpp:1571-1584), interior cells
__device__ __forceinline__ void copy_old_body(const Dev& d, unsigned blk) {
  long long t = (long long)blk * kBlock + threadIdx.x;
  if (t >= (long long)d.NX * d.NY) return;
  const int i = (int)(t / d.NY) + d.G, j = (int)(t % d.NY) + d.G;
  const long long c = (long long)i * d.ny + j;
#pragma unroll
  for (int v = 0; v < 4; v++) d.Uold[v * d.nc + c] = d.U[v * d.nc + c];
}
__global__ void __launch_bounds__(kBlock) k_copy_old(Dev d) { copy_old_body(d, blockIdx.x); }
// the ghost frame (every cell outside the interior) of `from` into d.U
__global__ void __launch_bounds__(kBlock) k_copy_ghosts(Dev d, const double* from) {
  const long long c = gtid();
  if (c >= d.nc) return;
  const int i = (int)(c / d.ny), j = (int)(c % d.ny);
  if (i >= d.G && i < d.nx - d.G && j >= d.G && j < d.ny - d.G) return;
#pragma unroll
  for (int v = 0; v < 4; v++) d.U[v * d.nc + c] = from[v * d.nc + c];
}

// ------------------------------------------------------ normal reconstruction
// WENO5_AO_with_df (reconstruction.cpp:1731-1859) for one cell along DIR;
// the stencil is already in the x-like frame (YchangetoX for y).
__device__ __forceinline__ void weno5_normal(double* lo, double* ro, const double (*w)[4], const double* alpha,
                                             double h, double ih, int recon_var, const Gas& g) {
  double df[4];
  df5(df, alpha);
  double pn1[4], p0[4], pp1[4];
  convar_to_primvar(pn1, w[1], g);
  convar_to_primvar(p0, w[2], g);
  convar_to_primvar(pp1, w[3], g);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double base[4];
#pragma unroll
    for (int i = 0; i < 4; i++) base[i] = side == 0 ? 0.5 * (pn1[i] + p0[i]) : 0.5 * (pp1[i] + p0[i]);
    double c[5][4];
    CharBase cb;
    if (recon_var == 0) {
#pragma unroll
      for (int s = 0; s < 5; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) c[s][i] = w[s][i];
    } else {
      char_base(cb, base, g);
#pragma unroll
      for (int s = 0; s < 5; s++) to_char(c[s], cb, w[s]);
    }
    double var[4], d1[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (side == 0) weno5_left(var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h, ih);
      else weno5_right(var[i], d1[i], c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], df, h, ih);
    }
    double* out = side == 0 ? lo : ro;
    if (recon_var == 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        out[i] = var[i];
        out[4 + i] = d1[i];
      }
    } else {
      from_char(out, cb, var, g);
      from_char(out + 4, cb, d1, g);
    }
  }
  const bool rr = reduce_order(ro, g);
  const bool rl = reduce_order(lo, g);
  if (rl || rr) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
      ro[m] = w[2][m];
      lo[m] = w[2][m];
      ro[4 + m] = 0.0;
      lo[4 + m] = 0.0;
    }
  }
}

// DIR 0: cells i in [G-1, nx-G], all j; writes x faces i (side 1) and i+1
// (side 0).  DIR 1: cells j in [G-1, ny-G], all i; y faces j and j+1.
// (reconstruction.cpp:1415-1425 with the guards of :1703, :1715)
#ifndef GKS_NORMAL_WAVES
#define GKS_NORMAL_WAVES 1
#endif
template <int DIR> __device__ __forceinline__ void normal_body(const Dev& d, unsigned blk) {
  long long t = (long long)blk * kBlock + threadIdx.x;
  int i, j;
  if (DIR == 0) {
    if (t >= (long long)(d.NX + 2) * d.ny) return;
    i = (int)(t / d.ny) + d.G - 1;
    j = (int)(t % d.ny);
  } else {
    if (t >= (long long)d.nx * (d.NY + 2)) return;
    i = (int)(t / (d.NY + 2));
    j = (int)(t % (d.NY + 2)) + d.G - 1;
  }
  double w[5][4], a[5];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const long long c = DIR == 0 ? (long long)(i - 2 + s) * d.ny + j : (long long)i * d.ny + (j - 2 + s);
    double u[4];
#pragma unroll
    for (int v = 0; v < 4; v++) u[v] = d.U[v * d.nc + c];
    if (DIR == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) w[s][v] = u[v];
      a[s] = d.ax[c];
    } else {
      w[s][0] = u[0];  // YchangetoX, basic_function.cpp:1192-1198
      w[s][1] = u[2];
      w[s][2] = -u[1];
      w[s][3] = u[3];
      a[s] = d.ay[c];
    }
  }
  double lo[8], ro[8];
  weno5_normal(lo, ro, w, a, DIR == 0 ? d.dx : d.dy, DIR == 0 ? d.idx : d.idy, d.recon_var, d.gas);
  double* line = DIR == 0 ? d.line_x : d.line_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const long long fa = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
  const long long fb = DIR == 0 ? (long long)(i + 1) * d.ny + j : (long long)i * (d.ny + 1) + j + 1;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    line[(8 + k) * nf + fa] = lo[k];  // side 1 (right state) of the face at the cell's left
    line[k * nf + fb] = ro[k];        // side 0 (left state) of the face at the cell's right
  }
}
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES))) k_normal(Dev d) {
  normal_body<DIR>(d, xcd_block());
}

// -------------------------------------------------- tangential reconstruction
// weno_5th_ao_with_df_tangential, gausspoint 2 (reconstruction.cpp:2578-2740)
// for one side of one face.  line[5][8]: convar, der1x of the 5 faces.
__device__ __forceinline__ void weno5_tangent(double (*gp)[12], const double (*line)[8], const double* alpha,
                                              double h, double ih, int recon_var, const Gas& g) {
  double df[4];
  df5(df, alpha);
  if (recon_var == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++)
      weno5_2gauss(gp[0][i], gp[0][8 + i], gp[1][i], gp[1][8 + i], df, line[0][i], line[1][i], line[2][i],
                   line[3][i], line[4][i], h, ih);
  } else {
    double base[4];
    convar_to_primvar(base, line[2], g);
    CharBase cb;
    char_base(cb, base, g);
    double c[5][4];
#pragma unroll
    for (int s = 0; s < 5; s++) to_char(c[s], cb, line[s]);
    double var[2][4], d1[2][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      weno5_2gauss(var[0][i], d1[0][i], var[1][i], d1[1][i], df, c[0][i], c[1][i], c[2][i], c[3][i], c[4][i], h,
                   ih);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      from_char(gp[q], cb, var[q], g);
      from_char(gp[q] + 8, cb, d1[q], g);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double t0, t1;
    weno5_2gauss(gp[0][4 + i], t0, gp[1][4 + i], t1, df, line[0][4 + i], line[1][4 + i], line[2][4 + i],
                 line[3][4 + i], line[4][4 + i], h, ih);
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    if (reduce_order(gp[q], g)) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        gp[q][v] = line[2][v];
        gp[q][4 + v] = line[2][4 + v];
        gp[q][8 + v] = 0.0;
      }
    }
  }
}

// Face ranges of the flux (flux_function.cpp:321-342):
//   DIR 0: x faces i in [G, G+NX], j in [G, G+NY)
//   DIR 1: y faces i in [G, G+NX), j in [G, G+NY]
// Tangential stencils (reconstruction.cpp:2553-2576): x face (i,j) uses x
// faces (i, j-2..j+2), alpha_y of cells (i-1, .) for side 0 and (i, .) for
// side 1, h = dy; y face (i,j) uses y faces (i+2..i-2, j) (local tangent is
// -x), alpha_x of cells (., j-1) / (., j), h = dx.
template <int DIR> __global__ void __launch_bounds__(kBlock) k_tangent(Dev d) {
  long long t = gtid();
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  if (t >= 2LL * ni * nj) return;
  const int side = (int)(t / ((long long)ni * nj));
  const long long r = t % ((long long)ni * nj);
  const int i = (int)(r / nj) + d.G, j = (int)(r % nj) + d.G;
  const double* line = DIR == 0 ? d.line_x : d.line_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  double L[5][8], a[5];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    long long f, c;
    if (DIR == 0) {
      f = (long long)i * d.ny + (j - 2 + s);
      c = (long long)(side == 0 ? i - 1 : i) * d.ny + (j - 2 + s);
      a[s] = d.ay[c];
    } else {
      f = (long long)(i + 2 - s) * (d.ny + 1) + j;
      c = (long long)(i + 2 - s) * d.ny + (side == 0 ? j - 1 : j);
      a[s] = d.ax[c];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) L[s][k] = line[(side * 8 + k) * nf + f];
  }
  double gp[2][12];
  weno5_tangent(gp, L, a, DIR == 0 ? d.dy : d.dx, DIR == 0 ? d.idy : d.idx, d.recon_var, d.gas);
  double* gst = DIR == 0 ? d.gst_x : d.gst_y;
  const long long f0 = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int k = 0; k < 12; k++) gst[((side * d.gp + q) * 12 + k) * nf + f0] = gp[q][k];
}

// ---------------------------------------------------- learned WENO5 normal
// cell t of k_normal's range along DIR, and its stencil (x-like frame)
template <int DIR> __device__ __forceinline__ void normal_cell(const Dev& d, long long t, int& i, int& j) {
  if (DIR == 0) {
    i = (int)(t / d.ny) + d.G - 1;
    j = (int)(t % d.ny);
  } else {
    i = (int)(t / (d.NY + 2));
    j = (int)(t % (d.NY + 2)) + d.G - 1;
  }
}
template <int DIR> __device__ __forceinline__ long long normal_count(const Dev& d) {
  return DIR == 0 ? (long long)(d.NX + 2) * d.ny : (long long)d.nx * (d.NY + 2);
}
template <int DIR>
__device__ __forceinline__ void normal_stencil(const Dev& d, int i, int j, double (*w)[4], double* a) {
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const long long c = DIR == 0 ? (long long)(i - 2 + s) * d.ny + j : (long long)i * d.ny + (j - 2 + s);
    double u[4];
#pragma unroll
    for (int v = 0; v < 4; v++) u[v] = d.U[v * d.nc + c];
    if (DIR == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) w[s][v] = u[v];
      a[s] = d.ax[c];
    } else {
      w[s][0] = u[0];  // YchangetoX
      w[s][1] = u[2];
      w[s][2] = -u[1];
      w[s][3] = u[3];
      a[s] = d.ay[c];
    }
  }
}

// classifier training samples: features and label of every stride-th cell
template <int DIR>
__global__ void __launch_bounds__(kBlock) k_label(Dev d, int stride, double tol, float* X, float* Y) {
  const long long t = gtid();
  const long long ns = (normal_count<DIR>(d) + stride - 1) / stride;
  if (t >= ns) return;
  int i, j;
  normal_cell<DIR>(d, t * stride, i, j);
  double w[5][4], a[5];
  normal_stencil<DIR>(d, i, j, w, a);
  double q[2][5];
  stencil_rho_p(q, w, d.gas);
  float x[16];
  stencil_features(x, q);
#pragma unroll
  for (int k = 0; k < 16; k++) X[t * 16 + k] = x[k];
  Y[t] = normal_label(w, a, DIR == 0 ? d.dx : d.dy, DIR == 0 ? d.idx : d.idy, d.recon_var, d.gas, tol);
}

// counts of linear / all reconstructions of a wave, added by its first
// active lane to one of kStatSlots slot pairs (spread by block so that the
// adds do not queue on one address); the host sums the slots
constexpr int kStatSlots = 1024;
// lstat word that accumulates the x (+0) / y (+1) flux faces the lists held
constexpr int kListedWord = 4 * kStatSlots;
__device__ __forceinline__ void count_linear(unsigned long long* stat, bool lin) {
  const unsigned long long act = __ballot(1), bl = __ballot(lin);
  if ((threadIdx.x & 63) == __ffsll((long long)act) - 1) {
    unsigned long long* slot = stat + 2 * ((blockIdx.x * 4 + (threadIdx.x >> 6)) % kStatSlots);
    atomicAdd(slot, (unsigned long long)__popcll(bl));
    atomicAdd(slot + 1, (unsigned long long)__popcll(act));
  }
}

// appends `value` of every lane with `listed` to list (one atomic per wave,
// ranks by lane order); every active lane of the wave must call it
__device__ __forceinline__ void list_append(unsigned* count, unsigned* list, bool listed, unsigned value) {
  const unsigned long long bl = __ballot(listed);
  if (!bl) return;
  const int lead = __ffsll((long long)bl) - 1;
  unsigned base = 0;
  if ((int)(threadIdx.x & 63) == lead) base = atomicAdd(count, (unsigned)__popcll(bl));
  base = (unsigned)__shfl((int)base, lead);
  const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(bl >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bl, 0u));
  if (listed) list[base + rank] = value;
}
// the counter of shard s of the dir list whose shard counts start at cnt
__device__ __forceinline__ unsigned* shard_count(unsigned* cnt, int dir, int s) {
  return cnt + (dir * kListShards + s) * kCountStride;
}
// the shard tile blk of a sweep of ntiles tiles appends to: contiguous runs
// of tiles share a shard (at most ceil(ntiles / kListShards) tiles each),
// so that a shard's entries stay close in the grid and, with xcd_block's
// mapping, each counter takes the appends of one XCD's workgroups
__device__ __forceinline__ int list_shard(unsigned blk, unsigned ntiles) {
  return (int)((unsigned long long)blk * kListShards / ntiles);
}
// the prefix record of the dir list (per_tile entries per tile, the
// shards' tiles one after the other) into pre (Dev::lpre layout): the
// shards' first tiles, the tile count, the shards' entry counts; the shard
// counters are cleared for the list's next sweep (the list's launch reads
// the record).  One lane; returns the tile count.
__device__ __forceinline__ unsigned list_prefix(unsigned* cnt, int dir, int per_tile, unsigned* pre) {
  unsigned c[kListShards];
#pragma unroll
  for (int k = 0; k < kListShards; k++) c[k] = cnt[(dir * kListShards + k) * kCountStride];
  unsigned acc = 0;
#pragma unroll
  for (int k = 0; k < kListShards; k++) {
    pre[k] = acc;
    pre[kListShards + 1 + k] = c[k];
    cnt[(dir * kListShards + k) * kCountStride] = 0u;
    acc += (c[k] + per_tile - 1) / per_tile;
  }
  pre[kListShards] = acc;
  return acc;
}
// tile b of a list: its shard s, the tile tb within the shard and the
// shard's count n; false past the last tile.  pre: the list's prefix
// record.  Uniform over a workgroup.
__device__ __forceinline__ bool list_tile(const unsigned* pre, unsigned b, int& s, unsigned& tb, unsigned& n) {
  if (b >= pre[kListShards]) return false;
  // the last shard starting at or before b (an empty shard starts where
  // the next one does)
  s = 0;
#pragma unroll
  for (int k = 1; k < kListShards; k++)
    if (pre[k] <= b) s = k;
  tb = b - pre[s];
  n = pre[kListShards + 1 + s];
  return true;
}
// entries of a list (pre: its prefix record)
__device__ __forceinline__ unsigned list_total(const unsigned* pre) {
  unsigned n = 0;
  for (int k = 0; k < kListShards; k++) n += pre[kListShards + 1 + k];
  return n;
}
// forms the tile prefixes of nkinds consecutive lists from `kind` (x and y
// of one kind: nkinds 2), one lane each (single-problem path)
__global__ void __launch_bounds__(64) k_list_prefix(Dev d, int kind, int nkinds, int per_tile) {
  if ((int)threadIdx.x < nkinds) {
    const int k = kind + (int)threadIdx.x;
    list_prefix(k < KL_FLUX_X ? d.ncount : d.tcount, k & 1, per_tile, d.lpre + k * kListRecord);
  }
}

// the two face records a cell's normal reconstruction writes (side 1 of
// the face at its left, side 0 of the face at its right)
template <int DIR>
__device__ __forceinline__ void store_normal_faces(const Dev& d, int i, int j, const double* lo, const double* ro) {
  double* line = DIR == 0 ? d.line_x : d.line_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const long long fa = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
  const long long fb = DIR == 0 ? (long long)(i + 1) * d.ny + j : (long long)i * (d.ny + 1) + j + 1;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    line[(8 + k) * nf + fa] = lo[k];
    line[k * nf + fb] = ro[k];
  }
}

// the linear-weight normal reconstruction of one cell (the fifth-order
// candidate on the conservative values, the order-reduction test kept):
// lo the state at its left face (side 1 of that face), ro at its right
// face (side 0); w[5][4] its stencil in the x-like frame
__device__ __forceinline__ void linear_normal(double* lo, double* ro, const double (*w)[4], double h, double ih,
                                              const Gas& g) {
#pragma unroll
  for (int v = 0; v < 4; v++) {
    weno5_linear_face(0, lo[v], lo[4 + v], w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], h, ih);
    weno5_linear_face(1, ro[v], ro[4 + v], w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], h, ih);
  }
  const bool rr = reduce_order(ro, g);
  const bool rl = reduce_order(lo, g);
  if (rl || rr) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
      ro[m] = w[2][m];
      lo[m] = w[2][m];
      ro[4 + m] = 0.0;
      lo[4 + m] = 0.0;
    }
  }
}

// k_normal with the classifier: a stencil the network calls smooth (logit
// below the threshold) with a zero DF sensor takes the linear weights, i.e.
// the fifth-order candidate, formed directly on the conservative variables
// (the characteristic projection commutes with a fixed linear combination);
// the order-reduction test is kept.  Every other stencil is listed for
// k_normal_list (WENO5-AO-DF), so this kernel carries no WENO registers.
template <int DIR>
__device__ __forceinline__ void normal_learned_body(const Dev& d, unsigned blk, unsigned ntiles) {
  __shared__ __align__(16) short xs[kBlock * 16];
  __shared__ float ls[kBlock];
  __shared__ MlpLds s_mlp;
  // the network's operands load while the stencils do
  mlp_stage(s_mlp, d.mlp, d.mlpw);
  const long long t = (long long)blk * kBlock + threadIdx.x;
  const bool valid = t < normal_count<DIR>(d);
  int i, j;
  normal_cell<DIR>(d, valid ? t : 0, i, j);
  double w[5][4], a[5];
  float x[16];
  if (valid) {
    normal_stencil<DIR>(d, i, j, w, a);
    double q[2][5];
    stencil_rho_p(q, w, d.gas);
    stencil_features(x, q);
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = 0.0f;
  }
  const int wv = threadIdx.x >> 6;
  __syncthreads();
  const float z = mlp_wave_logit(x, MlpShared{&s_mlp}, xs + wv * 64 * 16, ls + wv * 64);
  if (!valid) return;
  double amax = 0.0;
#pragma unroll
  for (int s = 0; s < 5; s++) amax = fmax(amax, a[s]);
  const bool lin = z < d.mlp_thr && amax == 0.0;
  count_linear(d.lstat, lin);
  d.lin[DIR][(long long)i * d.ny + j] = lin ? 1 : 0;
  // the other stencils run WENO5-AO-DF in k_normal_list, in waves of
  // troubled cells only
  const int sh = list_shard(blk, ntiles);
  list_append(shard_count(d.ncount, DIR, sh), d.nlist[DIR] + (size_t)sh * d.ncap[DIR], !lin, (unsigned)t);
  // linear cells' line values: formed by the smooth-face flux launch when
  // it is set to (fused_lines)
  if (!lin || (d.fused_lines >> DIR & 1)) return;
  double lo[8], ro[8];
  linear_normal(lo, ro, w, DIR == 0 ? d.dx : d.dy, DIR == 0 ? d.idx : d.idy, d.gas);
  store_normal_faces<DIR>(d, i, j, lo, ro);
}
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES)))
k_normal_learned(Dev d) {
  normal_learned_body<DIR>(d, xcd_block(), gridDim.x);
}
// Both classifier sweeps in one launch, for a learned path whose smooth-face
// flux launch forms every linear line value (so the sweeps store none): a
// workgroup takes kClsRows x kClsCols cells of the whole frame (ghost cells
// included), forms rho and p of each cell its stencils reach once, into LDS
// (in FP32, as the features take them; p once in each sweep's x-like frame,
// whose operation order differs), and each thread runs the network on its
// cell's x stencil and on its y stencil where the cell is in that sweep's
// range (normal_cell), setting the flag and listing the cell as
// normal_learned_body does.  A stencil cell outside the frame only feeds a
// cell outside both ranges.
__device__ __forceinline__ void classify_body(const Dev& d, unsigned blk, unsigned ntiles) {
  constexpr int SR = kClsRows + 4, SC = kClsCols + 4, NS = SR * SC, NU = (NS + kBlock - 1) / kBlock;
  static_assert(kClsRows * kClsCols == kBlock, "a thread per cell");
  __shared__ __align__(16) short xs[kBlock * 16];
  __shared__ float ls[kBlock];
  __shared__ MlpLds s_mlp;
  __shared__ float srho[NS], spx[NS], spy[NS];
  __shared__ unsigned char sz[NS];  // bit 0: alpha_x not above zero, bit 1: alpha_y
  mlp_stage(s_mlp, d.mlp, d.mlpw);
  const long long tcols = cls_tile_cols(d);
  const int i0 = (int)(blk / tcols) * kClsRows, j0 = (int)(blk % tcols) * kClsCols;
  {
    // staged cell e: (i0 - 2 + e / SC, j0 - 2 + e % SC); every load is
    // issued before the first conversion
    double u[NU][4], ax[NU], ay[NU];
#pragma unroll
    for (int r = 0; r < NU; r++) {
      const int e = threadIdx.x + r * kBlock, gi = i0 - 2 + e / SC, gj = j0 - 2 + e % SC;
      if (e < NS && gi >= 0 && gi < d.nx && gj >= 0 && gj < d.ny) {
        const long long c = (long long)gi * d.ny + gj;
#pragma unroll
        for (int v = 0; v < 4; v++) u[r][v] = d.U[v * d.nc + c];
        ax[r] = d.ax[c];
        ay[r] = d.ay[c];
      } else {
        u[r][0] = 1.0;
        u[r][1] = u[r][2] = 0.0;
        u[r][3] = 1.0;
        ax[r] = ay[r] = 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < NU; r++) {
      const int e = threadIdx.x + r * kBlock;
      if (e < NS) {
        double px[4], py[4];
        convar_to_primvar(px, u[r], d.gas);
        const double wy[4] = {u[r][0], u[r][2], -u[r][1], u[r][3]};  // YchangetoX
        convar_to_primvar(py, wy, d.gas);
        srho[e] = (float)u[r][0];
        spx[e] = (float)px[3];
        spy[e] = (float)py[3];
        sz[e] = (unsigned char)((ax[r] > 0.0 ? 0 : 1) | (ay[r] > 0.0 ? 0 : 2));
      }
    }
  }
  __syncthreads();
  const int ti = threadIdx.x / kClsCols, tj = threadIdx.x % kClsCols, i = i0 + ti, j = j0 + tj;
  const bool inside = i < d.nx && j < d.ny;
  const bool vx = inside && i >= d.G - 1 && i <= d.NX + d.G, vy = inside && j >= d.G - 1 && j <= d.NY + d.G;
  const int wv = threadIdx.x >> 6, sh = list_shard(blk, ntiles);
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
    const bool valid = dir == 0 ? vx : vy;
    float x[16];
    unsigned zero = 1u << dir;
    if (valid) {
      float q[2][5];
#pragma unroll
      for (int s = 0; s < 5; s++) {
        const int e = dir == 0 ? (ti + s) * SC + tj + 2 : (ti + 2) * SC + tj + s;
        q[0][s] = srho[e];
        q[1][s] = dir == 0 ? spx[e] : spy[e];
        zero &= sz[e];
      }
      stencil_features_f(x, q);
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) x[k] = 0.0f;
    }
    const float z = mlp_wave_logit(x, MlpShared{&s_mlp}, xs + wv * 64 * 16, ls + wv * 64);
    if (valid) {
      const bool lin = z < d.mlp_thr && zero != 0;
      count_linear(d.lstat, lin);
      d.lin[dir][(long long)i * d.ny + j] = lin ? 1 : 0;
      const long long t = dir == 0 ? (long long)(i - d.G + 1) * d.ny + j : (long long)i * (d.NY + 2) + j - d.G + 1;
      list_append(shard_count(d.ncount, dir, sh), d.nlist[dir] + (size_t)sh * d.ncap[dir], !lin, (unsigned)t);
    }
  }
}
__global__ void __launch_bounds__(kBlock) k_classify(Dev d) { classify_body(d, xcd_block(), gridDim.x); }
// WENO5-AO-DF over the cells k_normal_learned listed, one thread each, a
// list tile per workgroup; launched with the normal sweep's block count +
// kListShards (a bound on the list's tiles, whose number is known only on
// the device: k_list_prefix forms it; blocks past it return at once)
template <int DIR> __device__ __forceinline__ void normal_list_body(const Dev& d, unsigned blk) {
  int s;
  unsigned tb, n;
  if (!list_tile(d.lpre + (DIR == 0 ? KL_NORMAL_X : KL_NORMAL_Y) * kListRecord, blk, s, tb, n)) return;
  const long long k = (long long)tb * kBlock + threadIdx.x;
  if (k >= (long long)n) return;
  int i, j;
  normal_cell<DIR>(d, (long long)d.nlist[DIR][(size_t)s * d.ncap[DIR] + k], i, j);
  double w[5][4], a[5];
  normal_stencil<DIR>(d, i, j, w, a);
  double lo[8], ro[8];
  weno5_normal(lo, ro, w, a, DIR == 0 ? d.dx : d.dy, DIR == 0 ? d.idx : d.idy, d.recon_var, d.gas);
  store_normal_faces<DIR>(d, i, j, lo, ro);
}
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES)))
k_normal_list(Dev d) {
  normal_list_body<DIR>(d, blockIdx.x);
}
// both directions' listed cells in one launch: workgroups [0, nbx) take the
// x list's tiles, the rest the y list's (one launch tail instead of two)
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES)))
k_normal_list_xy(Dev d, unsigned nbx) {
  if (blockIdx.x < nbx) normal_list_body<0>(d, blockIdx.x);
  else normal_list_body<1>(d, blockIdx.x - nbx);
}

// logits of n feature rows (parity of the MFMA inference)
__global__ void __launch_bounds__(kBlock) k_mlp_infer(const float* P, const unsigned short* W, long long n, const float* X,
                                                   float* out) {
  __shared__ __align__(16) short xs[kBlock * 16];
  __shared__ float ls[kBlock];
  const long long t = gtid();
  const bool valid = t < n;
  float x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = valid ? X[t * 16 + k] : 0.0f;
  const int wv = threadIdx.x >> 6;
  const float z = mlp_wave_logit(x, P, W, xs + wv * 64 * 16, ls + wv * 64);
  if (valid) out[t] = z;
}

// ------------------------------------------------- WENO7 / WENO9 (normal)
// WENO7_AO_with_df / WENO9_AO_with_df (reconstruction.cpp:1894-2028,
// 2063-2211) for one cell along DIR; w[0..2r] in the x-like frame
template <int ORDER>
__device__ __forceinline__ void weno79_normal(double* lo, double* ro, const double (*w)[4], const double* alpha,
                                              double h, double ih, int recon_var, const Gas& g) {
  constexpr int r = (ORDER - 1) / 2, NS = ORDER == 7 ? 5 : 6;
  double df[NS];
  df79<ORDER>(df, alpha + r);
  double pn1[4], p0[4], pp1[4];
  convar_to_primvar(pn1, w[r - 1], g);
  convar_to_primvar(p0, w[r], g);
  convar_to_primvar(pp1, w[r + 1], g);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double base[4];
#pragma unroll
    for (int i = 0; i < 4; i++) base[i] = side == 0 ? 0.5 * (pn1[i] + p0[i]) : 0.5 * (pp1[i] + p0[i]);
    double c[2 * r + 1][4];
    CharBase cb;
    if (recon_var == 0) {
#pragma unroll
      for (int s = 0; s < 2 * r + 1; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) c[s][i] = w[s][i];
    } else {
      char_base(cb, base, g);
#pragma unroll
      for (int s = 0; s < 2 * r + 1; s++) to_char(c[s], cb, w[s]);
    }
    double var[4], d1[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      double line[2 * r + 1];
#pragma unroll
      for (int s = 0; s < 2 * r + 1; s++) line[s] = c[s][i];
      if (side == 0) weno79_face<ORDER, 0>(var[i], d1[i], line + r, df, h, ih);
      else weno79_face<ORDER, 1>(var[i], d1[i], line + r, df, h, ih);
    }
    double* out = side == 0 ? lo : ro;
    if (recon_var == 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        out[i] = var[i];
        out[4 + i] = d1[i];
      }
    } else {
      from_char(out, cb, var, g);
      from_char(out + 4, cb, d1, g);
    }
  }
  const bool rr = reduce_order(ro, g);
  const bool rl = reduce_order(lo, g);
  if (rl || rr) {
#pragma unroll
    for (int m = 0; m < 4; m++) {
      ro[m] = w[r][m];
      lo[m] = w[r][m];
      ro[4 + m] = 0.0;
      lo[4 + m] = 0.0;
    }
  }
}

// the cells and faces of k_normal, with stencils of radius r
template <int DIR, int ORDER> __global__ void __launch_bounds__(kBlock) k_normal79(Dev d) {
  constexpr int r = (ORDER - 1) / 2;
  long long t = gtid();
  int i, j;
  if (DIR == 0) {
    if (t >= (long long)(d.NX + 2) * d.ny) return;
    i = (int)(t / d.ny) + d.G - 1;
    j = (int)(t % d.ny);
  } else {
    if (t >= (long long)d.nx * (d.NY + 2)) return;
    i = (int)(t / (d.NY + 2));
    j = (int)(t % (d.NY + 2)) + d.G - 1;
  }
  double w[2 * r + 1][4], a[2 * r + 1];
#pragma unroll
  for (int s = 0; s < 2 * r + 1; s++) {
    const long long c = DIR == 0 ? (long long)(i - r + s) * d.ny + j : (long long)i * d.ny + (j - r + s);
    double u[4];
#pragma unroll
    for (int v = 0; v < 4; v++) u[v] = d.U[v * d.nc + c];
    if (DIR == 0) {
#pragma unroll
      for (int v = 0; v < 4; v++) w[s][v] = u[v];
      a[s] = d.ax[c];
    } else {
      w[s][0] = u[0];  // YchangetoX
      w[s][1] = u[2];
      w[s][2] = -u[1];
      w[s][3] = u[3];
      a[s] = d.ay[c];
    }
  }
  double lo[8], ro[8];
  weno79_normal<ORDER>(lo, ro, w, a, DIR == 0 ? d.dx : d.dy, DIR == 0 ? d.idx : d.idy, d.recon_var, d.gas);
  double* line = DIR == 0 ? d.line_x : d.line_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const long long fa = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
  const long long fb = DIR == 0 ? (long long)(i + 1) * d.ny + j : (long long)i * (d.ny + 1) + j + 1;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    line[(8 + k) * nf + fa] = lo[k];
    line[k * nf + fb] = ro[k];
  }
}

// --------------------------------------------- WENO7 / WENO9 (tangential)
// weno_7th/9th_ao_with_df_tangential (reconstruction.cpp:2867-3018,
// 3209-3360) for one side of one face: NG = 3 / 4 Gauss points.  The
// convar components go first (values and der1y, then the order-reduction
// test, which reads only them), der1x after, so the line's der1x values are
// loaded only when needed.  WENO9, characteristic variables, right side:
// the outermost two points keep the left side's projection (`quirk`: left
// side convar of line points 0, r, 2r), as the reference leaves them
// (:3306-3312).
// the tangential reconstruction of one side of x / y face (i, j): its NG
// Gauss-point states (convar, der1x, der1y) go to put(q, component, value)
template <int DIR, int ORDER, class Put>
__device__ __forceinline__ void tangent79_side(const Dev& d, int side, int i, int j, const Put& put) {
  constexpr int r = (ORDER - 1) / 2, NS = ORDER == 7 ? 5 : 6, NG = ORDER == 7 ? 3 : 4;
  const double* line = DIR == 0 ? d.line_x : d.line_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  long long fs[2 * r + 1];
  double a[2 * r + 1];
#pragma unroll
  for (int s = 0; s < 2 * r + 1; s++) {
    long long c;
    if (DIR == 0) {
      fs[s] = (long long)i * d.ny + (j - r + s);
      c = (long long)(side == 0 ? i - 1 : i) * d.ny + (j - r + s);
      a[s] = d.ay[c];
    } else {
      fs[s] = (long long)(i + r - s) * (d.ny + 1) + j;
      c = (long long)(i + r - s) * d.ny + (side == 0 ? j - 1 : j);
      a[s] = d.ax[c];
    }
  }
  double df[NS];
  df79<ORDER>(df, a + r);
  const double h = DIR == 0 ? d.dy : d.dx, ih = DIR == 0 ? d.idy : d.idx;
  double L[2 * r + 1][4];
#pragma unroll
  for (int s = 0; s < 2 * r + 1; s++)
#pragma unroll
    for (int k = 0; k < 4; k++) L[s][k] = line[(side * 8 + k) * nf + fs[s]];
  double gv[NG][4], gy[NG][4];
  if (d.recon_var == 0) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double w[2 * r + 1], v[NG], vx[NG];
#pragma unroll
      for (int s = 0; s < 2 * r + 1; s++) w[s] = L[s][k];
      weno79_gauss<ORDER>(v, vx, w + r, df, h, ih);
#pragma unroll
      for (int q = 0; q < NG; q++) {
        gv[q][k] = v[q];
        gy[q][k] = vx[q];
      }
    }
  } else {
    double base[4];
    convar_to_primvar(base, L[r], d.gas);
    CharBase cb;
    char_base(cb, base, d.gas);
    double c[2 * r + 1][4];
#pragma unroll
    for (int s = 0; s < 2 * r + 1; s++) to_char(c[s], cb, L[s]);
    if (ORDER == 9 && side == 1) {
      double q0[4], qr[4], q2r[4], bl[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        q0[k] = line[k * nf + fs[0]];
        qr[k] = line[k * nf + fs[r]];
        q2r[k] = line[k * nf + fs[2 * r]];
      }
      convar_to_primvar(bl, qr, d.gas);
      CharBase cl;
      char_base(cl, bl, d.gas);
      to_char(c[0], cl, q0);
      to_char(c[2 * r], cl, q2r);
    }
    double var[NG][4], d1[NG][4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double w[2 * r + 1], v[NG], vx[NG];
#pragma unroll
      for (int s = 0; s < 2 * r + 1; s++) w[s] = c[s][k];
      weno79_gauss<ORDER>(v, vx, w + r, df, h, ih);
#pragma unroll
      for (int q = 0; q < NG; q++) {
        var[q][k] = v[q];
        d1[q][k] = vx[q];
      }
    }
#pragma unroll
    for (int q = 0; q < NG; q++) {
      from_char(gv[q], cb, var[q], d.gas);
      from_char(gy[q], cb, d1[q], d.gas);
    }
  }
  bool red[NG];
#pragma unroll
  for (int q = 0; q < NG; q++) red[q] = reduce_order(gv[q], d.gas);
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      put(q, k, red[q] ? L[r][k] : gv[q][k]);
      put(q, 8 + k, red[q] ? 0.0 : gy[q][k]);
    }
  // der1x along the face
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double w[2 * r + 1], v[NG], vx[NG];
#pragma unroll
    for (int s = 0; s < 2 * r + 1; s++) w[s] = line[(side * 8 + 4 + k) * nf + fs[s]];
    weno79_gauss<ORDER>(v, vx, w + r, df, h, ih);
#pragma unroll
    for (int q = 0; q < NG; q++) put(q, 4 + k, red[q] ? w[r] : v[q]);
  }
}

template <int DIR, int ORDER> __global__ void __launch_bounds__(kBlock) k_tangent79(Dev d) {
  long long t = gtid();
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  if (t >= 2LL * ni * nj) return;
  const int side = (int)(t / ((long long)ni * nj));
  const long long rr = t % ((long long)ni * nj);
  const int i = (int)(rr / nj) + d.G, j = (int)(rr % nj) + d.G;
  double* gst = DIR == 0 ? d.gst_x : d.gst_y;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const long long f0 = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
  tangent79_side<DIR, ORDER>(d, side, i, j, [&](int q, int k, double v) {
    gst[((side * d.gp + q) * 12 + k) * nf + f0] = v;
  });
}

// -------------------------------------------------------------------- flux
// Centre state + GKS2D per (face, Gauss point); the S2O4 combination of the
// Gauss point's flux for this stage (time_advance.cpp:380-426, weighted by the
// Gauss weight) rotated to the global frame (Local_to_Global,
// basic_function.cpp:1226-1233), so the cell update reads one 4-vector per
// (face, Gauss point); and Calculate_alpha_k of the point
// (reconstruction.cpp:1466-1495; y faces back to the global frame with
// XchangetoY first).
//
// S2O4_2D (time_advance.cpp:221-230): stage 0 uses 0.5 f0 + 1/8 derf0;
// stage 1 uses (f0 + 1/6 derf0) + (0 f1 + 1/3 derf1).  Stage 0 stores the
// first bracket (times the weight) for stage 1; the reference's running sum
// starts from 0.0, which adds nothing.
// the flux of one (face, Gauss point) from its two Gauss-point states, and
// everything stored per point: the stage-combined rotated flux, the stage-1
// share, the raw (f, derf) when kept, and Calculate_alpha_k
template <int DIR, class V>
__device__ __forceinline__ void flux_point_tail(const Dev& d, int q, long long f, long long nf, const V& L,
                                                const V& R, int stage, double dt, bool smooth = false) {
  // alpha_k first: only the four conservative values of each side stay
  // live into it, and none of them into the flux below
  {
    double pl[4], pr[4], cl[4], cr[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      cl[k] = L[k];
      cr[k] = R[k];
    }
    convar_to_primvar(pl, cl, d.gas);
    convar_to_primvar(pr, cr, d.gas);
    double a;
    if (DIR == 0) {
      a = alpha_k(pl, pr, d.gas);
    } else {
      const double tl[4] = {pl[0], -pl[2], pl[1], pl[3]};
      const double tr[4] = {pr[0], -pr[2], pr[1], pr[3]};
      a = alpha_k(tl, tr, d.gas);
    }
    (DIR == 0 ? d.akx : d.aky)[q * nf + f] = a;
  }
  double fo[4], dfo[4];
  if (smooth) gks_point_smooth(fo, dfo, L, R, dt, d.gas);
  else gks_point(fo, dfo, nullptr, L, R, dt, d.gas);
  if (d.keep_face_fluxes) {
    double* fl = DIR == 0 ? (stage == 0 ? d.flx[0] : d.flx[1]) : (stage == 0 ? d.fly[0] : d.fly[1]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      fl[(q * 8 + k) * nf + f] = fo[k];
      fl[(q * 8 + 4 + k) * nf + f] = dfo[k];
    }
  }
  // (selects, not d.gw[q]: a runtime index into the descriptor would force
  // a batched kernel's copy of it into scratch)
  const double w = q == 0 ? d.gw[0] : q == 1 ? d.gw[1] : q == 2 ? d.gw[2] : d.gw[3];
  double x[4];
  double* pc = DIR == 0 ? d.pcx : d.pcy;
  if (stage == 0) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      x[v] = w * (0.5 * fo[v] + 1.0 / 8.0 * dfo[v]);
      pc[(q * 4 + v) * nf + f] = w * (1.0 * fo[v] + 1.0 / 6.0 * dfo[v]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) x[v] = pc[(q * 4 + v) * nf + f] + w * (0.0 * fo[v] + 1.0 / 3.0 * dfo[v]);
  }
  double* xc = DIR == 0 ? d.xcx : d.xcy;
  // normal (1,0): identity; normal (0,1): (x1, x2) -> (-x2, x1)
  xc[(q * 4 + 0) * nf + f] = x[0];
  xc[(q * 4 + 1) * nf + f] = DIR == 0 ? x[1] : -x[2];
  xc[(q * 4 + 2) * nf + f] = DIR == 0 ? x[2] : x[1];
  xc[(q * 4 + 3) * nf + f] = x[3];
}

// flux_point_tail for the fused WENO5 launches, whose two Gauss points of a
// face are the lanes 2m (q = 0) and 2m + 1 (q = 1) of one wave.  The pair
// forms the face's DF value 0.5 (alpha_0 + alpha_1), thresholded as
// Calculate_alpha (reconstruction.cpp:1466-1521) does, so that k_alpha reads
// one value per face (the same sum in the same order: bit for bit the
// per-point form).  In the fast build the pair also sums the stage-combined
// fluxes of its two points (each lane stores two of the four rotated
// components) and stage 0's stage-1 share, halving what k_update reads;
// the IEEE build keeps one record per point and Update_with_gauss's
// summation order.  Both lanes of a pair must be active together.
template <int DIR, class V>
__device__ __forceinline__ void flux_pair_tail(const Dev& d, int q, long long f, long long nf, const V& L,
                                               const V& R, int stage, double dt, bool smooth) {
  {
    double pl[4], pr[4], cl[4], cr[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      cl[k] = L[k];
      cr[k] = R[k];
    }
    convar_to_primvar(pl, cl, d.gas);
    convar_to_primvar(pr, cr, d.gas);
    double a;
    if (DIR == 0) {
      a = alpha_k(pl, pr, d.gas);
    } else {
      const double tl[4] = {pl[0], -pl[2], pl[1], pl[3]};
      const double tr[4] = {pr[0], -pr[2], pr[1], pr[3]};
      a = alpha_k(tl, tr, d.gas);
    }
    const double ap = __shfl_xor(a, 1);
    if (q == 0) {
      // (not fused with the last product of alpha_k: see below)
#pragma clang fp contract(off)
      const double s = 0.5 * (a + ap);
      (DIR == 0 ? d.akx : d.aky)[f] = s < 2.0 ? 0.0 : s;
    }
  }
  double fo[4], dfo[4];
  if (smooth) gks_point_smooth(fo, dfo, L, R, dt, d.gas);
  else gks_point(fo, dfo, nullptr, L, R, dt, d.gas);
  if (d.keep_face_fluxes) {
    double* fl = DIR == 0 ? (stage == 0 ? d.flx[0] : d.flx[1]) : (stage == 0 ? d.fly[0] : d.fly[1]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      fl[(q * 8 + k) * nf + f] = fo[k];
      fl[(q * 8 + 4 + k) * nf + f] = dfo[k];
    }
  }
  const double w = q == 0 ? d.gw[0] : d.gw[1];
  double* pc = DIR == 0 ? d.pcx : d.pcy;
  double* xc = DIR == 0 ? d.xcx : d.xcy;
#if GKS_FAST
  // rotated (normal (0,1): (x1, x2) -> (-x2, x1)); lane q stores the
  // rotated components 2q, 2q + 1 of the pair's sums.  No contraction here:
  // whether a product and the pair's sum would fuse depends on how the
  // kernel around it compiled, and the batched and single launches must
  // round alike.
  {
#pragma clang fp contract(off)
  auto rot = [](const double* x, double* y) {
    y[0] = x[0];
    y[1] = DIR == 0 ? x[1] : -x[2];
    y[2] = DIR == 0 ? x[2] : x[1];
    y[3] = x[3];
  };
  auto pair_sum = [&](const double* y, double* out) {
#pragma unroll
    for (int m = 0; m < 2; m++) {
      const double mine = q == 0 ? y[m] : y[2 + m], theirs = q == 0 ? y[2 + m] : y[m];
      out[m] = mine + __shfl_xor(theirs, 1);
    }
  };
  double xl[4], xr[4], xs[2];
  if (stage == 0) {
    double pl[4], pr[4], ps[2];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      xl[v] = w * (0.5 * fo[v] + 1.0 / 8.0 * dfo[v]);
      pl[v] = w * (1.0 * fo[v] + 1.0 / 6.0 * dfo[v]);
    }
    rot(pl, pr);
    pair_sum(pr, ps);
#pragma unroll
    for (int m = 0; m < 2; m++) pc[(2 * q + m) * nf + f] = ps[m];
    rot(xl, xr);
    pair_sum(xr, xs);
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) xl[v] = w * (0.0 * fo[v] + 1.0 / 3.0 * dfo[v]);
    rot(xl, xr);
    pair_sum(xr, xs);
#pragma unroll
    for (int m = 0; m < 2; m++) xs[m] += pc[(2 * q + m) * nf + f];
  }
#pragma unroll
  for (int m = 0; m < 2; m++) xc[(2 * q + m) * nf + f] = xs[m];
  }
#else
  double x[4];
  if (stage == 0) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      x[v] = w * (0.5 * fo[v] + 1.0 / 8.0 * dfo[v]);
      pc[(q * 4 + v) * nf + f] = w * (1.0 * fo[v] + 1.0 / 6.0 * dfo[v]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) x[v] = pc[(q * 4 + v) * nf + f] + w * (0.0 * fo[v] + 1.0 / 3.0 * dfo[v]);
  }
  xc[(q * 4 + 0) * nf + f] = x[0];
  xc[(q * 4 + 1) * nf + f] = DIR == 0 ? x[1] : -x[2];
  xc[(q * 4 + 2) * nf + f] = DIR == 0 ? x[2] : x[1];
  xc[(q * 4 + 3) * nf + f] = x[3];
#endif
}

// waves per SIMD asked of the flux-point kernels (k_flux, k_tanflux)
#ifndef GKS_FLUX_WAVES
#define GKS_FLUX_WAVES 2
#endif
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_FLUX_WAVES)))
k_flux(Dev d, int stage, double dt) {
  long long t = gtid();
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  if (t >= (long long)d.gp * ni * nj) return;
  const int q = (int)(t / ((long long)ni * nj));
  const long long r = t % ((long long)ni * nj);
  const int i = (int)(r / nj) + d.G, j = (int)(r % nj) + d.G;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const long long f = DIR == 0 ? (long long)i * d.ny + j : (long long)i * (d.ny + 1) + j;
  const double* gst = DIR == 0 ? d.gst_x : d.gst_y;
  double L[12], R[12];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    L[k] = gst[((0 * d.gp + q) * 12 + k) * nf + f];
    R[k] = gst[((1 * d.gp + q) * 12 + k) * nf + f];
  }
  flux_point_tail<DIR>(d, q, f, nf, L, R, stage, dt);
}

// ------------------------------------------- tangential + flux, one launch
// WENO5 (2 Gauss points): a workgroup takes kTanFlux consecutive faces.
// Phase 1, thread (side, face): the tangential reconstruction of k_tangent,
// its Gauss-point states kept in LDS; phase 2, thread (Gauss point, face):
// the flux of k_flux from those states.  The Gauss-point states never go to
// HBM (48 doubles per face each way).
//
// A workgroup of 128 lanes takes 64 faces and asks for two waves per SIMD
// (at most 256 VGPRs): the fast build fits in about 250 VGPRs without
// scratch, about 0.88-0.95 ms per launch at 2048^2 (one wave per SIMD took
// 1.54 ms); the IEEE build spills a few hundred bytes per lane.
constexpr int kTanFlux = 64;
// LISTED (learned path only): the tile's faces are list[blk * TF ...], nlist
// entries in all, instead of the consecutive faces blk * TF ...; they are
// the faces k_tanflux_cheap left (not both sides linear), so the smooth-flow
// flux never applies and their linear-weight counts are already taken
template <int DIR, bool LEARNED, int TF, bool LISTED = false>
__device__ __forceinline__ void tanflux_body(const Dev& d, unsigned blk, int stage, double dt,
                                             const unsigned* list = nullptr, long long nlist = 0) {
  __shared__ double sg[2][12][TF][2];  // [side][component][face][gauss point]
  __shared__ unsigned char slin[2][LEARNED ? TF : 1];  // side took the linear weights
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  const long long nfaces = (long long)ni * nj;
  const int tl = threadIdx.x % TF, hi = threadIdx.x / TF;
  const long long slot = (long long)blk * TF + tl;
  const bool valid = LISTED ? slot < nlist : slot < nfaces;
  const long long r = LISTED ? (valid ? (long long)list[slot] : 0) : slot;
  const int i = valid ? (int)(r / nj) + d.G : d.G, j = valid ? (int)(r % nj) + d.G : d.G;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const int side = hi;
  double L[5][8], a[5];
  if (valid) {
    const double* line = DIR == 0 ? d.line_x : d.line_y;
#pragma unroll
    for (int s = 0; s < 5; s++) {
      long long f, c;
      if (DIR == 0) {
        f = (long long)i * d.ny + (j - 2 + s);
        c = (long long)(side == 0 ? i - 1 : i) * d.ny + (j - 2 + s);
        a[s] = d.ay[c];
      } else {
        f = (long long)(i + 2 - s) * (d.ny + 1) + j;
        c = (long long)(i + 2 - s) * d.ny + (side == 0 ? j - 1 : j);
        a[s] = d.ax[c];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) L[s][k] = line[(side * 8 + k) * nf + f];
    }
  }
  bool lin = false;
  if constexpr (LEARNED) {
    // the tangential stencil inherits the classification of the normal
    // stencils its line values come from: linear when the five cells' normal
    // stencils and the centre cell's stencil across them all took the linear
    // weights, and the DF sensor is zero on the tangential stencil
    if (valid) {
      unsigned smooth;
      double amax = 0.0;
#pragma unroll
      for (int s = 0; s < 5; s++) amax = fmax(amax, a[s]);
      if (DIR == 0) {
        const long long c0 = (long long)(side == 0 ? i - 1 : i) * d.ny + j;
        smooth = d.lin[1][c0];
#pragma unroll
        for (int s = 0; s < 5; s++) smooth &= d.lin[0][c0 - 2 + s];
      } else {
        const int cj = side == 0 ? j - 1 : j;
        smooth = d.lin[0][(long long)i * d.ny + cj];
#pragma unroll
        for (int s = 0; s < 5; s++) smooth &= d.lin[1][(long long)(i + 2 - s) * d.ny + cj];
      }
      lin = smooth != 0 && amax == 0.0;
      if (!LISTED) count_linear(d.lstat + 2 * kStatSlots, lin);
      slin[side][tl] = lin ? 1 : 0;
    }
  }
  if (valid) {
    const double h = DIR == 0 ? d.dy : d.dx, ih = DIR == 0 ? d.idy : d.idx;
    double gp[2][12];
    if (lin) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        weno5_linear_2gauss(gp[0][v], gp[0][8 + v], gp[1][v], gp[1][8 + v], L[0][v], L[1][v], L[2][v], L[3][v],
                            L[4][v], h, ih);
        double t0, t1;
        weno5_linear_2gauss(gp[0][4 + v], t0, gp[1][4 + v], t1, L[0][4 + v], L[1][4 + v], L[2][4 + v],
                            L[3][4 + v], L[4][4 + v], h, ih);
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (reduce_order(gp[q], d.gas)) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            gp[q][v] = L[2][v];
            gp[q][4 + v] = L[2][4 + v];
            gp[q][8 + v] = 0.0;
          }
        }
      }
    } else {
      weno5_tangent(gp, L, a, h, ih, d.recon_var, d.gas);
    }
#pragma unroll
    for (int k = 0; k < 12; k++)
#pragma unroll
      for (int q = 0; q < 2; q++) sg[side][k][tl][q] = gp[q][k];
  }
  __syncthreads();
  // phase 2: lane pair (2 fp, 2 fp + 1) takes the two Gauss points of face fp
  const int fp = threadIdx.x >> 1, q = threadIdx.x & 1;
  const long long slot2 = (long long)blk * TF + fp;
  const bool valid2 = LISTED ? slot2 < nlist : slot2 < nfaces;
  if (valid2) {
    const long long r2 = LISTED ? (long long)list[slot2] : slot2;
    const int i2 = (int)(r2 / nj) + d.G, j2 = (int)(r2 % nj) + d.G;
    const long long f2 = DIR == 0 ? (long long)i2 * d.ny + j2 : (long long)i2 * (d.ny + 1) + j2;
    const Strided L{&sg[0][0][fp][q], 2 * TF}, R{&sg[1][0][fp][q], 2 * TF};
    // both sides smooth: the smooth-flow flux (learned path)
    const bool smooth = LEARNED && !LISTED && d.smooth_flux && (slin[0][fp] & slin[1][fp]);
    flux_pair_tail<DIR>(d, q, f2, nf, L, R, stage, dt, smooth);
  }
}
template <int DIR, bool LEARNED>
__global__ void __launch_bounds__(2 * kTanFlux) __attribute__((amdgpu_waves_per_eu(GKS_FLUX_WAVES)))
k_tanflux(Dev d, int stage, double dt) {
  tanflux_body<DIR, LEARNED, kTanFlux>(d, xcd_block(), stage, dt);
}

// WENO7 / WENO9 tangential reconstruction + flux in one launch, as
// k_tanflux does for WENO5: a workgroup of 128 lanes takes 64 consecutive
// faces.  Phase 1, lane (side, face): tangent79_side into an LDS tile of the
// NG Gauss-point states of both sides (2 x 12 x 64 x NG doubles: 36 KB for
// WENO7, 48 KB for WENO9); phase 2: k_flux's flux point for every (Gauss
// point, face) of the tile, in two rounds of 128 (WENO7: 192 points, the
// second round half full; WENO9: 256).  The Gauss-point states never reach
// HBM.  WENO7 runs two waves per SIMD, WENO9 one (as its separate
// tangential launch: at two its registers spill).
template <int DIR, int ORDER>
__global__ void __launch_bounds__(2 * kTanFlux) __attribute__((amdgpu_waves_per_eu(ORDER == 9 ? 1 : GKS_FLUX_WAVES)))
k_tanflux79(Dev d, int stage, double dt) {
  constexpr int NG = ORDER == 7 ? 3 : 4, TF = kTanFlux;
  __shared__ double sg[2][12][TF][NG];  // [side][component][face][gauss point]
  const unsigned blk = xcd_block();
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  const long long nfaces = (long long)ni * nj;
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  {
    const int tl = threadIdx.x % TF, side = threadIdx.x / TF;
    const long long r = (long long)blk * TF + tl;
    if (r < nfaces) {
      const int i = (int)(r / nj) + d.G, j = (int)(r % nj) + d.G;
      tangent79_side<DIR, ORDER>(d, side, i, j, [&](int q, int k, double v) { sg[side][k][tl][q] = v; });
    }
  }
  __syncthreads();
  // two rounds written out (a loop around the flux point spills)
  auto point = [&](int it) {
    const int q = it / TF, fp = it - q * TF;
    const long long r2 = (long long)blk * TF + fp;
    if (it >= NG * TF || r2 >= nfaces) return;
    const int i2 = (int)(r2 / nj) + d.G, j2 = (int)(r2 % nj) + d.G;
    const long long f2 = DIR == 0 ? (long long)i2 * d.ny + j2 : (long long)i2 * (d.ny + 1) + j2;
    const Strided L{&sg[0][0][fp][q], TF * NG}, R{&sg[1][0][fp][q], TF * NG};
    flux_point_tail<DIR>(d, q, f2, nf, L, R, stage, dt);
  };
  static_assert(NG * TF <= 4 * TF, "at most two rounds");
  point(threadIdx.x);
  point(threadIdx.x + 2 * TF);
}

// ------------------------------------ learned path: the flux in two launches
// With the smooth-flow flux on, the learned path's flux sweep runs in two
// launches, so that the many smooth faces do not carry the register budget
// (two waves per SIMD) of the full flux point.  k_tanflux_cheap takes every
// face: a face whose two tangential stencils are linear (the tangential
// stencil rule of tanflux_body) gets its Gauss-point states on each side
// from the linear weights and the smooth-flow flux; every other face is
// appended to tlist[DIR] (one wave-aggregated atomic per wave).
// k_tanflux_list then runs tanflux_body over the listed faces (WENO5-AO-DF
// or linear per side, full GKS2D) as a grid-stride loop over tiles of the
// list.  Each face gets the result k_tanflux<DIR, true> gives it; the list
// order varies from run to run, the results do not.
// the linear-weight rule for the tangential stencil of (side, face (i, j))
template <int DIR> __device__ __forceinline__ bool tangential_linear(const Dev& d, int side, int i, int j) {
  double amax = 0.0;
  unsigned smooth;
  if (DIR == 0) {
    const long long c0 = (long long)(side == 0 ? i - 1 : i) * d.ny + j;
#pragma unroll
    for (int s = 0; s < 5; s++) amax = fmax(amax, d.ay[c0 - 2 + s]);
    smooth = d.lin[1][c0];
#pragma unroll
    for (int s = 0; s < 5; s++) smooth &= d.lin[0][c0 - 2 + s];
  } else {
    const int cj = side == 0 ? j - 1 : j;
#pragma unroll
    for (int s = 0; s < 5; s++) amax = fmax(amax, d.ax[(long long)(i + 2 - s) * d.ny + cj]);
    smooth = d.lin[0][(long long)i * d.ny + cj];
#pragma unroll
    for (int s = 0; s < 5; s++) smooth &= d.lin[1][(long long)(i + 2 - s) * d.ny + cj];
  }
  return smooth != 0 && amax == 0.0;
}
// Two phases over a tile of TF faces, R = TF / kCheapCols rows (i) by
// kCheapCols faces (j).  Phase 1, thread (side, face): the tangential rule
// for its side and, when linear, that side's two Gauss-point states from
// the linear weights into LDS; the stencils' line values, and the DF
// sensor and linear flags of their source cells, are staged in LDS first
// (each value loaded once per tile).  Phase 2, lane pair per face: a face
// with both sides linear gets the smooth-flow flux from the two states;
// the first TF threads list every other face.
//
// With fused_lines the line values are not loaded: the tile forms them
// from the cell averages (the linear reconstruction of the normal sweep,
// a thread per source cell, from averages staged in LDS), and a listed
// face stores the ones of its stencil's linear source cells for the list
// launch (the WENO cells' values come from k_normal_list).
template <int DIR, int TF>
__device__ __forceinline__ void tanflux_cheap_body(const Dev& d, unsigned blk, unsigned ntiles, int stage,
                                                   double dt) {
  constexpr int C = kCheapCols, R = TF / C;
  static_assert(R * C == TF, "whole rows per tile");
  // staged extents.  x faces (stencils along j): line values of the R face
  // rows at columns j0 - 2 .. j0 + C + 1 (XW), their source cells' rows
  // i0 - 1 .. i0 + R - 1, and those cells' normal stencils, rows i0 - 3 ..
  // i0 + R + 1.  y faces (stencils across rows): line values of rows i0 - 2
  // .. i0 + R + 1 at columns j0 .. j0 + C - 1, their source cells' columns
  // j0 - 1 .. j0 + C - 1, and those cells' stencils, columns j0 - 3 .. j0 +
  // C + 1.
  constexpr int XW = C + 4;
  constexpr int LR = DIR == 0 ? R : R + 4, LW = DIR == 0 ? XW : C;      // line values [LR][16][LW]
  constexpr int AR = DIR == 0 ? R + 1 : R + 4, AW = DIR == 0 ? XW : C + 1;  // source flags [AR][AW]
  constexpr int UR = DIR == 0 ? R + 5 : R + 4, UW = DIR == 0 ? XW : C + 5;  // cell averages [UR][4][UW]
  constexpr int kSg = 2 * 12 * TF * 2, kStage = LR * 16 * LW + AR * AW + UR * 4 * UW;
  __shared__ double smem[kSg > kStage ? kSg : kStage];
  double(&sg)[2][12][TF][2] = *reinterpret_cast<double(*)[2][12][TF][2]>(smem);  // [side][component][face][point]
  double* st = smem;                // line values
  double* sa = st + LR * 16 * LW;   // DF sensor of the source cells (alpha_y for x faces, alpha_x for y)
  double* su = sa + AR * AW;        // cell averages (fused_lines)
  __shared__ unsigned char slin[2][TF];
  __shared__ unsigned char sfl[AR * AW];  // linear flag of the source cells' normal stencils
  const int ni = DIR == 0 ? d.NX + 1 : d.NX, nj = DIR == 0 ? d.NY : d.NY + 1;
  const long long tcols = cheap_tile_cols(d, DIR);
  const int bi = (int)(blk / tcols), bj = (int)(blk % tcols);
  const int i0 = d.G + bi * R, j0 = d.G + bj * C;
  // tile-local face t: whether it exists, its indices and its ordinal
  // (row-major over the interior faces)
  auto face_of = [&](int t, bool& ok, int& fi, int& fj) -> long long {
    const int ri = bi * R + t / C, cj = bj * C + t % C;
    ok = ri < ni && cj < nj;
    fi = (ok ? ri : 0) + d.G;
    fj = (ok ? cj : 0) + d.G;
    return (long long)ri * nj + cj;
  };
  const int tl = threadIdx.x % TF, hi = threadIdx.x / TF;
  const int rl = tl / C, cl = tl % C;
  bool valid;
  int i, j;
  const long long r = face_of(tl, valid, i, j);
  const long long nf = DIR == 0 ? d.nfx : d.nfy;
  const bool fused = (d.fused_lines >> DIR) & 1;
  auto clamp = [](int v, int hi_) { return v < 0 ? 0 : (v < hi_ ? v : hi_ - 1); };
  {
    const int side = hi;
    // without the smooth-flow flux every face goes to the list
    const bool want = valid && d.smooth_flux;
    // every load of a thread is issued before the first LDS store
    constexpr int NA = (AR * AW + 2 * TF - 1) / (2 * TF);
    constexpr int NU = (UR * 4 * UW + 2 * TF - 1) / (2 * TF), NL = (LR * 16 * LW + 2 * TF - 1) / (2 * TF);
    double va[NA];
    unsigned char vl[NA];
#pragma unroll
    for (int u = 0; u < NA; u++) {
      const int e = threadIdx.x + u * 2 * TF;
      if (e < AR * AW) {
        const int ar = e / AW, ac = e - ar * AW;
        const long long g = DIR == 0 ? (long long)clamp(i0 - 1 + ar, d.nx) * d.ny + clamp(j0 - 2 + ac, d.ny)
                                     : (long long)clamp(i0 - 2 + ar, d.nx) * d.ny + clamp(j0 - 1 + ac, d.ny);
        va[u] = DIR == 0 ? d.ay[g] : d.ax[g];
        vl[u] = d.lin[DIR][g];
      }
    }
    // the tangential rule's flag of the centre cell across (x faces: the
    // source cell's y stencil; y faces: its x stencil)
    const unsigned char across =
        !valid ? 0
               : (DIR == 0 ? d.lin[1][(long long)(i - 1 + side) * d.ny + j]
                           : d.lin[0][(long long)i * d.ny + j - 1 + side]);
    if (fused) {
      double vu[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = threadIdx.x + u * 2 * TF;
        if (e < UR * 4 * UW) {
          const int ur = e / (4 * UW), v = (e / UW) & 3, uc = e % UW;
          const long long g = DIR == 0 ? (long long)clamp(i0 - 3 + ur, d.nx) * d.ny + clamp(j0 - 2 + uc, d.ny)
                                       : (long long)clamp(i0 - 2 + ur, d.nx) * d.ny + clamp(j0 - 3 + uc, d.ny);
          vu[u] = d.U[v * d.nc + g];
        }
      }
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = threadIdx.x + u * 2 * TF;
        if (e < UR * 4 * UW) su[e] = vu[u];
      }
    } else if (d.smooth_flux) {
      const double* line = DIR == 0 ? d.line_x : d.line_y;
      double vv[NL];
#pragma unroll
      for (int u = 0; u < NL; u++) {
        const int e = threadIdx.x + u * 2 * TF;
        if (e < LR * 16 * LW) {
          const int lr = e / (16 * LW), k = (e / LW) & 15, lc = e % LW;
          // x faces: rows 0 .. nx, columns 0 .. ny - 1; y faces: rows 0 ..
          // nx - 1, columns 0 .. ny
          const long long g = DIR == 0
                                  ? (long long)clamp(i0 + lr, d.nx + 1) * d.ny + clamp(j0 - 2 + lc, d.ny)
                                  : (long long)clamp(i0 - 2 + lr, d.nx) * (d.ny + 1) + clamp(j0 + lc, d.ny + 1);
          vv[u] = line[k * nf + g];
        }
      }
#pragma unroll
      for (int u = 0; u < NL; u++) {
        const int e = threadIdx.x + u * 2 * TF;
        if (e < LR * 16 * LW) st[e] = vv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < NA; u++) {
      const int e = threadIdx.x + u * 2 * TF;
      if (e < AR * AW) {
        sa[e] = va[u];
        sfl[e] = vl[u];
      }
    }
    __syncthreads();
    if (fused) {
      // the linear reconstruction of every source cell of the tile: x
      // faces, cell (i0 - 1 + cr, j0 - 2 + cc) with its stencil rows cr ..
      // cr + 4, its lower face side 1 of face row cr - 1 and its upper face
      // side 0 of face row cr; y faces, (row lr, column c) forms both sides
      // of line record (lr, c): side 0 the upper face of cell j0 + c - 1
      // (averages c .. c + 4), side 1 the lower face of cell j0 + c (c + 1
      // .. c + 5), in the x-like frame of the y sweep
      constexpr int NI = DIR == 0 ? AR * XW : LR * C;
#pragma unroll
      for (int u = 0; u < (NI + 2 * TF - 1) / (2 * TF); u++) {
        const int e = threadIdx.x + u * 2 * TF;
        if (e < NI) {
          if constexpr (DIR == 0) {
            const int cr = e / XW, cc = e - cr * XW;
            double w[5][4], lo[8], ro[8];
#pragma unroll
            for (int s = 0; s < 5; s++)
#pragma unroll
              for (int v = 0; v < 4; v++) w[s][v] = su[((cr + s) * 4 + v) * UW + cc];
            linear_normal(lo, ro, w, d.dx, d.idx, d.gas);
            if (cr > 0) {
#pragma unroll
              for (int k = 0; k < 8; k++) st[((cr - 1) * 16 + 8 + k) * LW + cc] = lo[k];
            }
            if (cr < R) {
#pragma unroll
              for (int k = 0; k < 8; k++) st[(cr * 16 + k) * LW + cc] = ro[k];
            }
          } else {
            const int lr = e / C, c = e - lr * C;
#pragma unroll
            for (int sd = 0; sd < 2; sd++) {
              double w[5][4], lo[8], ro[8];
#pragma unroll
              for (int s = 0; s < 5; s++) {
                const double* a = su + lr * 4 * UW + c + sd + s;
                w[s][0] = a[0];  // YchangetoX
                w[s][1] = a[2 * UW];
                w[s][2] = -a[UW];
                w[s][3] = a[3 * UW];
              }
              linear_normal(lo, ro, w, d.dy, d.idy, d.gas);
#pragma unroll
              for (int k = 0; k < 8; k++) st[(lr * 16 + sd * 8 + k) * LW + c] = sd == 0 ? ro[k] : lo[k];
            }
          }
        }
      }
      __syncthreads();
    }
    // stencil entry s of this face and side: x faces, line record (i, j - 2
    // + s) at staged (rl, cl + s), source cell row rl + side; y faces, line
    // record (i + 2 - s, j) at staged (rl + 4 - s, cl), source cell (rl + 4
    // - s, cl + side)
    auto lidx = [&](int s, int k) {
      return DIR == 0 ? (rl * 16 + side * 8 + k) * LW + cl + s : ((rl + 4 - s) * 16 + side * 8 + k) * LW + cl;
    };
    auto aidx = [&](int s) { return DIR == 0 ? (rl + side) * AW + cl + s : (rl + 4 - s) * AW + cl + side; };
    double L[5][8];
    if (valid && (want || fused)) {
#pragma unroll
      for (int s = 0; s < 5; s++)
#pragma unroll
        for (int k = 0; k < 8; k++) L[s][k] = st[lidx(s, k)];
    }
    bool lin = false;
    if (valid) {
      // tangential_linear<DIR> on the staged values
      double amax = 0.0;
      unsigned smooth = across;
#pragma unroll
      for (int s = 0; s < 5; s++) {
        amax = fmax(amax, sa[aidx(s)]);
        smooth &= sfl[aidx(s)];
      }
      lin = smooth != 0 && amax == 0.0;
    }
    slin[side][tl] = lin ? 1 : 0;
    // the Gauss-point tile below overwrites the staged values
    __syncthreads();
    if (valid) count_linear(d.lstat + 2 * kStatSlots, lin);
    if (fused && valid && !(d.smooth_flux && slin[0][tl] && slin[1][tl])) {
      // a listed face: the line values of its stencil's linear source
      // cells are stored for the list launch, each by the first listed face
      // of the tile that reads it (x faces: entry s is read by faces j - 4
      // + s .. j of the row, so it is stored unless one of faces tl - 1 ..
      // tl - (4 - s) is listed; y faces: entry s by rows i - s .. i + 4 -
      // s of the column, stored unless one of rows i - 1 .. i - s is)
      bool store[5];
      if constexpr (DIR == 0) {
        bool seen = false;
#pragma unroll
        for (int k = 0; k <= 4; k++) {
          if (k > 0 && cl - k >= 0 && !(d.smooth_flux && slin[0][tl - k] && slin[1][tl - k])) seen = true;
          store[4 - k] = !seen;
        }
      } else {
        bool seen = false;
#pragma unroll
        for (int s = 0; s <= 4; s++) {
          if (s > 0 && rl - s >= 0 && !(d.smooth_flux && slin[0][tl - s * C] && slin[1][tl - s * C])) seen = true;
          store[s] = !seen;
        }
      }
      double* line = DIR == 0 ? d.line_x : d.line_y;
#pragma unroll
      for (int s = 0; s < 5; s++) {
        if (store[s] && sfl[aidx(s)]) {
          const long long f = DIR == 0 ? (long long)i * d.ny + j - 2 + s : (long long)(i + 2 - s) * (d.ny + 1) + j;
#pragma unroll
          for (int k = 0; k < 8; k++) line[(side * 8 + k) * nf + f] = L[s][k];
        }
      }
    }
    if (lin && want) {
      const double h = DIR == 0 ? d.dy : d.dx, ih = DIR == 0 ? d.idy : d.idx;
      double gp[2][12];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        weno5_linear_2gauss(gp[0][v], gp[0][8 + v], gp[1][v], gp[1][8 + v], L[0][v], L[1][v], L[2][v], L[3][v],
                            L[4][v], h, ih);
        double t0, t1;
        weno5_linear_2gauss(gp[0][4 + v], t0, gp[1][4 + v], t1, L[0][4 + v], L[1][4 + v], L[2][4 + v],
                            L[3][4 + v], L[4][4 + v], h, ih);
      }
#pragma unroll
      for (int q = 0; q < 2; q++) {
        if (reduce_order(gp[q], d.gas)) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            gp[q][v] = L[2][v];
            gp[q][4 + v] = L[2][4 + v];
            gp[q][8 + v] = 0.0;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 12; k++)
#pragma unroll
        for (int q = 0; q < 2; q++) sg[side][k][tl][q] = gp[q][k];
    }
  }
  __syncthreads();
  // the first TF threads (whole waves: TF is a multiple of 64) list the
  // faces of the tile that are not both sides linear
  {
    const bool cheap = valid && d.smooth_flux && (slin[0][tl] & slin[1][tl]);
    const int sh = list_shard(blk, ntiles);
    if (hi == 0)
      list_append(shard_count(d.tcount, DIR, sh), d.tlist[DIR] + (size_t)sh * d.tcap[DIR], valid && !cheap,
                  (unsigned)r);
  }
  // phase 2: lane pair (2 fp, 2 fp + 1) takes the two Gauss points of face fp
  const int fp = threadIdx.x >> 1, q = threadIdx.x & 1;
  bool ok2;
  int i2, j2;
  (void)face_of(fp, ok2, i2, j2);
  if (!(ok2 && d.smooth_flux && (slin[0][fp] & slin[1][fp]))) return;
  const long long f2 = DIR == 0 ? (long long)i2 * d.ny + j2 : (long long)i2 * (d.ny + 1) + j2;
  const Strided L{&sg[0][0][fp][q], 2 * TF}, R2{&sg[1][0][fp][q], 2 * TF};
  flux_pair_tail<DIR>(d, q, f2, nf, L, R2, stage, dt, true);
}
// faces per tile of the single-problem smooth-face launch: 8 x 16 for x
// faces, 4 x 16 for y faces (each direction's faster shape)
template <int DIR> constexpr int kCheapFaces = DIR == 0 ? 128 : 64;
template <int DIR>
__global__ void __launch_bounds__(2 * kCheapFaces<DIR>) k_tanflux_cheap(Dev d, int stage, double dt) {
  tanflux_cheap_body<DIR, kCheapFaces<DIR>>(d, xcd_block(), gridDim.x, stage, dt);
}
// one tile of tlist[DIR] per workgroup, launched for as many tiles as there
// are faces (the count is only known on the device; workgroups past it
// return at once).  A grid-stride loop over the tiles instead spills: the
// body's register demand grows past the two-wave budget inside the loop.
// The count is uniform over the workgroup, so every thread reaches the same
// barriers.
// faces per tile of the single-problem list launch
constexpr int kListFaces = 128;
template <int DIR>
__global__ void __launch_bounds__(2 * kListFaces) __attribute__((amdgpu_waves_per_eu(GKS_FLUX_WAVES)))
k_tanflux_list(Dev d, int stage, double dt) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(d.lstat + kListedWord + DIR, (unsigned long long)list_total(d.lpre + (KL_FLUX_X + DIR) * kListRecord));
  int s;
  unsigned tb, n;
  if (list_tile(d.lpre + (KL_FLUX_X + DIR) * kListRecord, blockIdx.x, s, tb, n))
    tanflux_body<DIR, true, kListFaces, true>(d, tb, stage, dt, d.tlist[DIR] + (size_t)s * d.tcap[DIR], n);
}


// ------------------------------------------------------------------ update
// cell update of Update_with_gauss (time_advance.cpp:442-474): U = U_old +
// sum over Gauss points of len * (x_S - x_N + x_W - x_E), in the reference's
// summation order, divided by the cell area
// the update of cell t of the interior; u: its new state
__device__ __forceinline__ void update_cell(const Dev& d, long long t, double* u_out) {
  const int i = (int)(t / d.NY) + d.G, j = (int)(t % d.NY) + d.G;
  const long long c = (long long)i * d.ny + j;
  const long long fb = (long long)i * (d.ny + 1) + j, ft = fb + 1;
  const long long fl_ = (long long)i * d.ny + j, fr = (long long)(i + 1) * d.ny + j;
  double total[4] = {0.0, 0.0, 0.0, 0.0};
  // the fast build's fused flux stores one Gauss-point sum per face
  const int ng = GKS_FAST && d.presum ? 1 : d.gp;
  for (int q = 0; q < ng; q++) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const long long oy = (long long)(q * 4 + v) * d.nfy, ox = (long long)(q * 4 + v) * d.nfx;
      total[v] += d.dx * d.xcy[oy + fb];
      total[v] += -d.dx * d.xcy[oy + ft];
      total[v] += d.dy * d.xcx[ox + fl_];
      total[v] += -d.dy * d.xcx[ox + fr];
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double u = d.Uold[v * d.nc + c];
    u += total[v] / d.area;
    d.U[v * d.nc + c] = u;
    u_out[v] = u;
  }
}
__device__ __forceinline__ void update_body(const Dev& d, unsigned blk) {
  const long long t = (long long)blk * kBlock + threadIdx.x;
  double u[4];
  if (t < (long long)d.NX * d.NY) update_cell(d, t, u);
}
__global__ void __launch_bounds__(kBlock) k_update(Dev d) { update_body(d, blockIdx.x); }
// the update and, over the cells it wrote, what Convar_to_Primvar would
// find on the new state: the blow-up test and the largest Dtx speed, into
// the block's partial pair (as k_primvar: its test also covers the ghost
// cells, which the boundary fill makes copies or mirror images of interior
// cells, so it finds what this one finds)
__device__ __forceinline__ void update_stat_body(const Dev& d, unsigned blk, double* partial) {
  __shared__ double smax[kBlock];
  __shared__ int sbad[kBlock];
  const long long t = (long long)blk * kBlock + threadIdx.x;
  double tmax = 0.0;
  int bad = 0;
  if (t < (long long)d.NX * d.NY) {
    double u[4], p[4];
    update_cell(d, t, u);
    convar_to_primvar(p, u, d.gas);
    if (p[0] < 0 || p[3] < 0 || p[0] != p[0] || p[3] != p[3]) bad = 1;
    const double tmp = sqrt(p[1] * p[1] + p[2] * p[2]) + sqrt(d.gas.gamma * p[3] / p[0]);
    if (tmp > tmax) tmax = tmp;  // a NaN speed is skipped, as Dtx's comparison does
  }
  smax[threadIdx.x] = tmax;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double o = smax[threadIdx.x + s];
      if (o > smax[threadIdx.x]) smax[threadIdx.x] = o;
      sbad[threadIdx.x] |= sbad[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blk] = smax[0];
    partial[2 * blk + 1] = sbad[0] ? 1.0 : 0.0;
  }
}
__global__ void __launch_bounds__(kBlock) k_update_stat(Dev d, double* partial) {
  update_stat_body(d, blockIdx.x, partial);
}

// Update_alpha (reconstruction.cpp:1466-1534): the reference bounds j by
// nodex (:1528); kept for nodey >= nodex, clipped to nodey otherwise.
__device__ __forceinline__ void alpha_body(const Dev& d, unsigned blk) {
  const int nj = d.NY < d.NX ? d.NY : d.NX;
  long long t = (long long)blk * kBlock + threadIdx.x;
  if (t >= (long long)d.NX * nj) return;
  const int i = (int)(t / nj) + d.G, j = (int)(t % nj) + d.G;
  auto face_alpha = [&](const double* ak, long long nf, long long f) {
    if (d.presum) return ak[f];  // formed by the fused flux launch
    double s;
    if (d.gp == 2) s = 0.5 * (ak[f] + ak[nf + f]);
    else if (d.gp == 3) s = (ak[f] + ak[nf + f] + ak[2 * nf + f]) / 3.0;
    else s = 0.25 * (ak[f] + ak[nf + f] + ak[2 * nf + f] + ak[3 * nf + f]);
    return s < 2.0 ? 0.0 : s;
  };
  const double al = face_alpha(d.akx, d.nfx, (long long)i * d.ny + j);
  const double ar = face_alpha(d.akx, d.nfx, (long long)(i + 1) * d.ny + j);
  const double ad = face_alpha(d.aky, d.nfy, (long long)i * (d.ny + 1) + j);
  const double au = face_alpha(d.aky, d.nfy, (long long)i * (d.ny + 1) + j + 1);
  const long long c = (long long)i * d.ny + j;
  d.ax[c] = al + ar;
  d.ay[c] = ad + au;
}
__global__ void __launch_bounds__(kBlock) k_alpha(Dev d) { alpha_body(d, blockIdx.x); }

// leaf check of the linear-weight reconstructions (learned path): per
// 5-cell stencil the fifth-order candidate's values and derivatives at the
// centre cell's left and right faces and at its two tangential Gauss points
__global__ void __launch_bounds__(kBlock) k_linear_points(long long n, const double* W, double h, double* out) {
  long long t = gtid();
  if (t >= n) return;
  double w[5];
#pragma unroll
  for (int s = 0; s < 5; s++) w[s] = W[t * 5 + s];
  const double ih = 1.0 / h;
  double o[8];
  weno5_linear_face(0, o[0], o[1], w[0], w[1], w[2], w[3], w[4], h, ih);
  weno5_linear_face(1, o[2], o[3], w[0], w[1], w[2], w[3], w[4], h, ih);
  weno5_linear_2gauss(o[4], o[5], o[6], o[7], w[0], w[1], w[2], w[3], w[4], h, ih);
#pragma unroll
  for (int k = 0; k < 8; k++) out[t * 8 + k] = o[k];
}

// leaf parity: centre + GKS2D over n independent points
// (smooth: the smooth-flow flux instead, no centre output)
__global__ void __launch_bounds__(kBlock) k_flux_points(Gas g, long long n, const double* L, const double* R,
                                                         double dt, double* C, double* F, double* D, int smooth) {
  long long t = gtid();
  if (t >= n) return;
  double l[12], r[12], c[12], f[4], df[4];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    l[k] = L[t * 12 + k];
    r[k] = R[t * 12 + k];
  }
  if (smooth) {
    gks_point_smooth(f, df, l, r, dt, g);
  } else {
    gks_point(f, df, c, l, r, dt, g);
#pragma unroll
    for (int k = 0; k < 12; k++) C[t * 12 + k] = c[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    F[t * 4 + k] = f[k];
    D[t * 4 + k] = df[k];
  }
}

// ------------------------------------------------------------------ batches
// A batched launch covers the tiles of every problem of a batch: workgroup
// b (after the XCD mapping) finds its problem by a binary search of the
// work class's tile prefix table, and runs the same body as the single-
// problem kernel on that problem's device state.  Problems not active in
// this step return at once.
__device__ __forceinline__ int batch_slot(const BatchArgs& B, int kind, unsigned gb) {
  const unsigned* st = B.start + kind * (B.np + 1);
  int lo = 0, hi = B.np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (st[mid] <= gb) lo = mid;
    else hi = mid;
  }
  return lo;
}
// The problem's descriptor is copied out of global memory once: read
// through a reference, every field would be reloaded after each store the
// body makes (the compiler cannot rule out that a store aliases it).
// problem p's descriptor as a launch of B sees it (stage 1: U and Uold
// exchanged)
__device__ __forceinline__ Dev batch_dev(const BatchArgs& B, int p) {
  Dev d = B.dev[p];
  if (B.swap) {
    double* u = d.U;
    d.U = d.Uold;
    d.Uold = u;
  }
  return d;
}
#define GKS_BATCH_TILE(kind)                                    \
  const unsigned gb_ = xcd_block();                             \
  const int q_ = batch_slot(B, kind, gb_);                      \
  const int p_ = B.pid[q_];                                     \
  if (!B.st[p_].active) return;                                 \
  const Dev d = batch_dev(B, p_);                               \
  const unsigned blk = gb_ - B.start[(kind) * (B.np + 1) + q_]; \
  (void)d;                                                      \
  (void)blk

__global__ void __launch_bounds__(kBlock) kb_bc(BatchArgs B, int side) {
  GKS_BATCH_TILE(side < 2 ? KB_BC_LR : KB_BC_DU);
  bc_body(d, blk, side, B.st[p_].bc[side]);
}
// one workgroup per problem: the partials of its `kind` tiles
__global__ void __launch_bounds__(kBlock) kb_reduce(BatchArgs B, int slot, int kind, int flag) {
  const int q = blockIdx.x, p = B.pid[q];
  if (!B.st[p].active) return;
  const unsigned* st = B.start + kind * (B.np + 1);
  reduce_body(B.dev[p].partial, (int)(st[q + 1] - st[q]), B.dev[p].stat, slot, flag);
}
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES))) kb_normal(BatchArgs B) {
  GKS_BATCH_TILE(DIR == 0 ? KB_NORMAL_X : KB_NORMAL_Y);
  normal_body<DIR>(d, blk);
}
// the learned batch: both directions' classifiers of every problem (as
// k_classify; the batch forms every linear line value in the smooth-face
// flux launch)
static_assert((kFusedLineDirs & 3) == 3, "the learned batch forms every line value in the smooth-face launch");
__global__ void __launch_bounds__(kBlock) kb_classify(BatchArgs B) {
  GKS_BATCH_TILE(KB_NORMAL_X);
  const unsigned* st = B.start + KB_NORMAL_X * (B.np + 1);
  classify_body(d, blk, st[q_ + 1] - st[q_]);
}
// Learned path, work lists: after the sweep that fills every problem's
// list of one kind, kb_list_prefix forms each problem's tile prefix of
// that list and packs the lists' tiles at the front of the next launch
// (problem q's tiles start at lstart[q]); the list kernel is launched with
// as many workgroups as the sweep had + kListShards per problem (a bound),
// and the ones past lstart[np] return after one load.  One workgroup: a
// lane per problem, then lane 0 scans the problems' tile counts (np is a
// batch's problem count, tens).
__global__ void __launch_bounds__(64) kb_list_prefix(BatchArgs B, int kind, int per_tile) {
  unsigned* ls = B.lstart + kind * (B.np + 1);
  // each problem's tile prefix (one lane per problem), its tiles into ls
  for (int q = threadIdx.x; q < B.np; q += 64) {
    const int p = B.pid[q];
    const Dev& d = B.dev[p];
    unsigned* cnt = kind < KL_FLUX_X ? d.ncount : d.tcount;
    ls[q] = (B.st[p].active && cnt) ? list_prefix(cnt, kind & 1, per_tile, d.lpre + kind * kListRecord) : 0u;
  }
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x != 0) return;
  unsigned acc = 0;
  for (int q = 0; q < B.np; q++) {
    const unsigned t = ls[q];
    ls[q] = acc;
    acc += t;
  }
  ls[B.np] = acc;
}
#define GKS_BATCH_LIST_TILE(kind)                           \
  const unsigned* ls_ = B.lstart + (kind) * (B.np + 1);     \
  const unsigned gb_ = blockIdx.x;                          \
  if (gb_ >= ls_[B.np]) return;                             \
  int q_ = 0;                                               \
  {                                                         \
    int hi_ = B.np;                                         \
    while (hi_ - q_ > 1) {                                  \
      const int mid_ = (q_ + hi_) >> 1;                     \
      if (ls_[mid_] <= gb_) q_ = mid_;                      \
      else hi_ = mid_;                                      \
    }                                                       \
  }                                                         \
  const int p_ = B.pid[q_];                                 \
  const Dev d = batch_dev(B, p_);                           \
  const unsigned blk = gb_ - ls_[q_]
// the cells each problem's learned normal sweep listed (as k_normal_list)
template <int DIR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GKS_NORMAL_WAVES)))
kb_normal_list(BatchArgs B) {
  GKS_BATCH_LIST_TILE(DIR == 0 ? KL_NORMAL_X : KL_NORMAL_Y);
  normal_list_body<DIR>(d, blk);
}
template <int DIR, bool LEARNED, int TF>
__global__ void __launch_bounds__(2 * TF) __attribute__((amdgpu_waves_per_eu(GKS_FLUX_WAVES)))
kb_tanflux(BatchArgs B, int stage) {
  GKS_BATCH_TILE(DIR == 0 ? KB_FLUX_X : KB_FLUX_Y);
  tanflux_body<DIR, LEARNED, TF>(d, blk, stage, B.st[p_].dt);
}
// the learned flux sweep in two launches over every problem (as
// k_tanflux_cheap / k_tanflux_list): smooth faces, then each problem's
// list, one list tile per workgroup of that problem's face tiles
template <int DIR, int TF>
__global__ void __launch_bounds__(2 * TF) kb_tanflux_cheap(BatchArgs B, int stage) {
  GKS_BATCH_TILE(DIR == 0 ? KB_FLUX_X : KB_FLUX_Y);
  const unsigned* st = B.start + (DIR == 0 ? KB_FLUX_X : KB_FLUX_Y) * (B.np + 1);
  tanflux_cheap_body<DIR, TF>(d, blk, st[q_ + 1] - st[q_], stage, B.st[p_].dt);
}
template <int DIR, int TF>
__global__ void __launch_bounds__(2 * TF) __attribute__((amdgpu_waves_per_eu(GKS_FLUX_WAVES)))
kb_tanflux_list(BatchArgs B, int stage) {
  GKS_BATCH_LIST_TILE(DIR == 0 ? KL_FLUX_X : KL_FLUX_Y);
  if (blk == 0 && threadIdx.x == 0)
    atomicAdd(d.lstat + kListedWord + DIR, (unsigned long long)list_total(d.lpre + (KL_FLUX_X + DIR) * kListRecord));
  int s;
  unsigned tb, n;
  if (list_tile(d.lpre + (KL_FLUX_X + DIR) * kListRecord, blk, s, tb, n))
    tanflux_body<DIR, true, TF, true>(d, tb, stage, B.st[p_].dt, d.tlist[DIR] + (size_t)s * d.tcap[DIR], n);
}
// the update of stage 0 (B.swap 0) writes the second buffer from the
// state, stage 1's (B.swap 1) updates the state in place; with the new
// state's blow-up test and CFL speed into the problem's partials
__global__ void __launch_bounds__(kBlock) kb_update_stat(BatchArgs B) {
  GKS_BATCH_TILE(KB_UPDATE);
  Dev du = d;
  du.U = d.Uold;
  du.Uold = B.swap ? d.Uold : d.U;
  update_stat_body(du, blk, d.partial);
}
__global__ void __launch_bounds__(kBlock) kb_alpha(BatchArgs B) {
  GKS_BATCH_TILE(KB_ALPHA);
  if (!d.use_df || d.df_frozen) return;
  alpha_body(d, blk);
}

// start of a step: which problems run, and their dt (Get_CFL / Dtx,
// time_advance.cpp:264-305, from the stale speed of the last primvar, as
// gks2d_run computes it on the host; same operations, not contracted)
__global__ void __launch_bounds__(64) kb_step_begin(BatchArgs B) {
#pragma clang fp contract(off)
  for (int q = threadIdx.x; q < B.np; q += blockDim.x) {
    const int p = B.pid[q];
    BState& st = B.st[p];
    const Dev& d = B.dev[p];
    st.active = 0;
    if (st.done) continue;
    // the last stage-1 update's blow-up flag (or the initial state's) is
    // taken at this step's end, as the stage-0 primvar sweep of this step
    // would have raised it
    d.stat[6] = d.stat[5];
    d.stat[5] = 0.0;
    if (d.stat[4] != 0.0) {
      st.blowup = 1;
      st.done = 1;
      continue;
    }
    if (!(st.t < st.tstop) || st.steps >= st.max_steps) {
      st.done = 1;
      continue;
    }
    const double tmax = d.stat[2 * st.slot];
    double dt = d.dx;
    const double hs[2] = {d.dx, d.dy};
    for (int k = 0; k < 2; k++) {
      const double h = hs[k];
      if (tmax > st.cfl * h / dt) dt = st.cfl * h / tmax;
      if (dt > 0.25 * st.cfl * h * h / d.gas.mu && d.gas.tau_type == 1 && d.gas.mu > 0)
        dt = 0.25 * st.cfl * h * h / d.gas.mu;
    }
    if (st.t + dt > st.tstop) dt = st.tstop - st.t + 1e-16;
    st.dt = dt;
    st.active = 1;
  }
}
// end of a step: a problem whose state blew up stops without advancing t
__global__ void __launch_bounds__(64) kb_step_end(BatchArgs B) {
#pragma clang fp contract(off)
  for (int q = threadIdx.x; q < B.np; q += blockDim.x) {
    const int p = B.pid[q];
    BState& st = B.st[p];
    if (!st.active) continue;
    if (B.dev[p].stat[4] != 0.0 || B.dev[p].stat[6] != 0.0) {
      st.blowup = 1;
      st.done = 1;
      st.active = 0;
      continue;
    }
    st.t = st.t + st.dt;
    st.steps++;
    st.slot = 1;
  }
}

// explicit instantiations: the launchers below are host-only, so the device
// pass must be told which template kernels to emit
template __global__ void k_normal79<0, 7>(Dev);
template __global__ void k_normal79<1, 7>(Dev);
template __global__ void k_normal79<0, 9>(Dev);
template __global__ void k_normal79<1, 9>(Dev);
template __global__ void k_tangent79<0, 7>(Dev);
template __global__ void k_tangent79<1, 7>(Dev);
template __global__ void k_tangent79<0, 9>(Dev);
template __global__ void k_tangent79<1, 9>(Dev);
template __global__ void k_normal<0>(Dev);
template __global__ void k_normal<1>(Dev);
template __global__ void k_tangent<0>(Dev);
template __global__ void k_tangent<1>(Dev);
template __global__ void k_tanflux<0, false>(Dev, int, double);
template __global__ void k_tanflux<1, false>(Dev, int, double);
template __global__ void k_tanflux<0, true>(Dev, int, double);
template __global__ void k_tanflux<1, true>(Dev, int, double);
template __global__ void k_tanflux_cheap<0>(Dev, int, double);
template __global__ void k_tanflux_cheap<1>(Dev, int, double);
template __global__ void k_tanflux_list<0>(Dev, int, double);
template __global__ void k_tanflux_list<1>(Dev, int, double);
template __global__ void kb_normal<0>(BatchArgs);
template __global__ void kb_normal<1>(BatchArgs);
template __global__ void kb_tanflux<0, false, 64>(BatchArgs, int);
template __global__ void kb_tanflux<1, false, 64>(BatchArgs, int);
template __global__ void kb_tanflux<0, true, 64>(BatchArgs, int);
template __global__ void kb_tanflux<1, true, 64>(BatchArgs, int);
template __global__ void kb_tanflux<0, false, 128>(BatchArgs, int);
template __global__ void kb_tanflux<1, false, 128>(BatchArgs, int);
template __global__ void kb_tanflux<0, true, 128>(BatchArgs, int);
template __global__ void kb_tanflux<1, true, 128>(BatchArgs, int);
template __global__ void kb_tanflux_cheap<0, 64>(BatchArgs, int);
template __global__ void kb_tanflux_cheap<1, 64>(BatchArgs, int);
template __global__ void kb_tanflux_cheap<0, 128>(BatchArgs, int);
template __global__ void kb_tanflux_cheap<1, 128>(BatchArgs, int);
template __global__ void kb_tanflux_list<0, 64>(BatchArgs, int);
template __global__ void kb_tanflux_list<1, 64>(BatchArgs, int);
template __global__ void kb_tanflux_list<0, 128>(BatchArgs, int);
template __global__ void kb_tanflux_list<1, 128>(BatchArgs, int);
template __global__ void k_normal_learned<0>(Dev);
template __global__ void k_normal_learned<1>(Dev);
template __global__ void k_normal_list<0>(Dev);
template __global__ void k_normal_list<1>(Dev);
template __global__ void kb_normal_list<0>(BatchArgs);
template __global__ void kb_normal_list<1>(BatchArgs);
template __global__ void k_label<0>(Dev, int, double, float*, float*);
template __global__ void k_label<1>(Dev, int, double, float*, float*);
template __global__ void k_tanflux79<0, 7>(Dev, int, double);
template __global__ void k_tanflux79<1, 7>(Dev, int, double);
template __global__ void k_tanflux79<0, 9>(Dev, int, double);
template __global__ void k_tanflux79<1, 9>(Dev, int, double);
template __global__ void k_flux<0>(Dev, int, double);
template __global__ void k_flux<1>(Dev, int, double);

// ----------------------------------------------------------------- launchers
#if !defined(__HIP_DEVICE_COMPILE__)
static inline unsigned blocks_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

static hipError_t bc(const Dev& d, int side, int kind, hipStream_t s) {
  const long long n = (long long)d.G * (side < 2 ? d.ny : d.nx);
  hipLaunchKernelGGL(k_bc, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, side, kind);
  return hipGetLastError();
}
static hipError_t bc_all(const Dev& d, const int* kind, hipStream_t s) {
  if (d.NX < d.G || d.NY < d.G) {
    // a periodic source may be a ghost cell: the fills one after the other
    for (int side = 0; side < 4; side++) {
      const hipError_t e = bc(d, side, kind[side], s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const long long n = 2LL * d.G * d.ny + 2LL * d.G * d.NX;
  hipLaunchKernelGGL(k_bc_all, dim3(blocks_for(n)), dim3(kBlock), 0, s, d, kind[0], kind[1], kind[2], kind[3]);
  return hipGetLastError();
}
static hipError_t primvar(const Dev& d, const double* prim, int slot, int flag, hipStream_t s) {
  long long nb = blocks_for(d.nc);
  if (nb > kPartialBlocks) nb = kPartialBlocks;
  hipLaunchKernelGGL(k_primvar, dim3((unsigned)nb), dim3(kBlock), 0, s, d, prim, d.partial);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, (const double*)d.partial, (int)nb, d.stat, slot, flag);
  return hipGetLastError();
}
static hipError_t copy_old(const Dev& d, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_old, dim3(blocks_for((long long)d.NX * d.NY)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}
static hipError_t copy_ghosts(const Dev& d, const double* from, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_ghosts, dim3(blocks_for(d.nc)), dim3(kBlock), 0, s, d, from);
  return hipGetLastError();
}
static hipError_t normal(const Dev& d, int dir, hipStream_t s) {
  const unsigned nb = dir == 0 ? blocks_for((long long)(d.NX + 2) * d.ny) : blocks_for((long long)d.nx * (d.NY + 2));
  if (d.order == 7) {
    if (dir == 0) hipLaunchKernelGGL((k_normal79<0, 7>), dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL((k_normal79<1, 7>), dim3(nb), dim3(kBlock), 0, s, d);
  } else if (d.order == 9) {
    if (dir == 0) hipLaunchKernelGGL((k_normal79<0, 9>), dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL((k_normal79<1, 9>), dim3(nb), dim3(kBlock), 0, s, d);
  } else if (d.mlp && (d.fused_lines & 3) == 3) {
    // both sweeps' classifiers in the x call (k_classify stores no line
    // values: both directions form them in the smooth-face launch)
    if (dir == 0) hipLaunchKernelGGL(k_classify, dim3((unsigned)cls_tiles(d)), dim3(kBlock), 0, s, d);
  } else if (d.mlp) {
    // the listed cells follow in normal_list (whose prefix kernel leaves the
    // shard counts cleared for the next sweep)
    if (dir == 0) hipLaunchKernelGGL(k_normal_learned<0>, dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL(k_normal_learned<1>, dim3(nb), dim3(kBlock), 0, s, d);
  } else {
    if (dir == 0) hipLaunchKernelGGL(k_normal<0>, dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL(k_normal<1>, dim3(nb), dim3(kBlock), 0, s, d);
  }
  return hipGetLastError();
}
static hipError_t tangent(const Dev& d, int dir, hipStream_t s) {
  const unsigned nb = dir == 0 ? blocks_for(2LL * (d.NX + 1) * d.NY) : blocks_for(2LL * d.NX * (d.NY + 1));
  // with the fused launch the tangential reconstruction runs inside flux()
  if (d.fuse_tanflux) return hipSuccess;
  if (d.order == 7) {
    if (dir == 0) hipLaunchKernelGGL((k_tangent79<0, 7>), dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL((k_tangent79<1, 7>), dim3(nb), dim3(kBlock), 0, s, d);
  } else if (d.order == 9) {
    if (dir == 0) hipLaunchKernelGGL((k_tangent79<0, 9>), dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL((k_tangent79<1, 9>), dim3(nb), dim3(kBlock), 0, s, d);
  } else {
    if (dir == 0) hipLaunchKernelGGL(k_tangent<0>, dim3(nb), dim3(kBlock), 0, s, d);
    else hipLaunchKernelGGL(k_tangent<1>, dim3(nb), dim3(kBlock), 0, s, d);
  }
  return hipGetLastError();
}
static hipError_t flux(const Dev& d, int dir, int stage, double dt, hipStream_t s) {
  if (d.order == 5 && d.fuse_tanflux) {
    const long long nfaces = dir == 0 ? (long long)(d.NX + 1) * d.NY : (long long)d.NX * (d.NY + 1);
    const unsigned nb = (unsigned)((nfaces + kTanFlux - 1) / kTanFlux);
    if (d.mlp && d.smooth_flux && d.tcount) {
      // smooth faces, then the listed rest (at most nfaces of them)
      // (the listed faces follow in flux_list)
      if (dir == 0)
        hipLaunchKernelGGL(k_tanflux_cheap<0>, dim3((unsigned)cheap_tiles(d, 0, kCheapFaces<0>)),
                           dim3(2 * kCheapFaces<0>), 0, s, d, stage, dt);
      else
        hipLaunchKernelGGL(k_tanflux_cheap<1>, dim3((unsigned)cheap_tiles(d, 1, kCheapFaces<1>)),
                           dim3(2 * kCheapFaces<1>), 0, s, d, stage, dt);
    } else if (d.mlp) {
      if (dir == 0) hipLaunchKernelGGL((k_tanflux<0, true>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
      else hipLaunchKernelGGL((k_tanflux<1, true>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
    } else {
      if (dir == 0) hipLaunchKernelGGL((k_tanflux<0, false>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
      else hipLaunchKernelGGL((k_tanflux<1, false>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
    }
    return hipGetLastError();
  }
  if (d.fuse_tanflux) {
    const long long nfaces = dir == 0 ? (long long)(d.NX + 1) * d.NY : (long long)d.NX * (d.NY + 1);
    const unsigned nb = (unsigned)((nfaces + kTanFlux - 1) / kTanFlux);
    if (d.order == 7) {
      if (dir == 0) hipLaunchKernelGGL((k_tanflux79<0, 7>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
      else hipLaunchKernelGGL((k_tanflux79<1, 7>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
    } else {
      if (dir == 0) hipLaunchKernelGGL((k_tanflux79<0, 9>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
      else hipLaunchKernelGGL((k_tanflux79<1, 9>), dim3(nb), dim3(2 * kTanFlux), 0, s, d, stage, dt);
    }
    return hipGetLastError();
  }
  if (dir == 0)
    hipLaunchKernelGGL(k_flux<0>, dim3(blocks_for((long long)d.gp * (d.NX + 1) * d.NY)), dim3(kBlock), 0, s, d,
                       stage, dt);
  else
    hipLaunchKernelGGL(k_flux<1>, dim3(blocks_for((long long)d.gp * d.NX * (d.NY + 1))), dim3(kBlock), 0, s, d,
                       stage, dt);
  return hipGetLastError();
}
// dir 0 / 1: one direction's list; kBothDirs: both lists (one prefix
// launch; the listed cells in one launch)
static hipError_t normal_list(const Dev& d, int dir, hipStream_t s) {
  if (d.order != 5 || !d.mlp || !d.ncount) return hipSuccess;
  // a bound on each list's tiles: each shard adds at most one partial tile
  const unsigned nbx = kListShards + blocks_for((long long)(d.NX + 2) * d.ny);
  const unsigned nby = kListShards + blocks_for((long long)d.nx * (d.NY + 2));
  hipLaunchKernelGGL(k_list_prefix, dim3(1), dim3(64), 0, s, d, dir == 1 ? (int)KL_NORMAL_Y : (int)KL_NORMAL_X,
                     dir == kBothDirs ? 2 : 1, kBlock);
  if (dir == 0) hipLaunchKernelGGL(k_normal_list<0>, dim3(nbx), dim3(kBlock), 0, s, d);
  else if (dir == 1) hipLaunchKernelGGL(k_normal_list<1>, dim3(nby), dim3(kBlock), 0, s, d);
  else hipLaunchKernelGGL(k_normal_list_xy, dim3(nbx + nby), dim3(kBlock), 0, s, d, nbx);
  return hipGetLastError();
}
static hipError_t flux_list(const Dev& d, int dir, int stage, double dt, hipStream_t s) {
  if (d.order != 5 || !d.fuse_tanflux || !d.mlp || !d.smooth_flux || !d.tcount) return hipSuccess;
  const long long nfx = (long long)(d.NX + 1) * d.NY, nfy = (long long)d.NX * (d.NY + 1);
  const unsigned nbx = kListShards + (unsigned)((nfx + kListFaces - 1) / kListFaces);
  const unsigned nby = kListShards + (unsigned)((nfy + kListFaces - 1) / kListFaces);
  hipLaunchKernelGGL(k_list_prefix, dim3(1), dim3(64), 0, s, d, (int)KL_FLUX_X + (dir == 1 ? 1 : 0),
                     dir == kBothDirs ? 2 : 1, kListFaces);
  // (both: one prefix launch, then a launch per direction: one kernel
  // holding both directions' flux bodies exceeds two waves' registers)
  if (dir != 1) hipLaunchKernelGGL(k_tanflux_list<0>, dim3(nbx), dim3(2 * kListFaces), 0, s, d, stage, dt);
  if (dir != 0) hipLaunchKernelGGL(k_tanflux_list<1>, dim3(nby), dim3(2 * kListFaces), 0, s, d, stage, dt);
  return hipGetLastError();
}
static hipError_t update(const Dev& d, int stage, hipStream_t s) {
  (void)stage;  // the stage's combination is formed by k_flux
  hipLaunchKernelGGL(k_update, dim3(blocks_for((long long)d.NX * d.NY)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}
static hipError_t update_stat(const Dev& d, int slot, hipStream_t s) {
  const unsigned nb = blocks_for((long long)d.NX * d.NY);
  hipLaunchKernelGGL(k_update_stat, dim3(nb), dim3(kBlock), 0, s, d, d.partial);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kBlock), 0, s, (const double*)d.partial, (int)nb, d.stat, slot, 4);
  return hipGetLastError();
}
static hipError_t alpha(const Dev& d, hipStream_t s) {
  const int nj = d.NY < d.NX ? d.NY : d.NX;
  hipLaunchKernelGGL(k_alpha, dim3(blocks_for((long long)d.NX * nj)), dim3(kBlock), 0, s, d);
  return hipGetLastError();
}
static hipError_t flux_points(const Gas& g, long long n, const double* L, const double* R, double dt, double* C,
                              double* F, double* D, int smooth, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_points, dim3(blocks_for(n)), dim3(kBlock), 0, s, g, n, L, R, dt, C, F, D, smooth);
  return hipGetLastError();
}

static hipError_t linear_points(long long n, const double* W, double h, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_linear_points, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, W, h, out);
  return hipGetLastError();
}

static hipError_t label(const Dev& d, int dir, int stride, double tol, float* X, float* Y, hipStream_t s) {
  const long long total = dir == 0 ? (long long)(d.NX + 2) * d.ny : (long long)d.nx * (d.NY + 2);
  const long long ns = (total + stride - 1) / stride;
  if (dir == 0) hipLaunchKernelGGL(k_label<0>, dim3(blocks_for(ns)), dim3(kBlock), 0, s, d, stride, tol, X, Y);
  else hipLaunchKernelGGL(k_label<1>, dim3(blocks_for(ns)), dim3(kBlock), 0, s, d, stride, tol, X, Y);
  return hipGetLastError();
}
static hipError_t infer(const float* P, const unsigned short* W, long long n, const float* X, float* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks_for(n)), dim3(kBlock), 0, s, P, W, n, X, out);
  return hipGetLastError();
}

// One stage of the batched step.  As gks2d_run's step: stage 0 writes its
// state to the second buffer, stage 1 reconstructs that buffer (B.swap) and
// updates the state in place (no CopyFluid_new_to_old); each update forms
// the blow-up test and CFL speed of the state it writes (no primvar sweep):
// stage 0's into slot 1 (the next step's CFL speed) with its verdict taken
// at this step's end, stage 1's into slot 0 with its verdict taken at the
// next step's end (stat[5], kb_step_begin).
template <bool LEARNED, int TF>
static void batch_stage(const BatchArgs& B0, const unsigned* tiles, int stage, hipStream_t s) {
  BatchArgs B = B0;
  B.swap = stage;
  // the list launches' bound: each problem's shards add at most one partial
  // tile each
  const unsigned lb = (unsigned)kListShards * (unsigned)B.np;
  (void)lb;
  for (int side = 0; side < 4; side++)
    hipLaunchKernelGGL(kb_bc, dim3(tiles[side < 2 ? KB_BC_LR : KB_BC_DU]), dim3(kBlock), 0, s, B, side);
  if (LEARNED) {
    // both directions' classifiers (KB_NORMAL_X counts their tiles)
    hipLaunchKernelGGL(kb_classify, dim3(tiles[KB_NORMAL_X]), dim3(kBlock), 0, s, B);
    hipLaunchKernelGGL(kb_list_prefix, dim3(1), dim3(64), 0, s, B, (int)KL_NORMAL_X, kBlock);
    hipLaunchKernelGGL(kb_normal_list<0>, dim3(tiles[KB_NORMAL_X] + lb), dim3(kBlock), 0, s, B);
  } else {
    hipLaunchKernelGGL(kb_normal<0>, dim3(tiles[KB_NORMAL_X]), dim3(kBlock), 0, s, B);
    hipLaunchKernelGGL(kb_normal<1>, dim3(tiles[KB_NORMAL_Y]), dim3(kBlock), 0, s, B);
  }
  if (LEARNED) {
    hipLaunchKernelGGL(kb_list_prefix, dim3(1), dim3(64), 0, s, B, (int)KL_NORMAL_Y, kBlock);
    hipLaunchKernelGGL(kb_normal_list<1>, dim3(tiles[KB_NORMAL_Y] + lb), dim3(kBlock), 0, s, B);
  }
  if (LEARNED) {
    hipLaunchKernelGGL((kb_tanflux_cheap<0, TF>), dim3(tiles[KB_FLUX_X]), dim3(2 * TF), 0, s, B, stage);
    hipLaunchKernelGGL(kb_list_prefix, dim3(1), dim3(64), 0, s, B, (int)KL_FLUX_X, TF);
    hipLaunchKernelGGL((kb_tanflux_list<0, TF>), dim3(tiles[KB_FLUX_X] + lb), dim3(2 * TF), 0, s, B, stage);
    hipLaunchKernelGGL((kb_tanflux_cheap<1, TF>), dim3(tiles[KB_FLUX_Y]), dim3(2 * TF), 0, s, B, stage);
    hipLaunchKernelGGL(kb_list_prefix, dim3(1), dim3(64), 0, s, B, (int)KL_FLUX_Y, TF);
    hipLaunchKernelGGL((kb_tanflux_list<1, TF>), dim3(tiles[KB_FLUX_Y] + lb), dim3(2 * TF), 0, s, B, stage);
  } else {
    hipLaunchKernelGGL((kb_tanflux<0, LEARNED, TF>), dim3(tiles[KB_FLUX_X]), dim3(2 * TF), 0, s, B, stage);
    hipLaunchKernelGGL((kb_tanflux<1, LEARNED, TF>), dim3(tiles[KB_FLUX_Y]), dim3(2 * TF), 0, s, B, stage);
  }
  hipLaunchKernelGGL(kb_update_stat, dim3(tiles[KB_UPDATE]), dim3(kBlock), 0, s, B);
  hipLaunchKernelGGL(kb_reduce, dim3(B.np), dim3(kBlock), 0, s, B, 1 - stage, (int)KB_UPDATE, stage == 0 ? 4 : 5);
  hipLaunchKernelGGL(kb_alpha, dim3(tiles[KB_ALPHA]), dim3(kBlock), 0, s, B);
}
template <bool LEARNED, int TF>
static void batch_step_t(const BatchArgs& B, const unsigned* tiles, hipStream_t s) {
  hipLaunchKernelGGL(kb_step_begin, dim3(1), dim3(64), 0, s, B);
  batch_stage<LEARNED, TF>(B, tiles, 0, s);
  batch_stage<LEARNED, TF>(B, tiles, 1, s);
  hipLaunchKernelGGL(kb_step_end, dim3(1), dim3(64), 0, s, B);
}
// the batch path is WENO5-AO-DF with the fused tangential + flux launch and
// the DF sensor update, as gks2d_run's stage sequence for weno_order 5
static hipError_t batch_step(const BatchArgs& B, const BatchShape& sh, const unsigned* tiles, hipStream_t s) {
  if (sh.learned) {
    if (sh.tanflux_faces == 64) batch_step_t<true, 64>(B, tiles, s);
    else batch_step_t<true, 128>(B, tiles, s);
  } else {
    if (sh.tanflux_faces == 64) batch_step_t<false, 64>(B, tiles, s);
    else batch_step_t<false, 128>(B, tiles, s);
  }
  return hipGetLastError();
}

const LaunchTable table = {bc,     bc_all,  primvar, copy_old,    copy_ghosts, normal,     tangent,      flux,
                          normal_list, flux_list, update, update_stat, alpha, flux_points, label, infer,
                          batch_step, linear_points};
#endif

}  // namespace GKS_VARIANT
