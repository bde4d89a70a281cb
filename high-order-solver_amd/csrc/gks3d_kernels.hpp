// gks3d_kernels.hpp — the kernels of the 3D GKS path (one S2O4 step of the
// reference's CubicTube loop, 3Dproblem.cpp:109-166, with lambda from the
// state and periodic boundaries added).
//
// Per stage: one boundary launch, the primitive-variable sweep + reduction
// (the CFL inputs and the blow-up flag), then per direction d = x, y, z
//   face   one thread per cell: WENO5-AO across the d faces
//          (WENO5_AO_splitting_3d, reconstruction.cpp:3774-3923) -> fv
//   line   one thread per (d face, side): the first tangential
//          reconstruction (WENO5_AO_of_line_value, :3925-4145) -> lv
//   flux   per d face that gets a flux, in one workgroup's LDS tile: the
//          second tangential reconstruction (WENO5_AO_of_point_value,
//          :4147-4399; one lane per face, line and side), then GKS3D
//          (flux_function.cpp:813-1163; one lane per face and Gauss point),
//          the S2O4 stage combination summed over the face's 4 points ->
//          xs[d] (and the stage-0 share pc[d])
// and finally the cell update (Update_gauss, time_advance.cpp:937-998).
// The face and line buffers are reused by the three directions.
//
// Every thread's index box is the reference's loop range for that array;
// threads are laid out with k (z) fastest so neighbouring lanes touch
// neighbouring cells / faces in every SoA component.
#pragma once
#include <hip/hip_runtime.h>

#include "gks3d_device.hpp"
#include "gks3d_launch.hpp"
// the troubled-interface classifier of the 2D path (stencil features, the
// MLP on the matrix cores) and its linear-weight reconstructions
#include "gks2d_learn_device.hpp"

#ifndef GKS_VARIANT
#error "GKS_VARIANT must name the kernel namespace"
#endif

namespace GKS_VARIANT {

using namespace gksd3;

constexpr int kBlock3 = 256;

// a box [lo, hi) of cell or face indices, k fastest
struct Box3 {
  int lo[3], n[3];
  __device__ __forceinline__ long long count() const { return (long long)n[0] * n[1] * n[2]; }
  __device__ __forceinline__ void at(long long t, int& i, int& j, int& k) const {
    k = lo[2] + (int)(t % n[2]);
    t /= n[2];
    j = lo[1] + (int)(t % n[1]);
    i = lo[0] + (int)(t / n[1]);
  }
};
__host__ __device__ inline Box3 make_box(int i0, int i1, int j0, int j1, int k0, int k1) {
  Box3 b;
  b.lo[0] = i0;
  b.lo[1] = j0;
  b.lo[2] = k0;
  b.n[0] = i1 > i0 ? i1 - i0 : 0;
  b.n[1] = j1 > j0 ? j1 - j0 : 0;
  b.n[2] = k1 > k0 ? k1 - k0 : 0;
  return b;
}

__device__ __forceinline__ long long cell3(const Dev3& d, int i, int j, int k) {
  return ((long long)i * d.ny + j) * d.nz + k;
}
__device__ __forceinline__ long long face3(const Dev3& d, int i, int j, int k) {
  return ((long long)i * (d.ny + 1) + j) * (d.nz + 1) + k;
}

// Workgroups go to the eight XCDs round robin, each with its own L2; the
// tile of workgroup b is the (b % 8)-th contiguous eighth's (b / 8)-th, so
// that tiles next to each other in the sweep's order run on one XCD at
// about the same time (a stencil's neighbouring rows then come through one
// L2).
__device__ __forceinline__ unsigned xcd_tile() {
  const unsigned nb = gridDim.x, b = blockIdx.x, q = nb / 8, r = nb % 8, x = b % 8;
  return x * q + (x < r ? x : r) + b / 8;
}

// ------------------------------------------------------------ boundaries
// free_boundary_* (boundary_condition.cpp:445-570) in the reference's side
// order x, y, z compose into one map per ghost cell: each coordinate outside
// the interior is clamped to the nearest interior layer (free) or wrapped
// by the interior size (periodic); sources are interior cells only, so one
// pass over the ghost cells gives the sequential result.
__global__ void __launch_bounds__(kBlock3) k3_bc(Dev3 d, int b0, int b1, int b2, int b3, int b4, int b5) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.nc) return;
  const int k = (int)(t % d.nz);
  const int j = (int)((t / d.nz) % d.ny);
  const int i = (int)(t / ((long long)d.nz * d.ny));
  const int G = d.G;
  auto map = [G](int x, int N, int lo_kind, int hi_kind) {
    if (x < G) return lo_kind == 0 ? G : x + N;
    if (x >= N + G) return hi_kind == 0 ? N + G - 1 : x - N;
    return x;
  };
  const int si = map(i, d.NX, b0, b1), sj = map(j, d.NY, b2, b3), sk = map(k, d.NZ, b4, b5);
  if (si == i && sj == j && sk == k) return;
  const long long src = cell3(d, si, sj, sk);
#pragma unroll
  for (int v = 0; v < 5; v++) d.U[v * d.nc + t] = d.U[v * d.nc + src];
}

// one side's fill, as the reference's boundary functions apply it (the ghost
// slab of that side over the full extent of the other two axes); sides in
// the reference's order x-, x+, y-, y+, z-, z+
__global__ void __launch_bounds__(kBlock3) k3_bc_side(Dev3 d, Box3 b, int side, int kind) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.count()) return;
  int i, j, k;
  b.at(t, i, j, k);
  const int G = d.G, axis = side >> 1;
  const bool lo = (side & 1) == 0;
  const int N = axis == 0 ? d.NX : (axis == 1 ? d.NY : d.NZ);
  const int x = axis == 0 ? i : (axis == 1 ? j : k);
  const int sx = kind == 0 ? (lo ? G : N + G - 1) : (lo ? x + N : x - N);
  const long long dst = cell3(d, i, j, k);
  const long long src = cell3(d, axis == 0 ? sx : i, axis == 1 ? sx : j, axis == 2 ? sx : k);
#pragma unroll
  for (int v = 0; v < 5; v++) d.U[v * d.nc + dst] = d.U[v * d.nc + src];
}

// ------------------------------------------------------- primitive sweep
// Convar_to_primvar over the interior (basic_function.cpp:775-804) reduced
// to what the next step's Get_CFL / Dtx_3d (time_advance.cpp:793-848) and
// the blow-up check need: the largest |u| + c, the smallest density, and
// whether any density or pressure is negative or NaN.
__global__ void __launch_bounds__(kBlock3) k3_primvar(Dev3 d) {
  __shared__ double s_max[kBlock3], s_min[kBlock3], s_bad[kBlock3];
  const long long n = (long long)d.NX * d.NY * d.NZ;
  double vmax = 0.0, rmin = 1e300, bad = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int k = d.G + (int)(t % d.NZ);
    const int j = d.G + (int)((t / d.NZ) % d.NY);
    const int i = d.G + (int)(t / ((long long)d.NZ * d.NY));
    const long long c = cell3(d, i, j, k);
    double u[5], p[5];
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = d.U[v * d.nc + c];
    prim3(p, u, d.gas);
    if (isnan(p[0]) || isnan(p[4]) || p[0] < 0 || p[4] < 0) bad = 1.0;
    const double overden = 1.0 / p[0];
    const double s = sqrt(p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) + sqrt(d.gas.gamma * p[4] * overden);
    if (s > vmax) vmax = s;
    if (p[0] < rmin) rmin = p[0];
  }
  s_max[threadIdx.x] = vmax;
  s_min[threadIdx.x] = rmin;
  s_bad[threadIdx.x] = bad;
  __syncthreads();
  for (int w = kBlock3 / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
      s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + w]);
      s_bad[threadIdx.x] = fmax(s_bad[threadIdx.x], s_bad[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.part[blockIdx.x] = s_max[0];
    d.part[kParts3 + blockIdx.x] = s_min[0];
    d.part[2 * kParts3 + blockIdx.x] = s_bad[0];
  }
}
// one workgroup of kParts3 lanes: the partials -> stat (the blow-up flag
// only ever rises; the host clears it)
__global__ void __launch_bounds__(kParts3) k3_reduce(Dev3 d, int nparts) {
  __shared__ double s_max[kParts3], s_min[kParts3], s_bad[kParts3];
  const int t = threadIdx.x;
  s_max[t] = t < nparts ? d.part[t] : 0.0;
  s_min[t] = t < nparts ? d.part[kParts3 + t] : 1e300;
  s_bad[t] = t < nparts ? d.part[2 * kParts3 + t] : 0.0;
  __syncthreads();
  for (int w = kParts3 / 2; w > 0; w >>= 1) {
    if (t < w) {
      s_max[t] = fmax(s_max[t], s_max[t + w]);
      s_min[t] = fmin(s_min[t], s_min[t + w]);
      s_bad[t] = fmax(s_bad[t], s_bad[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    d.stat[0] = s_max[0];
    d.stat[1] = s_min[0];
    d.stat[2] = fmax(d.stat[2], s_bad[0]);
  }
}

// CopyFluid_new_to_old (basic_function.cpp:1586-1603): interior cells
__global__ void __launch_bounds__(kBlock3) k3_copy_old(Dev3 d, Box3 b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.count()) return;
  int i, j, k;
  b.at(t, i, j, k);
  const long long c = cell3(d, i, j, k);
#pragma unroll
  for (int v = 0; v < 5; v++) d.Uold[v * d.nc + c] = d.U[v * d.nc + c];
}

// ------------------------------------------------------ learned classifier
// linear / all counts of one wave's reconstructions of a kind (0 face, 1
// line, 2 point), added by its first active lane to one of kStat3Slots
// slot pairs
__device__ __forceinline__ void count3(unsigned long long* stat, int kind, bool lin) {
  const unsigned long long act = __ballot(1), bl = __ballot(lin);
  if ((threadIdx.x & 63) == __ffsll((long long)act) - 1) {
    unsigned long long* slot =
        stat + 2 * (kind * kStat3Slots + (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kStat3Slots));
    atomicAdd(slot, (unsigned long long)__popcll(bl));
    atomicAdd(slot + 1, (unsigned long long)__popcll(act));
  }
}
// Every cell of the box, along x, y and z: the 16 features of the 5-point
// (rho, p) stencil and the MLP (one wave evaluates its 64 cells per
// direction on the matrix cores); lin3[dir][cell] = 1 when the logit is below
// the threshold (the 3D scheme has no DF sensor).  Along dir only cells
// whose dir index is in [G-1, n-G+1) are classified (their stencils stay in
// the block; every read of lin3[dir] is at such a cell).  All lanes of a
// wave run the MLP; the other lanes carry zero features and store nothing.
__global__ void __launch_bounds__(kBlock3) k3_classify(Dev3 d, Box3 b) {
  __shared__ __align__(16) short xs[kBlock3 * 16];
  __shared__ float ls[kBlock3];
  __shared__ gksd::MlpLds s_mlp;
  gksd::mlp_stage(s_mlp, d.mlp, d.mlpw);  // loads while the stencils do
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < b.count();
  int i, j, k;
  b.at(valid ? t : 0, i, j, k);
  const long long c = cell3(d, i, j, k);
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int dir = 0; dir < 3; dir++) {
    const long long cs = dir == 0 ? (long long)d.ny * d.nz : (dir == 1 ? (long long)d.nz : 1LL);
    const int cd = dir == 0 ? i : (dir == 1 ? j : k), nd = dir == 0 ? d.nx : (dir == 1 ? d.ny : d.nz);
    const bool ok = valid && cd >= d.G - 1 && cd < nd - d.G + 1;
    float x[16];
    if (ok) {
      double q[2][5];
#pragma unroll
      for (int s = 0; s < 5; s++) {
        double u[5], p[5];
#pragma unroll
        for (int v = 0; v < 5; v++) u[v] = d.U[v * d.nc + c + (s - 2) * cs];
        prim3(p, u, d.gas);
        q[0][s] = p[0];
        q[1][s] = p[4];
      }
      gksd::stencil_features(x, q);
    } else {
#pragma unroll
      for (int m = 0; m < 16; m++) x[m] = 0.0f;
    }
    if (dir == 0) __syncthreads();
    const float z = gksd::mlp_wave_logit(x, gksd::MlpShared{&s_mlp}, xs + wv * 64 * 16, ls + wv * 64);
    if (ok) d.lin3[dir * d.nc + c] = z < d.mlp_thr ? 1 : 0;
  }
}
// the five-entry tangential rule: entry q's source cell (the left cell of
// its face for side 0, the right one for side 1) took the linear weights
// across the face, and the centre entry's source cell took them along the
// tangential direction TDIR
template <int DIR, int TDIR>
__device__ __forceinline__ bool line_sources_linear(const Dev3& d, int i, int j, int k, int side, int sign) {
  const int si = i - (side == 0 && DIR == 0), sj = j - (side == 0 && DIR == 1), sk = k - (side == 0 && DIR == 2);
  unsigned ok = 1;
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int o = sign * (q - 2);
    ok &= d.lin3[DIR * d.nc + cell3(d, si + (TDIR == 0) * o, sj + (TDIR == 1) * o, sk + (TDIR == 2) * o)];
  }
  return ok && d.lin3[TDIR * d.nc + cell3(d, si, sj, sk)];
}
// tangential_side with the linear weights (the fifth-order candidate at the
// two Gauss points), on the conservative values: the characteristic
// projection commutes with a fixed linear combination
template <int NC, class ST>
__device__ __forceinline__ void tangential_side_linear(double (*out)[NC + 5], const ST& st, const HInv& h) {
  constexpr int NF = NC / 5;
#pragma unroll
  for (int fld = 0; fld < NF; fld++)
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const int m = 5 * fld + v;
      double g1, g1x, g2, g2x;
      gksd::weno5_linear_2gauss(g1, g1x, g2, g2x, st(0, m), st(1, m), st(2, m), st(3, m), st(4, m), h.h, h.r1);
      out[0][m] = g1;
      out[1][m] = g2;
      if (fld == 0) {
        out[0][NC + v] = g1x;
        out[1][NC + v] = g2x;
      }
    }
}

// ------------------------------------------------------------ face values
// WENO5_AO (reconstruction.cpp:3803-3923) of one cell across its two DIR
// faces, in the face frame (Ydirection / Zdirection for y / z).  LEARNED: a
// cell whose DIR stencil the classifier called smooth takes the linear
// weights (the fifth-order candidate, on the conservative values).
template <int DIR, bool LEARNED>
__global__ void __launch_bounds__(kBlock3) k3_face(Dev3 d, Box3 b) {
  const long long t = (long long)xcd_tile() * blockDim.x + threadIdx.x;
  if (t >= b.count()) return;
  int i, j, k;
  b.at(t, i, j, k);
  const Gas3& g = d.gas;
  const long long c = cell3(d, i, j, k);
  const long long cs = DIR == 0 ? (long long)d.ny * d.nz : (DIR == 1 ? (long long)d.nz : 1LL);
  const HInv h = hinv(DIR == 0 ? d.dx : (DIR == 1 ? d.dy : d.dz));
  double w[5][5];
#pragma unroll
  for (int s = 0; s < 5; s++) {
    double u[5];
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = d.U[v * d.nc + c + (s - 2) * cs];
    to_frame(w[s], u, DIR);
  }
  double lo[10], ro[10];
  bool lin = false;
  if constexpr (LEARNED) {
    lin = d.lin3[DIR * d.nc + c] != 0;
    count3(d.lstat3, 0, lin);
  }
  if (lin) {
#pragma unroll
    for (int v = 0; v < 5; v++) {
      gksd::weno5_linear_face(0, lo[v], lo[5 + v], w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], h.h, h.r1);
      gksd::weno5_linear_face(1, ro[v], ro[5 + v], w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], h.h, h.r1);
    }
  } else {
    double pn1[5], p0[5], pp1[5];
    prim3(pn1, w[1], g);
    prim3(p0, w[2], g);
    prim3(pp1, w[3], g);
#pragma unroll
    for (int side = 0; side < 2; side++) {
      double base[5];
#pragma unroll
      for (int v = 0; v < 5; v++) base[v] = 0.5 * ((side == 0 ? pn1[v] : pp1[v]) + p0[v]);
      double r[5][5], var[5], der1[5];
      if (d.recon_var == 0) {
#pragma unroll
        for (int s = 0; s < 5; s++)
#pragma unroll
          for (int v = 0; v < 5; v++) r[s][v] = w[s][v];
      } else {
        double sm[25];
        char_base3(sm, base, g);
#pragma unroll
        for (int s = 0; s < 5; s++) to_char3(r[s], sm, w[s]);
      }
#pragma unroll
      for (int v = 0; v < 5; v++) {
        if (side == 0) weno5_left3(var[v], der1[v], r[0][v], r[1][v], r[2][v], r[3][v], r[4][v], h);
        else weno5_right3(var[v], der1[v], r[0][v], r[1][v], r[2][v], r[3][v], r[4][v], h);
      }
      double* out = side == 0 ? lo : ro;
      if (d.recon_var == 0) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          out[v] = var[v];
          out[5 + v] = der1[v];
        }
      } else {
        const RightEig e = right_eig(base, g);
        from_char3(out, e, var);
        from_char3(out + 5, e, der1);
      }
    }
  }
  if (reduce3(ro, g) || reduce3(lo, g)) {
#pragma unroll
    for (int v = 0; v < 5; v++) {
      lo[v] = w[2][v];
      ro[v] = w[2][v];
      lo[5 + v] = 0.0;
      ro[5 + v] = 0.0;
    }
  }
  // the cell's left face gets its right-side state, the right face its left
  const long long fa = face3(d, i, j, k);
  const long long fb = face3(d, i + (DIR == 0), j + (DIR == 1), k + (DIR == 2));
#pragma unroll
  for (int m = 0; m < 10; m++) {
    d.fv[(10 + m) * d.nf + fa] = lo[m];
    d.fv[m * d.nf + fb] = ro[m];
  }
}

// ------------------------------------------------ tangential reconstruction
// One side of WENO5_AO_for_line_value (NC = 10: convar, der1x -> + der1y;
// reconstruction.cpp:3955-4145) or WENO5_AO_for_point_value (NC = 15:
// convar, der1x, der1y -> + der1z; :4163-4399).  st(q, m) reads stencil
// entry q (0..4) component m; out[p][NC + 5]: per Gauss point p the
// reconstructed components, then the new tangential derivative.  The first
// field (convar) takes the derivative, the carried derivatives only values;
// in characteristic variables every field is projected on the base of the
// centre entry's primitive variables.
template <int NC, class ST>
__device__ __forceinline__ void tangential_side(double (*out)[NC + 5], const ST& st, const HInv& h, int recon_var,
                                                const Gas3& g) {
  constexpr int NF = NC / 5;
  double c2[5], base[5];
#pragma unroll
  for (int v = 0; v < 5; v++) c2[v] = st(2, v);
  prim3(base, c2, g);
  if (recon_var == 0) {
#pragma unroll
    for (int fld = 0; fld < NF; fld++)
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const int m = 5 * fld + v;
        double d0, d1;
        if (fld == 0)
          weno5_2gauss3<true>(out[0][m], d0, out[1][m], d1, st(0, m), st(1, m), st(2, m), st(3, m), st(4, m), h);
        else
          weno5_2gauss3<false>(out[0][m], d0, out[1][m], d1, st(0, m), st(1, m), st(2, m), st(3, m), st(4, m), h);
        if (fld == 0) {
          out[0][NC + v] = d0;
          out[1][NC + v] = d1;
        }
      }
    return;
  }
  double sm[25];
  char_base3(sm, base, g);
  const RightEig e = right_eig(base, g);
#pragma unroll
  for (int fld = 0; fld < NF; fld++) {
    double r[5][5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      double c[5];
#pragma unroll
      for (int v = 0; v < 5; v++) c[v] = st(q, 5 * fld + v);
      to_char3(r[q], sm, c);
    }
    double var[2][5], der[2][5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      if (fld == 0)
        weno5_2gauss3<true>(var[0][v], der[0][v], var[1][v], der[1][v], r[0][v], r[1][v], r[2][v], r[3][v], r[4][v],
                            h);
      else
        weno5_2gauss3<false>(var[0][v], der[0][v], var[1][v], der[1][v], r[0][v], r[1][v], r[2][v], r[3][v],
                             r[4][v], h);
    }
#pragma unroll
    for (int p = 0; p < 2; p++) {
      from_char3(out[p] + 5 * fld, e, var[p]);
      if (fld == 0) from_char3(out[p] + NC, e, der[p]);
    }
  }
}

// stencil readers: entry q at face f0 + (q - 2) * step (reversed stencils
// pass a negative step), component m of side `side` in an SoA buffer
struct FvStencil {
  const double* p;
  long long nf, f0, step;
  int side;
  __device__ __forceinline__ double operator()(int q, int m) const {
    return p[(side * 10 + m) * nf + f0 + (q - 2) * step];
  }
};
struct LvStencil {
  const double* p;
  long long nf, f0, step;
  int line, side;
  __device__ __forceinline__ double operator()(int q, int m) const {
    return p[((line * 2 + side) * 15 + m) * nf + f0 + (q - 2) * step];
  }
};

// tiles of k3_line<DIR>, kLineFaces faces each: x and z faces, whose line
// stencil runs along j (nz + 1 faces apart), consecutive faces of the box
// (k fastest), consecutive tiles on one XCD so that a stencil's rows come
// through one L2; y faces, whose stencil runs along i ((ny + 1)(nz + 1)
// faces apart), kLineFaces consecutive faces of one (j, k) plane with i
// fastest from tile to tile
constexpr int kLineFaces = kBlock3 / 2;
__host__ __device__ inline long long line_tiles(const Box3& b, int dir) {
  const long long plane = (long long)b.n[1] * b.n[2];
  return dir == 1 ? (long long)b.n[0] * ((plane + kLineFaces - 1) / kLineFaces)
                  : ((long long)b.n[0] * plane + kLineFaces - 1) / kLineFaces;
}

// WENO5_AO_of_line_value (reconstruction.cpp:3925-3953): x faces along y
// (h = dy), y faces along -x (h = dx, the stencil reversed), z faces along
// y (h = dy).  One thread per (face, side); the two sides of a face are
// neighbouring lanes, which share their order-reduction verdict: both lines
// of both sides fall back to the face value (tangential derivative 0) when
// any of the four reconstructed states has lambda <= 0.
template <int DIR, bool LEARNED>
__global__ void __launch_bounds__(kBlock3) k3_line(Dev3 d, Box3 b) {
  const int fl = threadIdx.x >> 1, side = threadIdx.x & 1;
  const unsigned tile = xcd_tile();
  long long fi;  // the face's index in the box (k fastest)
  if (DIR == 1) {
    const long long plane = (long long)b.n[1] * b.n[2];
    const unsigned ti = tile % (unsigned)b.n[0], tp = tile / (unsigned)b.n[0];
    const long long p = (long long)tp * kLineFaces + fl;
    if (p >= plane) return;  // both lanes of a pair leave together
    fi = (long long)ti * plane + p;
  } else {
    fi = (long long)tile * kLineFaces + fl;
    if (fi >= b.count()) return;
  }
  int i, j, k;
  b.at(fi, i, j, k);
  const long long f = face3(d, i, j, k);
  const long long step = DIR == 1 ? -(long long)(d.ny + 1) * (d.nz + 1) : (long long)(d.nz + 1);
  const HInv h = hinv(DIR == 1 ? d.dx : d.dy);
  double res[2][15];  // [line]
  bool lin = false;
  if constexpr (LEARNED) {
    // line direction: y for x and z faces, -x for y faces
    lin = DIR == 1 ? line_sources_linear<1, 0>(d, i, j, k, side, -1) : line_sources_linear<DIR, 1>(d, i, j, k, side, 1);
    count3(d.lstat3, 1, lin);
    d.lin_line[(long long)side * d.nf + f] = lin ? 1 : 0;
  }
  if (lin) tangential_side_linear<10>(res, FvStencil{d.fv, d.nf, f, step, side}, h);
  else tangential_side<10>(res, FvStencil{d.fv, d.nf, f, step, side}, h, d.recon_var, d.gas);
  int red = (reduce3(res[0], d.gas) || reduce3(res[1], d.gas)) ? 1 : 0;
  red |= __shfl_xor(red, 1);
  if (red) {
#pragma unroll
    for (int l = 0; l < 2; l++) {
#pragma unroll
      for (int m = 0; m < 10; m++) res[l][m] = d.fv[(side * 10 + m) * d.nf + f];
#pragma unroll
      for (int m = 10; m < 15; m++) res[l][m] = 0.0;
    }
  }
#pragma unroll
  for (int l = 0; l < 2; l++)
#pragma unroll
    for (int m = 0; m < 15; m++) d.lv[((l * 2 + side) * 15 + m) * d.nf + f] = res[l][m];
}

// WENO5_AO_of_point_value (reconstruction.cpp:4147-4161) and GKS3D at the
// faces that get a flux (Calculate_flux, flux_function.cpp:760-811), fused
// through LDS: a workgroup of kPfBlock lanes takes kPfFaces faces.
// Phase 1, one lane per (face, line, side): the second tangential
// reconstruction — x and y faces along z (h = dz), z faces along -x
// (h = dx, reversed) — into the workgroup's LDS tile; the two sides are
// neighbouring lanes sharing the order-reduction verdict (fallback: the line
// value, der1z 0).  Phase 2, one lane per (face, Gauss point): GKS3D
// (flux_function.cpp:813-1163) reading its left and right states from the
// tile, and the S2O4 stage combination of Update_gauss
// (time_advance.cpp:867-935) — stage 0 x = 0.5 f + 1/8 derf (keeping
// 1.0 f + 1/6 derf for stage 1), stage 1 x = that + 1/3 derf — summed over
// the face's four points by lane shuffles.  The tile (kPfFaces x 2 sides x 4
// points x 20 values = 40 KB) never leaves the CU; four workgroups fill
// the 160 KB of a CU's LDS.
constexpr int kPfBlock = 128;
constexpr int kPfFaces = kPfBlock / 4;
constexpr int kPfRow = 20;
struct LdsSide {
  const double* p;
  __device__ __forceinline__ double operator[](int k) const { return p[k]; }
};
// tiles of k3_pflux<DIR>: x and y faces, kPfFaces consecutive faces of the
// box (k fastest: the point stencil runs along k, inside the tile); z faces,
// whose point stencil runs along i, (ny + 1)(nz + 1) faces apart, kPfFaces
// consecutive faces of one (j, k) plane with i fastest from tile to tile, so
// that the five tiles of a stencil read their line values through one L2
__host__ __device__ inline long long pflux_tiles(const Box3& b, int dir) {
  const long long plane = (long long)b.n[1] * b.n[2];
  return dir == 2 ? (long long)b.n[0] * ((plane + kPfFaces - 1) / kPfFaces)
                  : ((long long)b.n[0] * plane + kPfFaces - 1) / kPfFaces;
}
template <int DIR, bool LEARNED>
__global__ void __launch_bounds__(kPfBlock) k3_pflux(Dev3 d, Box3 b, int stage, double dt) {
  __shared__ double s_pt[kPfFaces][2][4][kPfRow];
  const int lane = threadIdx.x;
  const int fl = lane >> 2;
  long long fi;  // the face's index in the box (k fastest)
  bool active;   // the four lanes of a face agree
  if (DIR == 2) {
    const long long plane = (long long)b.n[1] * b.n[2];
    const unsigned t = xcd_tile(), ti = t % (unsigned)b.n[0], tp = t / (unsigned)b.n[0];
    const long long p = (long long)tp * kPfFaces + fl;
    active = p < plane;
    fi = (long long)ti * plane + p;
  } else {
    fi = (long long)blockIdx.x * kPfFaces + fl;
    active = fi < b.count();
  }
  int i = 0, j = 0, k = 0;
  if (active) b.at(fi, i, j, k);
  const long long f = active ? face3(d, i, j, k) : 0;
  if (active) {
    const int l = (lane >> 1) & 1, side = lane & 1;
    const long long step = DIR == 2 ? -(long long)(d.ny + 1) * (d.nz + 1) : 1LL;
    const HInv h = hinv(DIR == 2 ? d.dx : d.dz);
    double res[2][20];  // [point]
    bool lin = false;
    if constexpr (LEARNED) {
      // point direction: z for x and y faces, -x for z faces; the five line
      // values took the linear weights and the centre source cell's stencil
      // along the point direction did
      const int si = i - (side == 0 && DIR == 0), sj = j - (side == 0 && DIR == 1), sk = k - (side == 0 && DIR == 2);
      unsigned ok = d.lin3[(DIR == 2 ? 0 : 2) * d.nc + cell3(d, si, sj, sk)];
#pragma unroll
      for (int q = 0; q < 5; q++) ok &= d.lin_line[(long long)side * d.nf + f + (q - 2) * step];
      lin = ok != 0;
    }
    if constexpr (LEARNED) count3(d.lstat3, 2, lin);
    if (lin) tangential_side_linear<15>(res, LvStencil{d.lv, d.nf, f, step, l, side}, h);
    else tangential_side<15>(res, LvStencil{d.lv, d.nf, f, step, l, side}, h, d.recon_var, d.gas);
    int red = (reduce3(res[0], d.gas) || reduce3(res[1], d.gas)) ? 1 : 0;
    red |= __shfl_xor(red, 1);
    if (red) {
#pragma unroll
      for (int p = 0; p < 2; p++) {
#pragma unroll
        for (int m = 0; m < 15; m++) res[p][m] = d.lv[((l * 2 + side) * 15 + m) * d.nf + f];
#pragma unroll
        for (int m = 15; m < 20; m++) res[p][m] = 0.0;
      }
    }
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int m = 0; m < 20; m++) s_pt[fl][side][2 * l + p][m] = res[p][m];
  }
  __syncthreads();
  const int q = lane & 3;
  double x[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, share[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (active) {
    double fo[5], dfo[5];
    gks3d_point_dev(fo, dfo, nullptr, LdsSide{&s_pt[fl][0][q][0]}, LdsSide{&s_pt[fl][1][q][0]}, dt, d.gas);
#pragma unroll
    for (int v = 0; v < 5; v++) {
      if (stage == 0) {
        x[v] = 0.0 + (0.5 * fo[v] + 1.0 / 8.0 * dfo[v]);
        share[v] = 0.0 + (1.0 * fo[v] + 1.0 / 6.0 * dfo[v]);
      } else {
        x[v] = 0.0 * fo[v] + 1.0 / 3.0 * dfo[v];
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 5; v++) {
    x[v] += __shfl_xor(x[v], 1);
    x[v] += __shfl_xor(x[v], 2);
    share[v] += __shfl_xor(share[v], 1);
    share[v] += __shfl_xor(share[v], 2);
  }
  if (active && q == 0) {
    double* xs = d.xs[DIR];
    double* pc = d.pc[DIR];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const long long o = (long long)v * d.nf + f;
      if (stage == 0) {
        xs[o] = x[v];
        pc[o] = share[v];
      } else {
        xs[o] = pc[o] + x[v];
      }
    }
  }
}

// the cell update of Update_gauss (time_advance.cpp:937-998): U = Uold +
// the Gauss-point sums (weights 1/4) of the faces, the y / z face frames
// rotated back
__global__ void __launch_bounds__(kBlock3) k3_update(Dev3 d, Box3 b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.count()) return;
  int i, j, k;
  b.at(t, i, j, k);
  const long long c = cell3(d, i, j, k), fc = face3(d, i, j, k);
  const long long fx = face3(d, i + 1, j, k), fy = face3(d, i, j + 1, k), fz = face3(d, i, j, k + 1);
  const double odx = 1 / d.dx, ody = 1 / d.dy, odz = 1 / d.dz;
  const double w = 0.25;
  auto X = [&](int dir, long long f, int v) { return d.xs[dir][(long long)v * d.nf + f]; };
  double tot[5];
  tot[0] = w * odz * (X(2, fc, 0) - X(2, fz, 0));
  tot[1] = w * odz * (-X(2, fc, 3) + X(2, fz, 3));
  tot[2] = w * odz * (X(2, fc, 2) - X(2, fz, 2));
  tot[3] = w * odz * (X(2, fc, 1) - X(2, fz, 1));
  tot[4] = w * odz * (X(2, fc, 4) - X(2, fz, 4));
  tot[0] += w * ody * (X(1, fc, 0) - X(1, fy, 0));
  tot[1] += w * ody * (-X(1, fc, 2) + X(1, fy, 2));
  tot[2] += w * ody * (X(1, fc, 1) - X(1, fy, 1));
  tot[3] += w * ody * (X(1, fc, 3) - X(1, fy, 3));
  tot[4] += w * ody * (X(1, fc, 4) - X(1, fy, 4));
#pragma unroll
  for (int v = 0; v < 5; v++) tot[v] += w * odx * (X(0, fc, v) - X(0, fx, v));
#pragma unroll
  for (int v = 0; v < 5; v++) d.U[v * d.nc + c] = d.Uold[v * d.nc + c] + tot[v];
}

// leaf parity: GKS3D at n Gauss points (left / right n x 20, row major)
struct RowSide {
  const double* p;
  __device__ __forceinline__ double operator[](int k) const { return p[k]; }
};
__global__ void __launch_bounds__(kBlock3) k3_flux_points(Gas3 g, long long n, const double* L, const double* R,
                                                          double dt, double* C, double* F, double* D) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double f[5], df[5], c[20];
  gks3d_point_dev(f, df, C ? c : nullptr, RowSide{L + t * 20}, RowSide{R + t * 20}, dt, g);
#pragma unroll
  for (int v = 0; v < 5; v++) {
    F[t * 5 + v] = f[v];
    D[t * 5 + v] = df[v];
  }
  if (C)
    for (int v = 0; v < 20; v++) C[t * 20 + v] = c[v];
}

// the device pass sees no launcher, so the direction templates are
// instantiated here explicitly
#define GKS3_INST(L)                                             \
  template __global__ void k3_face<0, L>(Dev3, Box3);            \
  template __global__ void k3_face<1, L>(Dev3, Box3);            \
  template __global__ void k3_face<2, L>(Dev3, Box3);            \
  template __global__ void k3_line<0, L>(Dev3, Box3);            \
  template __global__ void k3_line<1, L>(Dev3, Box3);            \
  template __global__ void k3_line<2, L>(Dev3, Box3);            \
  template __global__ void k3_pflux<0, L>(Dev3, Box3, int, double); \
  template __global__ void k3_pflux<1, L>(Dev3, Box3, int, double); \
  template __global__ void k3_pflux<2, L>(Dev3, Box3, int, double);
GKS3_INST(false)
GKS3_INST(true)
#undef GKS3_INST

// ------------------------------------------------------------ launchers
#if !defined(__HIP_DEVICE_COMPILE__)
static inline unsigned blocks_for(long long n) { return (unsigned)((n + kBlock3 - 1) / kBlock3); }
static inline long long box_count(const Box3& b) { return (long long)b.n[0] * b.n[1] * b.n[2]; }

static hipError_t launch_bc(const Dev3& d, const int* bc, hipStream_t s) {
  k3_bc<<<blocks_for(d.nc), kBlock3, 0, s>>>(d, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5]);
  return hipGetLastError();
}
static hipError_t launch_bc_side(const Dev3& d, int side, int kind, hipStream_t s) {
  const int G = d.G, axis = side >> 1;
  const bool lo = (side & 1) == 0;
  const int n = axis == 0 ? d.nx : (axis == 1 ? d.ny : d.nz);
  const int a0 = lo ? 0 : n - G, a1 = lo ? G : n;
  const Box3 b = make_box(axis == 0 ? a0 : 0, axis == 0 ? a1 : d.nx, axis == 1 ? a0 : 0, axis == 1 ? a1 : d.ny,
                          axis == 2 ? a0 : 0, axis == 2 ? a1 : d.nz);
  if (box_count(b) > 0) k3_bc_side<<<blocks_for(box_count(b)), kBlock3, 0, s>>>(d, b, side, kind);
  return hipGetLastError();
}
static hipError_t launch_primvar(const Dev3& d, hipStream_t s) {
  const long long n = (long long)d.NX * d.NY * d.NZ;
  long long nb = (n + kBlock3 - 1) / kBlock3;
  if (nb > kParts3) nb = kParts3;
  k3_primvar<<<(unsigned)nb, kBlock3, 0, s>>>(d);
  k3_reduce<<<1, kParts3, 0, s>>>(d, (int)nb);
  return hipGetLastError();
}
static Box3 interior(const Dev3& d) {
  return make_box(d.G, d.G + d.NX, d.G, d.G + d.NY, d.G, d.G + d.NZ);
}
static hipError_t launch_copy_old(const Dev3& d, hipStream_t s) {
  const Box3 b = interior(d);
  if (box_count(b) > 0) k3_copy_old<<<blocks_for(box_count(b)), kBlock3, 0, s>>>(d, b);
  return hipGetLastError();
}
// the index boxes of the reference's loops (reconstruction.cpp:3698-3753
// with the guards of :3779, 3928-3949, 4147-4161; flux_function.cpp:760-811)
// face values: the DIR index in [G-1, n-G], the others in [G-2, n-G+2)
static Box3 face_box(const Dev3& d, int dir) {
  const int G = d.G;
  return make_box(dir == 0 ? G - 1 : G - 2, dir == 0 ? d.nx - G + 1 : d.nx - G + 2, dir == 1 ? G - 1 : G - 2,
                  dir == 1 ? d.ny - G + 1 : d.ny - G + 2, dir == 2 ? G - 1 : G - 2,
                  dir == 2 ? d.nz - G + 1 : d.nz - G + 2);
}
// line values
static Box3 line_box(const Dev3& d, int dir) {
  const int G = d.G;
  if (dir == 0) return make_box(G, d.nx - G + 1, G, d.ny - G, G - 2, d.nz - G + 2);
  if (dir == 1) return make_box(G, d.nx - G, G, d.ny - G + 1, G - 2, d.nz - G + 2);
  return make_box(G - 2, d.nx - G + 2, G, d.ny - G, G, d.nz - G + 1);
}
// the faces that get a flux
static Box3 flux_box(const Dev3& d, int dir) {
  const int G = d.G;
  return make_box(G, G + d.NX + (dir == 0), G, G + d.NY + (dir == 1), G, G + d.NZ + (dir == 2));
}
static hipError_t launch_face(const Dev3& d, int dir, hipStream_t s) {
  const Box3 b = face_box(d, dir);
  if (box_count(b) == 0) return hipSuccess;
  const unsigned nb = blocks_for(box_count(b));
  if (d.mlp) {
    if (dir == 0) k3_face<0, true><<<nb, kBlock3, 0, s>>>(d, b);
    else if (dir == 1) k3_face<1, true><<<nb, kBlock3, 0, s>>>(d, b);
    else k3_face<2, true><<<nb, kBlock3, 0, s>>>(d, b);
  } else {
    if (dir == 0) k3_face<0, false><<<nb, kBlock3, 0, s>>>(d, b);
    else if (dir == 1) k3_face<1, false><<<nb, kBlock3, 0, s>>>(d, b);
    else k3_face<2, false><<<nb, kBlock3, 0, s>>>(d, b);
  }
  return hipGetLastError();
}
static hipError_t launch_line(const Dev3& d, int dir, hipStream_t s) {
  const Box3 b = line_box(d, dir);
  if (box_count(b) == 0) return hipSuccess;
  const unsigned nb = (unsigned)line_tiles(b, dir);
  if (d.mlp) {
    if (dir == 0) k3_line<0, true><<<nb, kBlock3, 0, s>>>(d, b);
    else if (dir == 1) k3_line<1, true><<<nb, kBlock3, 0, s>>>(d, b);
    else k3_line<2, true><<<nb, kBlock3, 0, s>>>(d, b);
  } else {
    if (dir == 0) k3_line<0, false><<<nb, kBlock3, 0, s>>>(d, b);
    else if (dir == 1) k3_line<1, false><<<nb, kBlock3, 0, s>>>(d, b);
    else k3_line<2, false><<<nb, kBlock3, 0, s>>>(d, b);
  }
  return hipGetLastError();
}
static hipError_t launch_flux(const Dev3& d, int dir, int stage, double dt, hipStream_t s) {
  const Box3 b = flux_box(d, dir);
  if (box_count(b) == 0) return hipSuccess;
  const unsigned nb = (unsigned)pflux_tiles(b, dir);
  if (d.mlp) {
    if (dir == 0) k3_pflux<0, true><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
    else if (dir == 1) k3_pflux<1, true><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
    else k3_pflux<2, true><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
  } else {
    if (dir == 0) k3_pflux<0, false><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
    else if (dir == 1) k3_pflux<1, false><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
    else k3_pflux<2, false><<<nb, kPfBlock, 0, s>>>(d, b, stage, dt);
  }
  return hipGetLastError();
}
static hipError_t launch_update(const Dev3& d, hipStream_t s) {
  const Box3 b = interior(d);
  if (box_count(b) > 0) k3_update<<<blocks_for(box_count(b)), kBlock3, 0, s>>>(d, b);
  return hipGetLastError();
}
static hipError_t launch_flux_points(const Gas3& g, long long n, const double* L, const double* R, double dt,
                                     double* C, double* F, double* D, hipStream_t s) {
  if (n > 0) k3_flux_points<<<blocks_for(n), kBlock3, 0, s>>>(g, n, L, R, dt, C, F, D);
  return hipGetLastError();
}

// every cell any direction's face values read: [G-2, n-G+2) on each axis
static hipError_t launch_classify(const Dev3& d, hipStream_t s) {
  const int G = d.G;
  const Box3 b = make_box(G - 2, d.nx - G + 2, G - 2, d.ny - G + 2, G - 2, d.nz - G + 2);
  if (d.mlp && box_count(b) > 0) k3_classify<<<blocks_for(box_count(b)), kBlock3, 0, s>>>(d, b);
  return hipGetLastError();
}

extern const Launch3dTable table3d = {launch_bc,     launch_bc_side,     launch_primvar, launch_copy_old,
                                      launch_face,   launch_line,        launch_flux,    launch_update,
                                      launch_flux_points, launch_classify};
#endif

}  // namespace GKS_VARIANT
