// gks2d_launch.hpp — device-state descriptor shared by the kernel variants and
// the C ABI layer, and the table of launchers each variant exports.
#pragma once
#include <hip/hip_runtime.h>

#include "gks2d_types.hpp"

struct Dev {
  int NX, NY, G, nx, ny, gp, recon_var, use_df;
  int order;             // WENO order: 5, 7 or 9 (gp 2, 3, 4)
  int fuse_tanflux;      // WENO5: tangential reconstruction inside the flux launch
  int keep_face_fluxes;  // store raw (f, derf) per face for parity checks
  int df_frozen;         // the DF sensor is never updated (gks2d_params.df_frozen)
  int presum;            // fused WENO5 flux: ak holds the face's DF value (Gauss-point mean, thresholded)
                         // and, in the fast build, xc / pc the face's Gauss-point sums (plane 0)
  double dx, dy, area;
  double idx, idy;  // 1/dx, 1/dy (fast build)
  double gw[4];
  gksd::Gas gas;
  long long nc, nfx, nfy;
  double* U;
  double* Uold;
  double* ax;
  double* ay;
  double* line_x;
  double* line_y;
  double* gst_x;
  double* gst_y;
  double* xcx;
  double* xcy;
  double* pcx;
  double* pcy;
  double* flx[2];
  double* fly[2];
  double* akx;
  double* aky;
  const float* mlp;             // classifier parameters (gks_learn.h), null: plain WENO5-AO
  const unsigned short* mlpw;   // its weights in the MFMA operand layout (mlp_pack.hpp)
  float mlp_thr;                // logit below which a smooth stencil takes the linear weights
  unsigned long long* lstat;    // 2 x 1024 (linear, total) slot pairs: normal, then tangential; then the
                                // faces the flux lists held (x, y)
  unsigned char* lin[2];        // per cell: its x / y normal stencil took the linear weights
  int smooth_flux;              // learned path: faces with two linear sides take the smooth-flow flux
  int fused_lines;              // learned path, bit DIR: the linear line values of that direction are
                                // formed by its smooth-face flux launch from the cell averages (which also
                                // stores the ones the listed faces read); the normal sweep stores none
  // learned path: the work lists, each in kListShards shards (a run of
  // tiles of the sweep that fills a list appends to one shard, so that the
  // appends of a sweep spread over kListShards counters)
  unsigned* tlist[2];           // x / y faces left to the full flux (interior face ordinals); shard s at
                                // tlist[dir] + s * tcap[dir]
  unsigned* tcount;             // their shard counts: shard s of dir at tcount[(dir * kListShards + s) *
                                // kCountStride] (cleared before each flux sweep)
  unsigned* nlist[2];           // cells whose x / y normal stencil runs WENO5-AO-DF (ordinals of the normal
                                // sweep's range); shard s at nlist[dir] + s * ncap[dir]
  unsigned* ncount;             // their shard counts, laid out as tcount's (cleared before each normal sweep)
  unsigned tcap[2], ncap[2];    // entries one shard holds
  unsigned* lpre;               // per list (BatchList order) kListRecord words: the first tile of each shard,
                                // at [kListShards] the list's tiles, then the shards' entries (formed between
                                // the sweep that fills the list and the list's launch, which clears the shard
                                // counts for the list's next sweep)
  double* partial;  // per-block (max speed, blow-up flag) pairs: max(kSweepPartials, update blocks) of them
  double* stat;     // [0..3]: max speed of slot 0 / 1 (+pad); [4]: sticky blow-up flag; batched steps:
                    // [5] the blow-up flag of the last stage-1 update (or of the initial state), [6] that
                    // flag of the step before, which the current step's end takes
};

// workgroup size of the sweep kernels and the number of per-block partials
// of the CFL speed reduction (the host sizes batched launches with them)
constexpr int kSweepBlock = 256;
constexpr int kSweepPartials = 1024;
// learned-path work lists: shards per list, and the distance of two shard
// counters in 32-bit words (one 128-byte line each: atomics on one line
// queue behind each other)
constexpr int kListShards = 16;
constexpr int kCountStride = 32;
// words of a list's prefix record (Dev::lpre): kListShards first tiles, the
// tile count, kListShards entry counts
constexpr int kListRecord = 48;
// learned path: directions (bit 0 x, bit 1 y) whose smooth-face flux
// launch forms the linear line values from the cell averages
constexpr int kFusedLineDirs = 3;
// the list launchers' direction argument for both directions' lists (one
// prefix launch for the two)
constexpr int kBothDirs = 2;
// learned path: the smooth-face flux launch takes tiles of TF faces,
// TF / kCheapCols rows (i) by kCheapCols faces (j)
constexpr int kCheapCols = 16;
__host__ __device__ inline long long cheap_tile_cols(const Dev& d, int dir) {
  const int nj = dir == 0 ? d.NY : d.NY + 1;  // faces per row
  return (nj + kCheapCols - 1) / kCheapCols;
}
__host__ __device__ inline long long cheap_tiles(const Dev& d, int dir, int tf) {
  const int rows = tf / kCheapCols, ni = dir == 0 ? d.NX + 1 : d.NX;
  return (long long)((ni + rows - 1) / rows) * cheap_tile_cols(d, dir);
}
// learned path with both directions' line values formed by the smooth-face
// flux launch: one classifier launch takes both sweeps, in tiles of
// kClsRows x kClsCols cells of the whole frame (ghost cells included)
constexpr int kClsCols = 16, kClsRows = kSweepBlock / kClsCols;
__host__ __device__ inline long long cls_tile_cols(const Dev& d) { return (d.ny + kClsCols - 1) / kClsCols; }
__host__ __device__ inline long long cls_tiles(const Dev& d) {
  return (long long)((d.nx + kClsRows - 1) / kClsRows) * cls_tile_cols(d);
}

// ------------------------------------------------------------------ batches
// Per-problem time-loop state of a batch, on the device: the driver loop of
// gks2d_run (dt from the stale stage-1 speed, clipped to tstop; stop at
// tstop, at max_steps or on blow-up) kept per problem so that a captured
// step needs no host round trip.
struct BState {
  double t, tstop, dt, cfl;
  int active;  // this step runs for the problem
  int done;    // reached tstop or max_steps, or blew up
  int slot;    // stat slot the next dt reads (the last primvar's)
  int steps, max_steps;
  int blowup;
  int bc[4];
};
// work classes of a batched launch; each has its own tile prefix table
enum BatchKind {
  KB_BC_LR = 0, KB_BC_DU, KB_NORMAL_X, KB_NORMAL_Y, KB_FLUX_X, KB_FLUX_Y, KB_UPDATE, KB_ALPHA, KB_COUNT
};
// start[k * (np + 1) + q]: first tile of the q-th problem in launch order for
// work class k (start[.. + np] = tiles in the launch); pid[q]: that problem.
// np counts the problems the launch covers (the unfinished ones when the
// tables were built); dev and st are indexed by problem id.
struct BatchArgs {
  const Dev* dev;
  BState* st;
  const unsigned* start;
  const int* pid;
  int np;
  // learned path: 4 x (np + 1) tile prefixes of the work lists (normal x,
  // normal y, flux x, flux y), in launch order, filled on the device after
  // the sweep that builds each list
  unsigned* lstart;
  // stage-1 launches of the batched step: every problem's U and Uold
  // exchanged (stage 0 writes its state to the second buffer)
  int swap;
};
enum BatchList { KL_NORMAL_X = 0, KL_NORMAL_Y, KL_FLUX_X, KL_FLUX_Y, KL_COUNT };
// tile shapes of a batch (chosen by gks2d_batch_autotune)
struct BatchShape {
  int tanflux_faces;  // faces per flux workgroup: 64 or 128
  int learned;        // the classifier runs (set on every problem)
};

// device phases timed by gks2d_timing (order of GKS_PHASE_* in gks2d.h)
enum Phase {
  PH_BC = 0, PH_PRIMVAR, PH_NORMAL_X, PH_NORMAL_Y, PH_TANGENT_X, PH_TANGENT_Y,
  PH_FLUX_X, PH_FLUX_Y, PH_UPDATE, PH_ALPHA, PH_COPY_OLD,
  PH_NORMAL_LIST_X, PH_NORMAL_LIST_Y, PH_FLUX_LIST_X, PH_FLUX_LIST_Y, PH_COUNT
};

struct LaunchTable {
  hipError_t (*bc)(const Dev&, int side, int kind, hipStream_t);
  // the four sides' fills (kind[4]: left, right, down, up) as bc one after
  // the other gives them
  hipError_t (*bc_all)(const Dev&, const int* kind, hipStream_t);
  // Convar_to_Primvar's blow-up test (flag: the stat word it raises, 4) and
  // the CFL speed into stat slot `slot` (none for slot < 0)
  hipError_t (*primvar)(const Dev&, const double* prim, int slot, int flag, hipStream_t);
  hipError_t (*copy_old)(const Dev&, hipStream_t);
  // the ghost frame of `from` into the state (U)
  hipError_t (*copy_ghosts)(const Dev&, const double* from, hipStream_t);
  hipError_t (*normal)(const Dev&, int dir, hipStream_t);
  hipError_t (*tangent)(const Dev&, int dir, hipStream_t);
  hipError_t (*flux)(const Dev&, int dir, int stage, double dt, hipStream_t);
  // learned path: the listed normal cells (after normal) and flux faces
  // (after flux) of direction dir; nothing otherwise
  hipError_t (*normal_list)(const Dev&, int dir, hipStream_t);  // dir kBothDirs: both lists
  hipError_t (*flux_list)(const Dev&, int dir, int stage, double dt, hipStream_t);  // likewise
  hipError_t (*update)(const Dev&, int stage, hipStream_t);
  // the update, and the new state's blow-up test and CFL speed into stat
  // slot `slot` (as primvar's)
  hipError_t (*update_stat)(const Dev&, int slot, hipStream_t);
  hipError_t (*alpha)(const Dev&, hipStream_t);
  hipError_t (*flux_points)(const gksd::Gas&, long long n, const double* L, const double* R, double dt,
                            double* C, double* F, double* D, int smooth, hipStream_t);
  // classifier training samples of the cells reconstructed along dir
  hipError_t (*label)(const Dev&, int dir, int stride, double tol, float* X, float* Y, hipStream_t);
  // logits of n feature rows through the solver's MFMA inference
  hipError_t (*infer)(const float* P, const unsigned short* W, long long n, const float* X, float* out,
                        hipStream_t);
  // one S2O4 step of every active problem of a batch (capturable: no host
  // synchronisation, no memset); tiles[k] = start[k * (np + 1) + np]
  hipError_t (*batch_step)(const BatchArgs&, const BatchShape&, const unsigned* tiles, hipStream_t);
  // the linear-weight reconstructions of n 5-cell stencils (leaf check)
  hipError_t (*linear_points)(long long n, const double* W, double h, double* out, hipStream_t);
};

namespace ieee { extern const LaunchTable table; }
namespace fast { extern const LaunchTable table; }
