/*
 * gks1d.h — C ABI of the 1D GKS hot path on gfx950: WENO5-AO (first step) /
 * WENO5-AO-DF reconstruction, Center_collision, GKS (gks2nd), S2O4, DF
 * sensor, CFL and ghost fills of the reference's SodTubeProblem loop
 * (reference src/1Dproblem.cpp:79-148).  The whole time loop runs inside one
 * persistent workgroup on the device (a 1D grid of a few hundred cells is far
 * too small to fill the GPU; keeping it in one workgroup removes every launch
 * and host round trip).  Same conventions as gks2d.h.
 */
#ifndef GKS1D_H
#define GKS1D_H

#include <stdint.h>

#include "gks2d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* snapshot of the reference's 1D globals and Block1d (fluid_mesh.h:8-26,
 * basic_function.h:5-16, reconstruction.h:4-11); S2O4 (time_advance.cpp:20-30) */
typedef struct gks1d_params {
  int32_t nodex;           /* interior cells */
  int32_t ghost;           /* ghost layers (5) */
  int32_t recon_variable;  /* GKS_VAR_* */
  int32_t use_df;          /* is_using_df_factor: WENO5_AO_with_DF from step 1 on + Update_alpha */
  int32_t tau_type;        /* GKS_TAU_* */
  int32_t K;
  int32_t fast_math;       /* 0: IEEE kernels; 1: fast kernels */
  int32_t df_from_start;   /* 1: cellreconstruction = WENO5_AO_with_DF before the loop
                              (accuracy_sinwave_1d, 1Dproblem.cpp:587); 0: WENO5_AO */
  double gamma, mu, nu, pr, c1_euler, c2_euler;
  double left, right;      /* domain */
  double cfl;
} gks1d_params;

typedef struct gks1d_ctx gks1d_ctx;

/* Replaces the 1D Setflux_array / SetUniformMesh (basic_function.cpp:1262-1309). */
int gks1d_create(const gks1d_params* p, gks1d_ctx** out);
int gks1d_destroy(gks1d_ctx* ctx);
/* conservative state of all nx = nodex + 2*ghost cells (strided host array) */
int gks1d_set_convar(gks1d_ctx* ctx, const double* base, int64_t cell_stride_bytes, int64_t var_stride_bytes);
int gks1d_get_convar(gks1d_ctx* ctx, double* base, int64_t cell_stride_bytes, int64_t var_stride_bytes);
/* DF sensor (Fluid1d::alpha, fluid_mesh.h:38) */
int gks1d_get_alpha(gks1d_ctx* ctx, double* alpha, int64_t cell_stride_bytes);
/* The driver's time loop (1Dproblem.cpp:79-148): while (t < tstop && steps <
 * max_steps) { copy old; dt = Get_CFL (time_advance.cpp:75-91); two stages of
 * boundary fills (bc[0] left, bc[1] right), Reconstruction_within_cell,
 * Reconstruction_forg0, Calculate_flux, Update, Update_alpha }.  The
 * reconstruction is WENO5_AO at step 0 and WENO5_AO_with_DF afterwards when
 * use_df.  t_inout / step_inout carry block.t / block.step; dt_hist
 * (optional, max_steps entries) receives each step's dt.  max_steps is
 * limited to 10^7. */
int gks1d_run(gks1d_ctx* ctx, const int32_t bc[2], double tstop, int32_t max_steps, double* t_inout,
              int32_t* step_inout, double* dt_hist);
/* Allocation-free variant of gks1d_run.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * the same kernels as gks1d_run, without a context:
 *   U       device pointer, [3][nodex + 2*ghost] doubles (SoA, the layout of
 *           a context's state), read and advanced in place;
 *   alpha   device pointer, [nodex + 2*ghost] doubles, the DF sensor read and
 *           advanced in place (ghost entries must be 0), or NULL to start
 *           from a zero sensor as a freshly created context does;
 *   stream  a hipStream_t (NULL: the null stream);
 *   bc, tstop, max_steps, t_inout, step_inout, dt_hist as for gks1d_run.
 * Every other device buffer (U_old, face states and fluxes, loop status, dt
 * record) lives in a static per-GPU scratch array of 3.71 MB, sized for
 * nodex <= 16384; larger tubes get GKS_ERR_ARG.  It calls no hipMalloc /
 * hipFree.  All calls on one GPU share that GPU's scratch array, which keeps
 * the last call's contents: calls that overlap on one GPU (other threads or
 * streams) race and must be serialised by the caller.  The call returns
 * after its last launch has finished, like gks1d_run. */
int gks1d_run_noalloc(const gks1d_params* p, double* U, double* alpha, const int32_t bc[2], double tstop,
                      int32_t max_steps, double* t_inout, int32_t* step_inout, double* dt_hist, void* stream);
/* f[3], derf[3] of every face (nx+1 faces, face-major, 6 doubles each) from
 * the last step's stage 0 or 1 (Flux1d fluxes[i][stage], fluid_mesh.h:52-58) */
int gks1d_get_face_fluxes(gks1d_ctx* ctx, int32_t stage, double* out);

/* One sweep at a time, for a driver that keeps the reference's loop
 * (1Dproblem.cpp:79-148) on the host; each is one single-workgroup launch:
 *  gks1d_boundary       one side's ghost fill (side 0 left, 1 right; kind
 *                       GKS_BC_FREE, _PERIODIC or _REFLECT;
 *                       boundary_condition.cpp:4-64);
 *  gks1d_copy_old       CopyFluid_new_to_old;
 *  gks1d_cfl            Get_CFL (time_advance.cpp:75-91): the reference's
 *                       sequential Dtx scan over the current state, clipped
 *                       to tstop (tstop - t + 1e-15);
 *  gks1d_reconstruct    Reconstruction_within_cell with WENO5_AO (with_df 0)
 *                       or WENO5_AO_with_DF (1);
 *  gks1d_flux           Reconstruction_forg0 (Center_collision) +
 *                       Calculate_flux of stage s with GKS;
 *  gks1d_update         Update (S2O4) of stage s;
 *  gks1d_update_alpha   Update_alpha (the DF sensor). */
int gks1d_boundary(gks1d_ctx* ctx, int32_t side, int32_t kind);
int gks1d_copy_old(gks1d_ctx* ctx);
int gks1d_cfl(gks1d_ctx* ctx, double t, double tstop, double* dt);
int gks1d_reconstruct(gks1d_ctx* ctx, int32_t with_df);
int gks1d_flux(gks1d_ctx* ctx, int32_t stage, double dt);
int gks1d_update(gks1d_ctx* ctx, int32_t stage);
int gks1d_update_alpha(gks1d_ctx* ctx);
/* message of the last failed call on this thread */
const char* gks1d_last_error(void);
/* leaf parity: Center_collision (reconstruction.cpp:1375-1407) + GKS
 * (flux_function.cpp:16-276) for n points; left/right n x 6 (convar[3],
 * der1[3]); center n x 6; f/derf n x 3.  Host pointers. */
int gks1d_flux_points(const gks1d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf);

#ifdef __cplusplus
}
#endif

#endif /* GKS1D_H */
