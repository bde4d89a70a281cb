/*
 * gks3d.h — C ABI of the 3D GKS path on the GPU (libgks3d.so).
 *
 * Replaces the 3D driver loop of the reference (CubicTube,
 * src/3Dproblem.cpp:109-166: CopyFluid_new_to_old, Get_CFL, and per S2O4
 * stage the six boundary functions, Convar_to_primvar,
 * Reconstruction_within_cell (face, line and point values,
 * reconstruction.cpp:3698-3753), Calculate_flux with GKS3D
 * (flux_function.cpp:760-1163) and Update_gauss (time_advance.cpp:850-999)).
 *
 * The reference's 3D path does not run as shipped (Interface3d::gauss is
 * never allocated, basic_function.cpp:1896-1904, and GKS3D takes the
 * pressure for the Maxwellian's lambda, flux_function.cpp:832-838); this
 * path keeps its algorithm with lambda from the conservative state, decides
 * the line / point order-reduction fallback at every reconstruction, and
 * adds periodic boundaries.  Scheme: WENO5-AO (wenoz weights, no DF factor)
 * in conservative or characteristic variables, 2 x 2 Gauss points per face,
 * GKS3D gks2nd with the S2O4 (dt, dt/2) combination, Euler or
 * Navier-Stokes collision time.
 *
 * State arrays: conservative variables (rho, rho u, rho v, rho w, rho E) of
 * every cell including ghosts, cell (i, j, k) at (i*ny + j)*nz + k
 * (z fastest), with byte strides between cells and between variables as in
 * gks2d.h.  Status codes and boundary kinds (free, periodic) as in gks2d.h.
 */
#ifndef GKS3D_H
#define GKS3D_H

#include <stdint.h>

#include "gks2d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sides in the reference's call order (3Dproblem.cpp:137-142) */
enum { GKS_SIDE3_XLEFT = 0, GKS_SIDE3_XRIGHT = 1, GKS_SIDE3_YLEFT = 2, GKS_SIDE3_YRIGHT = 3,
       GKS_SIDE3_ZLEFT = 4, GKS_SIDE3_ZRIGHT = 5 };

typedef struct gks3d_params {
  int32_t nodex, nodey, nodez; /* interior cells */
  int32_t ghost;               /* ghost layers: 3 (the reference's CubicTube) or more */
  int32_t recon_variable;      /* GKS_VAR_* */
  int32_t tau_type;            /* GKS_TAU_* */
  int32_t K;                   /* internal degrees of freedom (2 for gamma 1.4) */
  int32_t fast_math;           /* 0: IEEE build; 1: contracted FMAs */
  double gamma, mu, nu;        /* Gamma, Mu, Nu */
  double c1_euler, c2_euler;   /* numerical collision time coefficients */
  double left, right, down, up, back, front; /* domain */
  double cfl;
} gks3d_params;

typedef struct gks3d_ctx gks3d_ctx;

int gks3d_create(const gks3d_params* p, gks3d_ctx** out);
int gks3d_destroy(gks3d_ctx* ctx);
int gks3d_set_convar(gks3d_ctx* ctx, const double* base, int64_t cell_stride, int64_t var_stride);
int gks3d_get_convar(gks3d_ctx* ctx, double* base, int64_t cell_stride, int64_t var_stride);
/* the primitive variables (rho, u, v, w, p) the first step's CFL reads
 * (Get_CFL reads the primvar of the last stage; at step 0 the initial
 * condition's) */
int gks3d_set_cfl_primvar(gks3d_ctx* ctx, const double* base, int64_t cell_stride, int64_t var_stride);
/* The driver loop: steps until *t reaches tstop or max_steps steps are
 * taken; bc[6] per side (GKS_BC_FREE or GKS_BC_PERIODIC).  Returns
 * GKS_ERR_BLOWUP if rho or p turns negative or NaN (that step is not
 * counted). */
int gks3d_run(gks3d_ctx* ctx, const int32_t bc[6], double tstop, int32_t max_steps, double* t, int32_t* step);
int gks3d_sync(gks3d_ctx* ctx);

/* One sweep at a time, for a driver that keeps the reference's loop
 * (3Dproblem.cpp:109-166) on the host:
 *  gks3d_boundary   one side's ghost fill (free_boundary_xleft ...,
 *                   boundary_condition.cpp:445-570; periodic added), side
 *                   GKS_SIDE3_*, kind GKS_BC_FREE or GKS_BC_PERIODIC;
 *  gks3d_primvar    Convar_to_primvar over the interior: the next CFL's
 *                   speed and density, the blow-up flag;
 *  gks3d_copy_old   CopyFluid_new_to_old;
 *  gks3d_cfl        Get_CFL / Dtx_3d (time_advance.cpp:793-848) from the last
 *                   gks3d_primvar (or gks3d_set_cfl_primvar), clipped to
 *                   tstop; GKS_ERR_BLOWUP (flag cleared) if the last sweeps
 *                   saw rho or p negative or NaN;
 *  gks3d_flux       Reconstruction_within_cell + Calculate_flux of stage s
 *                   (face, line and point values, GKS3D, the S2O4 stage sum);
 *  gks3d_update     Update_gauss of stage s. */
int gks3d_boundary(gks3d_ctx* ctx, int32_t side, int32_t kind);
int gks3d_primvar(gks3d_ctx* ctx);
int gks3d_copy_old(gks3d_ctx* ctx);
int gks3d_cfl(gks3d_ctx* ctx, double t, double tstop, double* dt);
int gks3d_flux(gks3d_ctx* ctx, int32_t stage, double dt);
int gks3d_update(gks3d_ctx* ctx, int32_t stage);
/* Device time per kernel phase (all three directions summed), measured with
 * HIP events on the context's stream (no counterpart in the reference, whose
 * only timer is omp_get_wtime around the loop, 3Dproblem.cpp:112-169).
 * gks3d_timing(ctx, 1) resets and starts accumulating; gks3d_get_timing
 * fills ms[GKS3_NPHASE] (summed milliseconds) and launches[GKS3_NPHASE]. */
enum {
  GKS3_PHASE_BC = 0, GKS3_PHASE_PRIMVAR, GKS3_PHASE_FACE, GKS3_PHASE_LINE,
  GKS3_PHASE_FLUX,  /* point values + GKS3D, one launch per direction */
  GKS3_PHASE_UPDATE, GKS3_PHASE_COPY_OLD,
  GKS3_PHASE_CLASSIFY,  /* learned path: the classifier over every cell, once per stage */
  GKS3_NPHASE
};
int gks3d_timing(gks3d_ctx* ctx, int enable);
/* Times only the phases whose bit (1 << GKS3_PHASE_x) is set in mask (0:
 * off; gks3d_timing(ctx, 1) is the mask of every phase).  Resets the
 * accumulators like gks3d_timing. */
int gks3d_timing_phases(gks3d_ctx* ctx, uint32_t mask);
int gks3d_get_timing(gks3d_ctx* ctx, double* ms, int64_t* launches);
/* Leaf parity: GKS3D at n Gauss points.  left/right: n x 20 (convar[5],
 * der1x[5], der1y[5], der1z[5] in the face frame); center (optional):
 * n x 20 collision state and slopes; f, derf: n x 5. */
int gks3d_flux_points(const gks3d_params* p, int64_t n, const double* left, const double* right, double dt,
                      double* center, double* f, double* derf);
/* The learned reconstruction in 3D (no counterpart in the reference): the
 * troubled-interface classifier of gks_learn.h (params: GKS_MLP_NPARAM
 * floats; the same network and stencil features as the 2D path) runs once
 * per stage on every cell's x, y and z stencils (BF16 MFMA inference, FP32
 * accumulation).  A face-value reconstruction takes the linear weights (the
 * fifth-order candidate) where its cell's stencil across the face has
 * logit < logit_threshold; a line (point) reconstruction takes them where
 * its five source values did and the centre source cell's stencil along
 * the line (point) direction did too; the order-reduction test is kept and
 * GKS3D is unchanged.  params == NULL disables it. */
int gks3d_set_classifier(gks3d_ctx* ctx, const float* params, float logit_threshold);
/* Counters since the classifier was set: counts[0] / [1] face-value
 * reconstructions that took the linear weights / all of them; [2] / [3]
 * line values (per face side); [4] / [5] point values (per face side and
 * line). */
int gks3d_classifier_counts(gks3d_ctx* ctx, int64_t counts[6]);
const char* gks3d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GKS3D_H */
