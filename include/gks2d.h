/*
 * gks2d.h — C ABI of the MI355X (gfx950) GKS hot path: WENO-AO interface
 * reconstruction, the GKS time-dependent flux, the two-stage fourth-order
 * (S2O4) update and the ghost-cell fills, on a device-resident FP64 state.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns a status code (GKS_OK == 0).  The functions replace the 2D sweep
 * functions of the reference solver one for one; each declaration cites the
 * reference function it stands in for (paths relative to the reference repo).
 *
 * State convention (same as the reference): cell (i, j), i in [0, nx),
 * j in [0, ny), nx = nodex + 2*ghost, ny = nodey + 2*ghost, flat index
 * i*ny + j (y fastest).  Conservative variables (rho, rho*u, rho*v, rho*E).
 * Host arrays are described by (base pointer, byte stride between cells,
 * byte stride between the 4 components), so both the reference's AoS
 * `Fluid2d` records and plain SoA arrays can be read and written in place.
 */
#ifndef GKS2D_H
#define GKS2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference prints and calls exit(0) instead) */
enum {
  GKS_OK = 0,
  GKS_ERR_ARG = 1,          /* bad argument / null pointer / size */
  GKS_ERR_UNSUPPORTED = 2,  /* scheme combination not implemented */
  GKS_ERR_HIP = 3,          /* HIP runtime error */
  GKS_ERR_BLOWUP = 4,       /* rho<0, p<0 or NaN seen (basic_function.cpp:752-773) */
  GKS_ERR_NOMEM = 5,
  GKS_ERR_STATE = 6         /* call out of order */
};

/* boundary kinds, reference boundary_condition.h:13-34 */
enum { GKS_BC_FREE = 0, GKS_BC_PERIODIC = 1, GKS_BC_REFLECT = 2 };
/* sides in the reference's call order (2Dproblem.cpp:139-142) */
enum { GKS_SIDE_LEFT = 0, GKS_SIDE_RIGHT = 1, GKS_SIDE_DOWN = 2, GKS_SIDE_UP = 3 };
/* reconstruction.h:4 */
enum { GKS_VAR_CONSERVATIVE = 0, GKS_VAR_CHARACTERISTIC = 1 };
/* basic_function.h:15 */
enum { GKS_TAU_EULER = 0, GKS_TAU_NS = 1 };

/*
 * Problem parameters: a snapshot of the reference's process-wide globals
 * (basic_function.h:5-16, fluid_mesh.h:78-80, reconstruction.h:4-11) and of
 * the Block2d geometry (fluid_mesh.h:82-121).  The time scheme is S2O4_2D
 * (time_advance.cpp:221-230) and the flux is GKS2D with gks2nd_2d
 * (flux_function.cpp:345-698).
 */
typedef struct gks2d_params {
  int32_t nodex, nodey;      /* interior cells */
  int32_t ghost;             /* ghost layers (5 for WENO5/7/9) */
  int32_t gausspoint;        /* 2 (WENO5), 3 (WENO7), 4 (WENO9, Lobatto) */
  int32_t weno_order;        /* 5, 7 or 9 */
  int32_t recon_variable;    /* GKS_VAR_* */
  int32_t use_df;            /* 1: *_with_df reconstructions + Update_alpha */
  int32_t tau_type;          /* GKS_TAU_* */
  int32_t K;                 /* internal degrees of freedom */
  int32_t prandtl_fix;       /* is_Prandtl_fix */
  double gamma, mu, nu, pr;  /* Gamma, Mu, Nu, Pr */
  double c1_euler, c2_euler; /* numerical collision time coefficients */
  double left, right, down, up; /* domain */
  double cfl;
  int32_t fast_math;         /* 0: IEEE parity build; 1: contracted FMAs */
  int32_t df_frozen;         /* 1: the DF sensor is never updated (no Update_alpha in
                                gks2d_step / gks2d_run / batches), as in the reference's
                                smooth accuracy driver sinwave_2d, 2Dproblem.cpp:1514-1519 */
} gks2d_params;

typedef struct gks2d_ctx gks2d_ctx;

/* Creates the device state (U, U_old, DF alpha, face buffers) for one block.
 * Replaces Setfluid / Setinterface_array / Setflux_gauss_array /
 * SetUniformMesh (basic_function.cpp:1310-1491) for the device side. */
int gks2d_create(const gks2d_params* p, gks2d_ctx** out);
int gks2d_destroy(gks2d_ctx* ctx);

/* Host <-> device copies of the conservative state of all nx*ny cells. */
int gks2d_set_convar(gks2d_ctx* ctx, const double* base, int64_t cell_stride_bytes,
                     int64_t var_stride_bytes);
int gks2d_get_convar(gks2d_ctx* ctx, double* base, int64_t cell_stride_bytes,
                     int64_t var_stride_bytes);
/* DF shock-sensor fields alpha_x / alpha_y (fluid_mesh.h:163-164). */
int gks2d_set_alpha(gks2d_ctx* ctx, const double* ax, const double* ay,
                    int64_t cell_stride_bytes);
int gks2d_get_alpha(gks2d_ctx* ctx, double* ax, double* ay, int64_t cell_stride_bytes);
/* Primitive variables the next gks2d_cfl reads (the reference's Get_CFL reads
 * fluids[].primvar, which at step 0 is what the initial condition wrote,
 * 2Dproblem.cpp:345-385). */
int gks2d_set_cfl_primvar(gks2d_ctx* ctx, const double* base, int64_t cell_stride_bytes,
                          int64_t var_stride_bytes);

/* CopyFluid_new_to_old (basic_function.cpp:1571-1584). */
int gks2d_copy_old(gks2d_ctx* ctx);
/* Get_CFL + Dtx (time_advance.cpp:264-305): dt from the primitive variables
 * of the last gks2d_primvar call (the reference's stale primvar), clamped to
 * tstop - t + 1e-16 on the last step.  Returns GKS_ERR_BLOWUP if any
 * gks2d_primvar since the previous call saw rho<0, p<0 or NaN. */
int gks2d_cfl(gks2d_ctx* ctx, double t, double tstop, double* dt_out);
/* One ghost fill: free_boundary_* / periodic_boundary_* / reflection_boundary_*
 * (boundary_condition.cpp:66-186, 367-442). */
int gks2d_boundary(gks2d_ctx* ctx, int side, int kind);
/* Convar_to_Primvar (basic_function.cpp:752-773): blow-up check over all
 * cells and the primitive state for the next gks2d_cfl. Non-blocking. */
int gks2d_primvar(gks2d_ctx* ctx);
/* Reconstruction_within_cell + Reconstruction_forg0
 * (reconstruction.cpp:1415-1437, 3612-3687). */
int gks2d_reconstruct(gks2d_ctx* ctx);
/* Calculate_flux with GKS2D (flux_function.cpp:317-698). */
int gks2d_flux(gks2d_ctx* ctx, int stage, double dt);
/* Update + Update_with_gauss (time_advance.cpp:307-474). */
int gks2d_update(gks2d_ctx* ctx, int stage);
/* Update_alpha (reconstruction.cpp:1523-1534). */
int gks2d_update_alpha(gks2d_ctx* ctx);

/* One whole S2O4 step = the driver's loop body (2Dproblem.cpp:128-158):
 * copy_old, then per stage the four boundary fills (bc[4] in GKS_SIDE_*
 * order), primvar, reconstruct, flux, update and (use_df) update_alpha. */
int gks2d_step(gks2d_ctx* ctx, double dt, const int32_t bc[4]);
/* The driver's time loop: while (t < tstop && steps < max_steps) {dt = cfl;
 * step}.  Keeps the device busy by reading dt back asynchronously.
 * t_inout / step_inout carry block.t / block.step.  The states, dt and
 * step counts equal those of gks2d_step in the loop; its steps launch no
 * copy_old (stage 0 writes its state to the second buffer, which afterwards
 * holds that state rather than a copy of the step's initial one) and form
 * Convar_to_Primvar's blow-up test and CFL speed in the updates (the
 * primvar sweep runs at the run's first stage only). */
int gks2d_run(gks2d_ctx* ctx, const int32_t bc[4], double tstop, int32_t max_steps,
              double* t_inout, int32_t* step_inout);

/* Blocks until all queued device work of ctx is done. */
int gks2d_sync(gks2d_ctx* ctx);

/* Device time per kernel phase, measured with HIP events on the context's
 * stream (no counterpart in the reference, whose only timer is clock() around
 * the whole loop, 2Dproblem.cpp:164-167).  gks2d_timing(ctx, 1) resets and
 * starts accumulating; gks2d_get_timing fills ms[GKS_NPHASE] (summed
 * milliseconds) and launches[GKS_NPHASE]. */
enum {
  GKS_PHASE_BC = 0, GKS_PHASE_PRIMVAR, GKS_PHASE_NORMAL_X, GKS_PHASE_NORMAL_Y,
  GKS_PHASE_TANGENT_X, GKS_PHASE_TANGENT_Y, GKS_PHASE_FLUX_X, GKS_PHASE_FLUX_Y,
  GKS_PHASE_UPDATE, GKS_PHASE_ALPHA, GKS_PHASE_COPY_OLD,
  /* learned path (gks_learn.h): the second launch of a normal / flux sweep,
   * over the cells / faces the first one listed for full WENO5-AO-DF */
  GKS_PHASE_NORMAL_LIST_X, GKS_PHASE_NORMAL_LIST_Y, GKS_PHASE_FLUX_LIST_X, GKS_PHASE_FLUX_LIST_Y,
  GKS_NPHASE
};
int gks2d_timing(gks2d_ctx* ctx, int enable);
/* Times only the phases whose bit (1 << GKS_PHASE_x) is set in mask (0: off;
 * gks2d_timing(ctx, 1) is the mask of every phase).  Each timed launch is
 * bracketed by two events on the stream, which cost the step a few
 * microseconds each, so a caller that only needs one kernel's duration
 * times only that phase.  Resets the accumulators like gks2d_timing. */
int gks2d_timing_phases(gks2d_ctx* ctx, uint32_t mask);
int gks2d_get_timing(gks2d_ctx* ctx, double* ms, int64_t* launches);

/* Leaf parity entry: GKS2D (flux_function.cpp:345-698) preceded by the
 * collision centre state (Center_all_collision_2d_multi,
 * reconstruction.cpp:3652-3687) for n independent points.  left/right/center
 * are n x 12 doubles (convar[4], der1x[4], der1y[4]) in the face-local
 * frame; f/derf are n x 4.  Host pointers. */
int gks2d_flux_points(const gks2d_params* p, int64_t n, const double* left,
                      const double* right, double dt, double* center, double* f,
                      double* derf);

/* Parity support (no counterpart in the reference, whose face records are
 * always on the host): with on=1, gks2d_flux also stores the raw f/derf of
 * every (face, Gauss point) so gks2d_get_face_fluxes can return them.  Off by
 * default: the update reads only the stage-combined flux. */
int gks2d_keep_face_fluxes(gks2d_ctx* ctx, int on);

/* Per-face fluxes of the last gks2d_flux call (needs gks2d_keep_face_fluxes
 * on during that call), for parity checks:
 * xf: (nx+1) x ny faces, yf: nx x (ny+1) faces, each face gausspoint x 8
 * doubles (f[4], derf[4]) in the face-local frame.  Faces outside the range
 * the reference computes are left untouched. */
int gks2d_get_face_fluxes(gks2d_ctx* ctx, double* xf, double* yf);

/* Name of the last error on this thread. */
const char* gks2d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GKS2D_H */
