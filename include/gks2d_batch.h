/*
 * gks2d_batch.h — many independent 2D problems stepped together.
 *
 * Not a reference interface: the reference runs one block per process and
 * keeps its scheme settings in globals (basic_function.cpp:3-13).  A batch
 * holds one gks2d_ctx per problem, each with its own gks2d_params snapshot
 * (grid, gas, CFL), so problems of different sizes and settings share one
 * set of launches: every sweep of a step is one launch over the tiles of all
 * problems, and a whole S2O4 step (dt, copy, two stages of fills, primitive
 * variables, reconstruction, flux, update, DF sensor) is captured once as a
 * HIP graph and replayed.  Each problem keeps the driver loop of gks2d_run:
 * its own dt from its own stale stage-1 speed, its own tstop and step limit,
 * so a batched problem's state equals a gks2d_run of that problem alone.
 *
 * Scope: WENO5-AO-DF with 2 Gauss points (the fused tangential + flux
 * launch), the same fast_math setting for all problems, and the learned
 * classifier on all problems or on none.
 */
#ifndef GKS2D_BATCH_H
#define GKS2D_BATCH_H

#include <stdint.h>

#include "gks2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gks2d_batch gks2d_batch;

/* n problems from n parameter snapshots (params[0..n-1]); 1 <= n <= 1024 */
int gks2d_batch_create(const gks2d_params* params, int32_t n, gks2d_batch** out);
int gks2d_batch_destroy(gks2d_batch* b);

/* The context of problem p (owned by the batch: do not destroy it).  All of
 * gks2d.h works on it between batch runs: initial conditions, state copies,
 * gks2d_set_classifier, single-problem runs. */
int gks2d_batch_problem(gks2d_batch* b, int32_t p, gks2d_ctx** out);

/* Tile shape: faces per flux workgroup (64 or 128) and problem order in the
 * launches (0: as created, 1: largest grid first). */
int gks2d_batch_set_shape(gks2d_batch* b, int32_t tanflux_faces, int32_t largest_first);

/* Times every tile shape over `steps` captured steps from the current state
 * (restored afterwards, so the run that follows is unchanged) and keeps the
 * fastest.  bc: 4 kinds per problem.  Outputs may be NULL. */
int gks2d_batch_autotune(gks2d_batch* b, const int32_t* bc, int32_t steps, int32_t* tanflux_faces,
                         int32_t* largest_first, double* ms_per_step);

/* Advances every problem p until t[p] reaches tstop[p] or it has taken
 * max_steps more steps, as gks2d_run does for one problem.  bc: 4 kinds per
 * problem; t and steps (n entries each) are read and written.  The host
 * checks for completion every `check_every` graph replays.  Returns
 * GKS_ERR_BLOWUP if a problem's state went bad (that problem stops there;
 * the others run on). */
int gks2d_batch_run(gks2d_batch* b, const int32_t* bc, const double* tstop, int32_t max_steps,
                    int32_t check_every, double* t, int32_t* steps);

#ifdef __cplusplus
}
#endif

#endif /* GKS2D_BATCH_H */
