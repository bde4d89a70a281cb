/*
 * gks_learn.h — the learned troubled-interface classifier of the 2D path.
 *
 * Not a reference interface: the reference has no learned component.  A
 * small MLP (16 inputs: the normalised 5-point density and pressure stencil
 * of a reconstruction and their relative spreads; two hidden layers of 32
 * ReLU units; one logit) predicts whether WENO5-AO's nonlinear weights would
 * move the reconstruction away from the linear (fifth-order) one.  Where the
 * network is confident the stencil is smooth and the DF sensor is zero over
 * the stencil, the reconstruction takes the linear weights; everywhere else
 * it runs the full WENO5-AO-DF of the reference.
 *
 * Parameter vector (float32, GKS_MLP_NPARAM entries, row-major):
 *   W1[16][32], b1[32], W2[32][32], b2[32], W3[32], b3
 * Training runs on the GPU (gks_mlp_train); inference inside the solver runs
 * on the matrix cores in BF16 with FP32 accumulation.
 */
#ifndef GKS_LEARN_H
#define GKS_LEARN_H

#include <stdint.h>

#include "gks2d.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GKS_MLP_NIN 16
#define GKS_MLP_NH 32
#define GKS_MLP_NPARAM (16 * 32 + 32 + 32 * 32 + 32 + 32 + 1)

/* Training samples from the current device state: for every stride-th
 * reconstruction cell along `dir` (0: x, 1: y; the cells k_normal
 * reconstructs), the 16 features and the label (1: the WENO5-AO-DF face
 * values or derivatives differ from the linear-weight ones by more than
 * label_tol times the component's stencil range, or the DF sensor is
 * nonzero on the stencil; 0: otherwise).  Writes at most `cap` samples;
 * *n_out = count. */
int gks2d_classifier_samples(gks2d_ctx* ctx, int32_t dir, int32_t stride, double label_tol, float* features,
                             float* labels, int64_t cap, int64_t* n_out);

/* Mean weighted binary cross-entropy of the logits and its gradient with
 * respect to the parameters, on the GPU in FP32 (n samples, host arrays). */
int gks_mlp_loss_grad(const float* params, int64_t n, const float* X, const float* y, const float* w, float* grad,
                      double* loss);

/* Adam on the GPU (beta1 0.9, beta2 0.999, eps 1e-8): `epochs` passes over
 * the n samples in a seeded shuffled order, mini-batches of `batch`.
 * params is updated in place; loss_per_epoch (optional, `epochs` entries)
 * receives each epoch's mean loss. */
int gks_mlp_train(float* params, int64_t n, const float* X, const float* y, const float* w, int32_t batch,
                  int32_t epochs, float lr, uint64_t seed, float* loss_per_epoch);

/* Logits of n samples through the solver's inference kernel (BF16 weights
 * and activations on the matrix cores, FP32 accumulation). */
int gks_mlp_infer_bf16(const float* params, int64_t n, const float* X, float* logits);

/* Enables the learned reconstruction in a 2D context (WENO5 only): a
 * normal stencil takes the linear weights where logit < threshold and the
 * DF sensor is zero on it; a tangential stencil takes them where the normal
 * stencils of the cells its line values come from (and the centre cell's
 * stencil across them) all did and its DF sensor is zero; a Gauss point
 * whose two sides both took them gets the smooth-flow GKS flux (see
 * gks2d_flux_points_smooth; gks2d_classifier_options turns that off).
 * params == NULL disables the classifier. */
int gks2d_set_classifier(gks2d_ctx* ctx, const float* params, float logit_threshold);

/* smooth_flux: 1 (default) or 0 (every Gauss point keeps GKS2D) */
int gks2d_classifier_options(gks2d_ctx* ctx, int32_t smooth_flux);

/* The smooth-flow flux of n Gauss points (layout of gks2d_flux_points):
 * interface state and slopes = the averages of the two sides, distribution
 * g0 [1 - tau (a u + b v + A) + t A], f = rho0 dt [<u psi g0> - tau
 * (<u psi g0 (a u + b v)> + <u psi g0 A>)], derf = rho0 dt^2 <u psi g0 A>:
 * what GKS2D gives when the two sides are equal, without its half-space
 * (erfc, exp) moments. */
int gks2d_flux_points_smooth(const gks2d_params* p, int64_t n, const double* left, const double* right, double dt,
                             double* f, double* derf);

/* The linear-weight reconstructions the learned paths use (2D and 3D), for
 * n 5-cell stencils (n x 5 cell averages, host): per stencil 8 doubles, the
 * fifth-order candidate's value and derivative at the centre cell's left
 * face, at its right face, then value and derivative at its two Gauss
 * points (-1/(2 sqrt 3) and +1/(2 sqrt 3) cell widths from the centre; the
 * reference's weno_5th_ao_with_df_left/right / _2gauss with the AO weights
 * at their linear values, reconstruction.cpp:480-657, 2742-2838), in the
 * build p->fast_math selects.  h: the cell width. */
int gks2d_linear_weights_points(const gks2d_params* p, int64_t n, const double* stencils, double h, double* out);

/* Reconstructions that took the linear weights / all reconstructions, over
 * the calls since the classifier was set (normal and tangential). */
int gks2d_classifier_stats(gks2d_ctx* ctx, int64_t* linear, int64_t* total);

/* Counters since the classifier was set: counts[0] / [1] normal
 * reconstructions that took the linear weights / all of them; [2] / [3]
 * the same for tangential stencils (one per face side); [4] / [5] the x /
 * y flux faces that went to the full-flux list (the others took the
 * smooth-flow flux). */
int gks2d_classifier_counts(gks2d_ctx* ctx, int64_t counts[6]);

/* message of the last failed gks_mlp_* call on this thread (the gks2d_*
 * functions above report through gks2d_last_error) */
const char* gks_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GKS_LEARN_H */
