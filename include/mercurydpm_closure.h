/*
 * mercurydpm_closure.h -- the learned closure of the granular rheology:
 * multilayer perceptrons (a small FP64 one, 4 -> 32 -> 32 -> 2 tanh, and the
 * configured one, 4 -> 3 x 128 GELU -> 2 with BF16 weights) from the
 * coarse-grained state of a flow (inertial number I, packing fraction phi,
 * pressure P, granular temperature T; standardised by the caller) to the
 * effective friction mu = -sigma_xz / sigma_zz and the packing fraction.  It
 * has no counterpart in the reference, which takes such closures as fixed
 * laws (mu(I) fits to CG output, Drivers/MercuryCG); the training data are the
 * reference's own CG fields (dem_cg_z).  Forward pass, gradients and Adam
 * updates run on the GPU; part of libmercurydpm_dem.so.
 */
#ifndef MERCURYDPM_CLOSURE_H
#define MERCURYDPM_CLOSURE_H

#include <stdint.h>

#include "mercurydpm_dem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DEM_MLP_HIDDEN 32
/* W1[32][4], b1[32], W2[32][32], b2[32], W3[2][32], b3[2], row-major, FP64 */
#define DEM_MLP_PARAMS 1282

typedef struct dem_mlp dem_mlp;

int dem_mlp_create(dem_mlp** out, int32_t device);
int dem_mlp_destroy(dem_mlp* mlp);
/* parameters in the layout above; setting them also resets the Adam state */
int dem_mlp_set_params(dem_mlp* mlp, const double* params);
int dem_mlp_get_params(dem_mlp* mlp, double* params);
/* y (n x 2) = network(x (n x 4)) */
int dem_mlp_forward(dem_mlp* mlp, int64_t n, const double* x, double* y);
/* loss = (1/n) sum over samples and outputs of (y - target)^2, and its
 * gradient with respect to every parameter (grad may be NULL) */
int dem_mlp_gradient(dem_mlp* mlp, int64_t n, const double* x, const double* target, double* grad, double* loss);
/* `epochs` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) at learning
 * rate lr; loss: the loss before the last step */
int dem_mlp_train(dem_mlp* mlp, int64_t n, const double* x, const double* target, int32_t epochs, double lr,
                  double* loss);

/* ---- the configured closure network ------------------------------------
 * 4 -> 128 -> 128 -> 128 -> 2, GELU hidden layers (0.5 z (1 + erf(z/sqrt 2))),
 * linear output; BF16 weights (rounded to nearest even from FP32 master
 * weights at every pass) and BF16 activations into the 128 x 128 layers,
 * which run on the matrix cores; FP32 accumulation, pre-activations, GELU
 * and Adam.  Same roles as dem_mlp_*, FP32 arrays. */
#define DEM_CLOSURE_HIDDEN 128
/* W1[128][4], b1[128], W2[128][128], b2[128], W3[128][128], b3[128],
 * W4[2][128], b4[2], row-major, FP32 */
#define DEM_CLOSURE_PARAMS 33922

typedef struct dem_closure dem_closure;

int dem_closure_create(dem_closure** out, int32_t device);
int dem_closure_destroy(dem_closure* net);
/* parameters in the layout above; setting them also resets the Adam state */
int dem_closure_set_params(dem_closure* net, const float* params);
int dem_closure_get_params(dem_closure* net, float* params);
/* y (n x 2) = network(x (n x 4)) */
int dem_closure_forward(dem_closure* net, int64_t n, const float* x, float* y);
/* loss = (1/n) sum over samples and outputs of (y - target)^2 and its
 * gradient with respect to every (FP32 master) parameter, the BF16 weights
 * standing in for them (grad may be NULL) */
int dem_closure_gradient(dem_closure* net, int64_t n, const float* x, const float* target, float* grad,
                         float* loss);
/* `epochs` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) at learning
 * rate lr; loss: the loss before the last step */
int dem_closure_train(dem_closure* net, int64_t n, const float* x, const float* target, int32_t epochs, float lr,
                      float* loss);

#ifdef __cplusplus
}
#endif

#endif
