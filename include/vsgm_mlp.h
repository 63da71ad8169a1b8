/*
 * vsgm_mlp.h — the learned planner's cost model on the GPU (part of
 * libvsgm_hip.so). No counterpart in the reference, whose planner is the
 * heuristic Plan::Optimize (src/query/plan.h:61-580); that planner stays the
 * fallback, and every plan the model picks must give the same count.
 *
 * Model: y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3, hidden width 64,
 * input width in_dim <= 128. W1, W2, w3 are used in BF16 (round to nearest
 * even of FP32 master copies); every product is accumulated in FP32; biases
 * are FP32. Loss: mean over the batch of (y - t)^2 / 2.
 *
 * Two kernel paths (vsgm_mlp_set_mfma): the default runs every matrix
 * product on the matrix cores (bf16 v_mfma_f32_32x32x8_bf16, FP32
 * accumulation), with the products' activations and backward deltas
 * rounded to BF16 as well (x, relu(z1), dy w3 [z2 > 0], (dz2 W2) [h1 > 0]);
 * the other keeps activations in FP32 (one wavefront per sample, scalar
 * FMAs).
 *
 * Parameters are one FP32 vector in this order:
 *   W1 [64][in_dim] | b1 [64] | W2 [64][64] | b2 [64] | w3 [64] | b3 [1].
 * Training is Adam on the FP32 masters; after each step the BF16 copies are
 * refreshed. All device buffers are context-pool allocations.
 */
#ifndef VSGM_MLP_H_
#define VSGM_MLP_H_

#include <stdint.h>

#include "vsgm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VSGM_MLP_HIDDEN 64
#define VSGM_MLP_MAX_IN 128

typedef struct vsgm_mlp vsgm_mlp;

/* New model with He-uniform weights from a counter-based generator keyed by seed. */
int vsgm_mlp_create(vsgm_ctx* ctx, int in_dim, uint64_t seed, vsgm_mlp** out);
int vsgm_mlp_free(vsgm_mlp* m);
int vsgm_mlp_param_count(const vsgm_mlp* m, uint64_t* n);
/* FP32 master parameters to / from the host (set also resets Adam's moments). */
int vsgm_mlp_get_params(vsgm_mlp* m, float* h_params, uint64_t n);
int vsgm_mlp_set_params(vsgm_mlp* m, const float* h_params, uint64_t n);
/* y[i] = model(x[i]) for a device batch x [batch][in_dim] (FP32). */
int vsgm_mlp_forward(vsgm_mlp* m, const float* d_x, uint64_t batch, float* d_y);
/* Gradient of the batch loss with respect to the parameters (as used, i.e.
 * through the BF16 weights), kept inside the model; *h_loss gets the loss.
 * With h_grad non-NULL the gradient is also copied to the host (n floats). */
int vsgm_mlp_backward(vsgm_mlp* m, const float* d_x, const float* d_t, uint64_t batch, float* h_loss, float* h_grad, uint64_t n);
/* 1 (default): matrix-core kernels; 0: FP32-activation kernels. */
int vsgm_mlp_set_mfma(vsgm_mlp* m, int on);
/* One Adam step with the last gradient (t counts steps from 1). */
int vsgm_mlp_adam_step(vsgm_mlp* m, float lr, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif /* VSGM_MLP_H_ */
