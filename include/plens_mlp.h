/*
 * plens_mlp.h -- learned troubled-cell indicator (no counterpart in the reference).
 *
 * A 125-64-64-1 MLP (GELU hidden layers, 0.5*sigmoid output) trained on the GPU by hand-written
 * forward / backward / Adam kernels, FP32 master weights rounded to BF16 for every forward pass and
 * FP32 accumulation. Inputs: the (N+1)^3 Legendre modes of rho*p of a cell divided by the cell mean
 * (mode 0), zero-padded to 125; target: the plain blender value of BlenderCalculator::get_blender
 * (dgsem/blender_calculator.cc:123-187) before wall override and diffusion.
 * Parameter packing (NPARAM = 12289 floats): W1[64][125], b1[64], W2[64][64], b2[64], w3[64], b3.
 */
#ifndef PLENS_MLP_H
#define PLENS_MLP_H
#include <stdint.h>
#include "plens_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plens_mlp plens_mlp;

int32_t plens_mlp_n_params(void);
int plens_mlp_create(const float* init_params, plens_mlp** out);
int plens_mlp_destroy(plens_mlp* m);
int plens_mlp_get_params(const plens_mlp* m, float* params);
/* device pointer of the BF16-rounded parameters, for plens_gpu_set_learned_indicator() */
const float* plens_mlp_device_bf16_params(const plens_mlp* m);
int plens_mlp_predict(plens_mlp* m, const float* X, int64_t n, float* y);
/* gradients of the mean squared error of one batch, no update */
int plens_mlp_grad(plens_mlp* m, const float* X, const float* target, int32_t n, float* grads, float* loss);
/* one Adam update (lr) from the gradients of one batch */
int plens_mlp_step(plens_mlp* m, const float* X, const float* target, int32_t n, float lr);
/* mini-batch Adam over `epochs` seeded permutations of (X, target); mean batch loss per epoch */
int plens_mlp_train(plens_mlp* m, const float* X, const float* target, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float* epoch_loss);

/* Solver side: use the network (device parameters from plens_mlp_device_bf16_params, which must
 * outlive its use; NULL = plain indicator) for the cell blender value. Network values below `floor`
 * count as 0 (the network's own alpha_min) and are capped at alpha_max; wherever
 * |alpha_mlp - alpha_plain| > tolerance the plain value is used. Degrees 3 and 4 only
 * (PLENS_E_INVALID otherwise). No counterpart in the reference, whose blender value is always
 * BlenderCalculator::get_blender (dgsem/blender_calculator.cc:123-187). */
int plens_gpu_set_learned_indicator(plens_gpu* h, const float* mlp_params_device, double tolerance, double floor);
/* cells evaluated by the network and fallbacks to the plain value since the last reset */
int plens_gpu_indicator_stats(plens_gpu* h, int64_t* evaluated, int64_t* fallbacks, int32_t reset);
/* features [n_cells][125] and cell blender values [n_cells] (before wall override and diffusion) of
 * the current state: plain values unless a network is set (training data: use a plain handle) */
int plens_gpu_indicator_samples(plens_gpu* h, float* feats, double* alpha_raw);

#ifdef __cplusplus
}
#endif
#endif
