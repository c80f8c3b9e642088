/*
 * rootpwa_amd_mlp.h -- the learned initializer's network on the GPU: a
 * multilayer perceptron (hidden layers with the exact GELU, linear output),
 * trained in FP32 with hand-written kernels (MFMA products, Adam) on a mean
 * squared error.  It maps per-bin data-amplitude moments to start values of
 * the production amplitudes; every prediction is refined by a plain fit, and
 * random starts remain the fallback.  No counterpart in the reference, which
 * starts every fit from random values (highLevelInterface/pwaNloptFit.cc:192-203).
 *
 * Parameters: W_0, b_0, ..., W_L, b_L, float32, W_l row-major [in][out]
 * (y = x W + b).  Host arrays are row-major [rows][features].
 */
#ifndef ROOTPWA_AMD_MLP_H
#define ROOTPWA_AMD_MLP_H

#include <stdint.h>

#include "rootpwa_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rpwa_mlp rpwa_mlp;

/* Glorot-uniform weights from seed (splitmix64 counters), zero biases */
int rpwa_mlp_create(rpwa_mlp** m, int device, uint32_t nIn, uint32_t nHidden, uint32_t nHiddenLayers,
                    uint32_t nOut, uint64_t seed);
int rpwa_mlp_destroy(rpwa_mlp* m);
int rpwa_mlp_nmb_params(const rpwa_mlp* m, uint64_t* nParams);
int rpwa_mlp_get_params(const rpwa_mlp* m, float* params);
int rpwa_mlp_set_params(rpwa_mlp* m, const float* params);
/* y [nRows][nOut] = network(x [nRows][nIn]) */
int rpwa_mlp_forward(rpwa_mlp* m, const float* x, uint32_t nRows, float* y);
/* mean squared error over x / target and its gradient [nParams] (grad may be NULL) */
int rpwa_mlp_loss_grad(rpwa_mlp* m, const float* x, const float* target, uint32_t nRows, float* loss,
                       float* grad);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8, learning rate lr) over shuffled
 * mini-batches of batchSize rows for nEpochs epochs; the optimiser state
 * persists across calls.  lossPerEpoch [nEpochs] (or NULL): mean batch loss. */
int rpwa_mlp_train(rpwa_mlp* m, const float* x, const float* target, uint32_t nRows, uint32_t batchSize,
                   uint32_t nEpochs, float lr, uint64_t seed, float* lossPerEpoch);

#ifdef __cplusplus
}
#endif
#endif
