/*
 * dem_surrogate.h -- MLP surrogate of the packing outcome of a sweep point.
 *
 * No counterpart in the reference (which has no parameter sweeps).  The
 * network maps the 8 sweep parameters of a packing (normalised by the
 * caller) to 2 outcomes (phi and Z, normalised by the caller):
 *   8 -> 128 -> 128 -> 128 -> 2, SiLU after the three hidden layers,
 *   34,434 parameters.
 * Weights are kept in FP32 (master copy, updated by Adam) and used as BF16
 * by the layer products, which run on the MFMA units with FP32
 * accumulation; activations enter each product as BF16.
 *
 * Parameter layout (FP32, sur_nparams() values): W1[128][8], b1[128],
 * W2[128][128], b2[128], W3[128][128], b3[128], W4[2][128], b4[2]; W[o][i]
 * multiplies input i of output o.
 *
 * All array arguments are host pointers; the calls copy them to and from
 * the device.  Every call returns DEM_OK (0) or a negative DEM_ERR_* code
 * with the message in dem_last_error().
 */
#ifndef DEM_SURROGATE_H
#define DEM_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

#define SUR_IN 8
#define SUR_HIDDEN 128
#define SUR_OUT 2

typedef struct sur_mlp sur_mlp;

int sur_nparams(void);
/* Glorot-uniform weights from a seeded generator, zero biases. */
int sur_create(unsigned long long seed, sur_mlp **out);
int sur_destroy(sur_mlp *m);
int sur_get_params(sur_mlp *m, float *params);
int sur_set_params(sur_mlp *m, const float *params);
/* y[n][2] = f(x[n][8]) with the BF16 weights. */
int sur_forward(sur_mlp *m, int n, const float *x, float *y);

/* Allocation-free sur_forward, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops of a sweep's
 * screening).  No counterpart in the reference.  x[n][8] and y[n][2] are
 * DEVICE pointers on the network's GPU; y equals sur_forward's output bit
 * for bit.  At most SUR_NOALLOC_MAX_ROWS rows per call (larger n gives
 * DEM_ERR_ARG).  The activations live in a fixed array of the library in
 * device memory (4.3 MB per GPU, held until the process ends) that every
 * call on that GPU shares: calls that overlap on one GPU race.  Runs on the
 * network's stream and returns once y is written. */
#define SUR_NOALLOC_MAX_ROWS 8192
int sur_forward_noalloc(sur_mlp *m, int n, const float *x, float *y);
/* Mean squared error over the n x 2 outputs against t[n][2], and its
 * gradient with respect to the FP32 parameters (grad may be NULL). */
int sur_loss_grad(sur_mlp *m, int n, const float *x, const float *t, float *loss, float *grad);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of size batch,
 * shuffled each epoch by a seeded generator; *loss = mean squared error
 * over all n points after the last epoch. */
int sur_train(sur_mlp *m, int n, const float *x, const float *t, int epochs, int batch, float lr,
              unsigned long long seed, float *loss);

#ifdef __cplusplus
}
#endif
#endif
