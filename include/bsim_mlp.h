/*
 * bsim_mlp.h -- C ABI of the rate surrogate in libbsim_amd.so: a ReLU
 * multilayer perceptron trained and evaluated on the GPU with bf16 MFMA
 * kernels (fp32 accumulation, fp32 master weights, Adam).
 *
 * No entry point here has a counterpart in the reference (CRAFT-THU/BSim has
 * no learned component); they serve sweep screening: the network is trained
 * on per-population rate trajectories recorded by the simulator and predicts
 * them for candidate sweep points, which the simulator then re-checks.
 *
 * Parameters are exchanged as one flat fp32 array in the padded layout: for
 * each linear layer l (0 = input layer), the weight matrix
 * W_l[dims[l + 1]][dims[l]] row-major, then the bias b_l[dims[l + 1]], with
 * dims[0] = 64 (inputs zero-padded), dims[1..L] = hidden, dims[L + 1] =
 * out_dim rounded up to 64 (bsim_mlp_layout).  Padded entries are zero.
 *
 * Error convention as in bsim_amd.h: 0 on success, a negative BSIM_E_* code
 * otherwise, with the message in bsim_last_error().  One handle is not
 * thread-safe; several handles may coexist.
 */
#ifndef BSIM_MLP_H
#define BSIM_MLP_H

#include <stdint.h>

#include "bsim_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bsim_mlp bsim_mlp;

typedef struct bsim_mlp_config {
    int32_t in_dim;      /* input features, 1..64 */
    int32_t hidden;      /* units per hidden layer, a multiple of 64, at most 4096 */
    int32_t layers;      /* hidden layers, 1..9 */
    int32_t out_dim;     /* outputs, 1..4096 */
    int32_t max_batch;   /* largest batch, a multiple of 256 */
    int32_t device;      /* HIP device */
    uint64_t seed;       /* He-uniform weight initialisation, zero biases */
    float beta1, beta2, eps;   /* Adam */
    float pad;
} bsim_mlp_config;

/* Allocates the model, its Adam state and workspaces for max_batch rows on
 * the device.  NULL on failure (bsim_last_error). */
bsim_mlp *bsim_mlp_create(const bsim_mlp_config *cfg);
void bsim_mlp_destroy(bsim_mlp *m);

/* Number of floats in the flat parameter array. */
int64_t bsim_mlp_num_params(const bsim_mlp *m);
/* Padded layer widths dims[0..L+1]; returns L + 1 (the number of linear
 * layers). */
int bsim_mlp_layout(const bsim_mlp *m, int32_t *dims, int32_t cap);
/* fp32 master parameters out / in.  Setting also writes the bf16 copy and
 * resets the Adam moments and step count. */
int bsim_mlp_get_params(bsim_mlp *m, float *out);
int bsim_mlp_set_params(bsim_mlp *m, const float *in);

/* Uploads a training set: x[n][in_dim], y[n][out_dim] (y kept in fp32, x
 * stored as bf16 on the device). */
int bsim_mlp_set_data(bsim_mlp *m, const float *x, const float *y, int64_t n);
/* steps Adam steps on minibatches of the uploaded set: rows[s * batch + i]
 * is the i-th row of step s.  Loss: mean squared error over the batch and
 * out_dim outputs; loss[s] receives step s's loss (before its update) when
 * loss is not NULL.  batch: a multiple of 256, at most max_batch. */
int bsim_mlp_train(bsim_mlp *m, const int32_t *rows, int32_t steps, int32_t batch, float lr,
                   float *loss);
/* bsim_mlp_train for callers that cannot allocate device memory (code under
 * HIP graph capture, tight training or serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, kernels and
 * results as bsim_mlp_train, but the per-call staging (the minibatch row
 * indices and the per-step losses) lives in a fixed device array of
 * BSIM_MLP_NOALLOC_WORDS 32-bit words per GPU, allocated with the library
 * and held until the process ends, instead of a hipMalloc per call:
 * round_up(steps, 64) + steps * batch must not exceed it (255 steps at batch
 * 4,096), or BSIM_E_ARG is returned.  Calls on one GPU share the array:
 * calls that overlap on one GPU (from several threads or handles) race. */
#define BSIM_MLP_NOALLOC_WORDS (1 << 20)
int bsim_mlp_train_noalloc(bsim_mlp *m, const int32_t *rows, int32_t steps, int32_t batch,
                           float lr, float *loss);

/* Gradient of the loss for one batch x[batch][in_dim], y[batch][out_dim],
 * in the flat parameter layout, without updating anything. */
int bsim_mlp_gradients(bsim_mlp *m, const float *x, const float *y, int32_t batch, float *grad,
                       float *loss);
/* Outputs y[n][out_dim] for inputs x[n][in_dim], any n. */
int bsim_mlp_predict(bsim_mlp *m, const float *x, int64_t n, float *y);

#ifdef __cplusplus
}
#endif

#endif
