/*
 * flysim_mlp.h -- C ABI of the rate surrogate: a small fully connected
 * network (GELU after every layer but the last) trained and evaluated on the
 * GPU with BF16 matrix-core GEMMs (FP32 accumulation, FP32 master weights,
 * Adam).  It maps sweep parameters (e.g. per-module external rates and
 * receptor gains) to per-module steady-state rates, learned from runs of the
 * simulator itself, and screens sweep points before the simulator re-runs the
 * flagged ones.  The reference has no counterpart (it runs one network per
 * process, code/main.cpp:67-264); callers are sweep drivers.
 *
 * Status codes and fs_last_error() as in flysim.h.  Parameters are exchanged
 * as one flat f32 vector: for each layer i, W_i[out][in] row-major, then
 * b_i[out].
 */
#ifndef FLYSIM_MLP_H
#define FLYSIM_MLP_H
#include <stdint.h>

#include "flysim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs_mlp fs_mlp;

/* sizes[0..n_layers]: widths, e.g. {54, 256, 256, 256, 50}; weights He-uniform
 * from seed, biases zero. */
int fs_mlp_create(const int32_t *sizes, int32_t n_layers, uint64_t seed, fs_mlp **out);
int fs_mlp_destroy(fs_mlp *m);
int fs_mlp_n_params(fs_mlp *m, int64_t *n);
int fs_mlp_get_params(fs_mlp *m, float *out);
/* Also resets the Adam moments and step count. */
int fs_mlp_set_params(fs_mlp *m, const float *in);

/* Mean-squared-error loss of rows x (f32 [rows][in]) against y (f32
 * [rows][out]) and its gradient with respect to the flat parameters (host
 * pointers; grad may be NULL).  No parameter update. */
int fs_mlp_loss_grad(fs_mlp *m, const float *x, const float *y, int32_t rows, float *loss, float *grad);

/* y[n][out] = network(x[n][in]); device_ptrs = 1: x and y are device
 * pointers (inputs resident in HBM).  *ms (optional) = GPU time of the
 * forward kernels measured with HIP events.  Networks whose padded widths
 * (multiples of 64) are all <= 256, at most 8 layers, run one fused kernel
 * for every layer; others one GEMM per layer. */
int fs_mlp_predict(fs_mlp *m, const float *x, int64_t n, float *y, int32_t device_ptrs, double *ms);

/* Allocation-free forward.  No counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  x (f32 [n][in]) and y
 * (f32 [n][out]) are device pointers.  Networks that fs_mlp_predict runs
 * with its fused kernel run that kernel here too, which keeps every
 * intermediate in registers and LDS; for the others the packed input and the
 * hidden activations of each chunk of rows live in one fixed 8 MiB device
 * array per GPU (resolved at fs_mlp_create), so the call allocates nothing.
 * Any n is processed in chunks of that array's size (a multiple of 64 rows;
 * the widths fs_mlp_create admits always fit one chunk).  Output equals
 * fs_mlp_predict's bit for bit.  Every call on one GPU shares the array and
 * leaves its contents behind: calls that overlap on one GPU (other threads or
 * other handles) race.  Returns after the GPU has finished. */
int fs_mlp_predict_noalloc(fs_mlp *m, const float *x, int64_t n, float *y);

/* Training set copied to the device (f32 [n][in], f32 [n][out]). */
int fs_mlp_set_data(fs_mlp *m, const float *x, const float *y, int64_t n);

typedef struct fs_adam {
  float lr, beta1, beta2, eps, weight_decay;
} fs_adam;

/* Minibatch Adam over the resident training set: each epoch draws a
 * permutation (shuffle = 0: identity order) and runs floor(n / batch) steps,
 * one captured graph per step (gather, forward, loss, backward, Adam).
 * epoch_loss[epochs] (optional) = mean minibatch loss of each epoch. */
int fs_mlp_train(fs_mlp *m, int32_t epochs, int32_t batch, const fs_adam *opt, uint64_t shuffle_seed,
                 int32_t shuffle, float *epoch_loss);

/* Device buffers for callers that keep inputs resident in HBM (e.g. the
 * sweep points of fs_mlp_predict with device_ptrs = 1).  The library's HIP
 * runtime owns them; dir: 0 host->device, 1 device->host, 2 device->device. */
int fs_device_alloc(int64_t bytes, void **ptr);
int fs_device_free(void *ptr);
int fs_device_copy(void *dst, const void *src, int64_t bytes, int32_t dir);

#ifdef __cplusplus
}
#endif
#endif
