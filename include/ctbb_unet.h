/*
 * ctbb_unet.h -- C ABI of the learned warm start: a 2.5D U-Net that maps
 * three adjacent slices of an early ICD image to a start image for the
 * centre slice.
 *
 * No counterpart in the reference: it replaces the start image of
 * initialize_recon_volume (src/initialize_recon_volume.cpp:51-76,180-183)
 * with a learned one; the caller checks the result with the ICD cost and
 * falls back to the plain start when the learned one is worse.
 *
 * Network: levels L (default 4) with base*2^l channels at level l; per
 * level two 3x3 convolutions with bias and ReLU; 2x2 max pooling down,
 * nearest 2x upsampling and channel concatenation with the encoder output
 * up; a final 1x1 convolution to one channel added to the centre input
 * slice (residual).  Activations and convolution weights are BF16 in NHWC,
 * accumulation FP32 (MFMA), master weights and Adam state FP32.  Image
 * height and width must be multiples of 2^(L-1).
 *
 * Parameters, flat FP32, in layer order: for every 3x3 convolution
 * (encoder levels 0..L-1, conv a then b; then decoder levels L-2..0, conv a
 * then b) its weights [3][3][cout][cin] (tap-major, row dy then dx, cin
 * fastest) then its bias [cout]; last the 1x1 convolution weights [base]
 * and its bias [1].  The encoder's level-0 conv a has cin = 8 (3 slices and
 * 5 zero channels).
 */
#ifndef CTBB_UNET_H
#define CTBB_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctbb_unet ctbb_unet;

/* Network for batches of up to max_batch slices of height x width; weights
 * He-initialised from the seed (final 1x1 layer zero: the untrained network
 * returns its centre input). */
int ctbb_unet_create(int levels, int base_channels, int max_batch, int height,
                     int width, uint64_t seed, ctbb_unet **out);
int ctbb_unet_free(ctbb_unet *u);
int ctbb_unet_num_params(const ctbb_unet *u, int64_t *n);
int ctbb_unet_get_params(const ctbb_unet *u, float *params);
int ctbb_unet_set_params(ctbb_unet *u, const float *params);

/* Forward + backward of the mean squared error over a batch.
 * input [batch][3][height][width] (slices k-1, k, k+1), target
 * [batch][height][width], FP32 host arrays.  loss: mean over pixels of
 * (output - target)^2.  grads (optional, [num_params]): gradients of the
 * loss, same order as the parameters.  The gradients are kept on the device
 * for the next ctbb_unet_adam_step. */
int ctbb_unet_grad(ctbb_unet *u, const float *input, const float *target,
                   int batch, double *loss, float *grads);

/* One Adam update (bias-corrected) with the gradients of the last
 * ctbb_unet_grad. */
int ctbb_unet_adam_step(ctbb_unet *u, double lr, double beta1, double beta2,
                        double eps);

/* output [batch][height][width] for input [batch][3][height][width]. */
int ctbb_unet_infer(ctbb_unet *u, const float *input, int batch, float *output);

/* Average device time (ms) of the last ctbb_unet_grad / ctbb_unet_infer and
 * the MFMA FLOPs it executed (3x3 convolutions only). */
int ctbb_unet_last_timing(const ctbb_unet *u, double *ms, double *conv_flops);

#ifdef __cplusplus
}
#endif
#endif
