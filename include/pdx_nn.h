/*
 * pdx_nn.h -- C ABI of the BF16 MFMA kernels of the learned warm start
 * (simul_amd/csrc/pdx_nn.hip, linked into libpdx.so).  No reference
 * counterpart: the reference has no learned component on this path.
 *
 * All pointers are device pointers; matrices are row-major; BF16 values are
 * uint16_t bit patterns; `stream` is a hipStream_t (NULL = the null stream).
 * Calls only enqueue work.  Return 0 on success, 1 for invalid arguments, 3 for
 * a launch error.
 */
#ifndef PDX_NN_H
#define PDX_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Y = act(X W + b): X (M x K) BF16, W (K x N) BF16, b (N) FP32 or NULL, relu 0/1;
 * writes Ybf (BF16) and/or Yf (FP32), both with row stride ldy >= N */
int pdx_nn_linear_fwd(const uint16_t *X, const uint16_t *W, const float *bias, int M, int K, int N, int relu,
                      uint16_t *Ybf, float *Yf, int ldy, void *stream);
/* dX = (dY W^T) * [act > 0]: dY (M x N, row stride ldy) BF16, W (K x N) BF16,
 * act (M x K) BF16 activations of the layer input (NULL: no ReLU mask); dX (M x K) BF16 */
int pdx_nn_linear_dgrad(const uint16_t *dY, const uint16_t *W, int M, int K, int N, int ldy, const uint16_t *act,
                        uint16_t *dX, void *stream);
/* dW = X^T dY (K x N, FP32) and db = column sums of dY (N, FP32, may be NULL); ldy must equal N.
 * The reduction over M runs in slices of rows_per_slice rows (multiple of 8) summed in slice order;
 * part holds ceil(M / rows_per_slice) * K * N floats of scratch */
int pdx_nn_linear_wgrad(const uint16_t *X, const uint16_t *dY, int M, int K, int N, int ldy, int rows_per_slice,
                        float *part, float *dW, float *db, void *stream);
/* weighted mean-squared error of the first 3 columns of Y (M x ldy, FP32) against T (M x 3):
 * dY (M x ldy BF16) = 2 w_i (y - t) / wsum (zero past column 3); loss_part (ceil(M/256)) partial sums
 * of w_i |y - t|^2 / wsum; w may be NULL (all 1) */
int pdx_nn_mse(const float *Y, int ldy, const float *T, const float *w, int M, float wsum, uint16_t *dY,
               float *loss_part, void *stream);
/* one Adam step (decoupled from nothing: L2 weight decay added to the gradient) on FP32 parameters,
 * refreshing their BF16 copy pbf when not NULL; step counts from 1 */
int pdx_nn_adam(float *p, const float *g, float *m, float *v, uint16_t *pbf, int64_t n, float lr, float beta1,
                float beta2, float eps, int64_t step, float weight_decay, void *stream);
/* FP32 -> BF16, round to nearest even */
int pdx_nn_to_bf16(const float *x, int64_t n, uint16_t *y, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PDX_NN_H */
