/* simulato_surrogate.h — the layers of the HAADF surrogate network (libsimulato_hip.so).
 *
 * Build-only: the reference has no learned part (its detectors are computed, never predicted). These are the
 * network's forward, backward and update steps as hand-written HIP kernels, for callers that train or run the
 * surrogate (simulatokit_amd/surrogate.py). All pointers are device memory; `stream` is a hipStream_t (NULL = the
 * default stream); every call only enqueues work. Activations are float32 [batch][channels][h][w] with h and w
 * multiples of 16; weights are bfloat16 bit patterns [cout][cin][5][5]; padding is circular (the cells are periodic).
 * Every reduction runs in a fixed order. Returns 0 or a negative SK_ERR_* status (message: sk_hip_last_error). */
#ifndef SIMULATO_SURROGATE_H
#define SIMULATO_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = bias + conv5(x, wt), ReLU applied when relu != 0 */
int sk_conv5_forward(const float* x, int batch, int cin, int h, int w, const uint16_t* wt, const float* bias, int cout,
                     int relu, float* y, void* stream);

/* The weights in the tap-major layout of the matrix-core convolution, out[tap][32][kp] with kp = 16 when the
 * convolution it serves has at most 16 input channels, else 32: out[tap][co][ci] = wt[co][ci][tap] (flip = 0; kp from
 * cin), or, for the backward pass over the data, the flipped transposed kernel out[tap][ci][co] = wt[co][ci][24 - tap]
 * (flip = 1; kp from cout); the rows and columns past the real channels are zero. cin, cout <= 32; out holds
 * 25 * 32 * kp bfloat16 values (at most 25,600). */
int sk_conv5_weights_tap_major(const uint16_t* wt, int cin, int cout, int flip, uint16_t* out, void* stream);

/* y = bias + conv5(x, w) on the matrix cores (v_mfma_f32_32x32x16_bf16: the bfloat16 weights exact, each activation
 * as a bfloat16 hi + lo pair, float32 accumulation), ReLU applied when relu != 0; wtm = the tap-major weights
 * (sk_conv5_weights_tap_major with flip = 0; with flip = 1 and dy as x, cin and cout swapped, this is the backward
 * pass over the data). cin, cout <= 32, w a multiple of 32, h a multiple of 16. */
int sk_conv5_forward_tap_major(const float* x, int batch, int cin, int h, int w, const uint16_t* wtm,
                               const float* bias, int cout, int relu, float* y, void* stream);

/* The floats of the workspace sk_conv5_backward_weights_mfma takes: 512 blocks' partial sums of 25 x 32 x 32, and
 * 32 x 64 float64 partial sums of the bias gradient */
#define SK_CONV5_MFMA_WORKSPACE_FLOATS (512 * 25 * 32 * 32 + 2 * 32 * 64)

/* dw [cout][cin][5][5] and db [cout] (db may be NULL) as sk_conv5_backward_weights computes them, on the matrix
 * cores: both operands as bfloat16 hi + lo pairs, three products (a relative 2^-17 per product against float32's
 * 2^-24), float32 sums per block of pixels, the blocks' sums added in a fixed order in float64. cin, cout <= 32, w a
 * multiple of 32, h a multiple of 16; workspace holds SK_CONV5_MFMA_WORKSPACE_FLOATS floats. */
int sk_conv5_backward_weights_mfma(const float* x, const float* dy, int batch, int cin, int h, int w, int cout,
                                   float* dw, float* db, float* workspace, void* stream);

/* dx = the transpose of conv5 applied to dy; wt_scratch holds cin * cout * 25 bfloat16 values */
int sk_conv5_backward_data(const float* dy, int batch, int cin, int h, int w, const uint16_t* wt, int cout, float* dx,
                           uint16_t* wt_scratch, void* stream);

/* dw [cout][cin][5][5] and db [cout] (db may be NULL) of sum(dy * conv5(x, w) + dy * b) */
int sk_conv5_backward_weights(const float* x, const float* dy, int batch, int cin, int h, int w, int cout, float* dw,
                              float* db, void* stream);

/* dy *= (y > 0), y the output of a ReLU layer */
int sk_relu_backward(float* dy, const float* y, long long n, void* stream);

/* out[b][u][v] = sum_{i,j} A[u][i] A[v][j] in[b][i][j]; A m x n, in [batch][n][n], tmp [batch][n][m], out [batch][m][m]
 * (the output's resampling to the scan grid, and with A transposed its backward pass) */
int sk_separable_map(const float* in, int batch, int n, const float* A, int m, float* tmp, float* out, void* stream);

/* grad = 2 (pred - label) / n; partials[(n + 255) / 256] receive the squared error's per-block sums (float64) */
int sk_mse_grad(const float* pred, const float* label, long long n, float* grad, double* partials, void* stream);

/* as sk_mse_grad with a weight per image: element i belongs to image i / per (n a multiple of per), grad =
 * 2 weights[i / per] (pred - label) / n and partials the per-block sums of weights[i / per] (pred - label)^2 (the
 * training loss that weights each image by the inverse of its squared norm, so every image counts by its relative
 * error) */
int sk_weighted_mse_grad(const float* pred, const float* label, const float* weights, long long per, long long n,
                         float* grad, double* partials, void* stream);

/* one Adam step (bias-corrected, step >= 1) of the float32 masters p with gradient g and moments m, v; p_bf16 (may be
 * NULL) receives the bfloat16 rounding (to nearest, ties to even) of the updated p */
int sk_adam_step(float* p, const float* g, float* m, float* v, uint16_t* p_bf16, long long n, float lr, float beta1,
                 float beta2, float eps, int step, void* stream);

/* out = bfloat16 rounding (to nearest, ties to even) of in */
int sk_to_bf16(const float* in, uint16_t* out, long long n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SIMULATO_SURROGATE_H */
