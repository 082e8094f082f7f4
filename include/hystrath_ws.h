/*
 * hystrath_ws.h -- C ABI of the learned warm-start's network kernels.
 *
 * No counterpart in the reference (hyStrath has no learned models): a small
 * convolutional encoder-decoder that maps fields sampled over a short window
 * to the long-average fields, used to shorten the sampling window and to
 * initialise particle populations.  Every entry point takes device pointers
 * (the caller owns the memory) and a HIP stream (NULL: the default stream),
 * returns 0 or a negative HS_E* code, and enqueues its work asynchronously.
 *
 * Layout: activations are NHWC images of bfloat16 (uint16_t bit patterns),
 * n images of h x w pixels, channel counts c in {16, 32, 64} (a layer with
 * fewer true channels keeps the rest zero); w must be a multiple of 32.
 * Master weights are FP32 [9][cin][cout] (tap = ky*3 + kx, HWIO order) with
 * an FP32 bias[cout]; the kernels read bfloat16 packings of them.  Every
 * product is a bfloat16 MFMA on gfx950 accumulating in FP32.
 */
#ifndef HYSTRATH_WS_H
#define HYSTRATH_WS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Packs FP32 master weights [9][cin][cout] into the two bfloat16 layouts the
 * convolutions read: forward [9][cout_pad][cin] (cout_pad = cout rounded up
 * to 32) and transposed-flipped for the input gradient [9][cin_pad][cout]
 * (tap 8 - t, cin_pad = cin rounded up to 32); padding is zero. */
int hs_ws_pack(const float* w, int32_t cin, int32_t cout, uint16_t* w_fwd, uint16_t* w_bwd, void* stream);

/* y = act(conv3x3(x, w) + bias), taps dilation pixels apart (1: the plain
 * 3x3), zero padding dilation, stride 1.  x has cin
 * channels, y has cout channels (true cout <= cout_pad of the packing).
 * relu: apply max(0, .).  mask (may be NULL): an NHWC bf16 tensor with cout
 * channels; the result is multiplied by (mask > 0), which turns the input
 * gradient of the next layer into the gradient of this layer's ReLU input.
 * bias may be NULL.  y (bf16) and/or y32 (FP32, same layout) may be NULL. */
int hs_ws_conv3x3(const uint16_t* x, const uint16_t* w_packed, const float* bias, const uint16_t* mask,
                  uint16_t* y, float* y32, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                  int32_t relu, int32_t dilation, void* stream);

/* Weight and bias gradients of y = conv3x3(x, w) + b given dy = dL/dy (the
 * gradient at the convolution's output, ReLU already applied):
 * dw[9][cin][cout] += sum over pixels of x(p + dilation * tap) dy(p),
 * db[cout] += sum dy.  Accumulates into dw and db (clear them first). */
int hs_ws_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, int32_t n, int32_t h,
                        int32_t w, int32_t cin, int32_t cout, int32_t dilation, void* stream);

/* Masked mean-squared error of pred (FP32 NHWC, c channels, of which the
 * first nf count) against target (same layout), over pixels whose mask (FP32
 * [n][h][w]) is non-zero: loss += sum mask*(pred - target)^2 / (nf * count),
 * dy (bf16, c channels) = 2*mask*(pred - target)/(nf * count), zero in the
 * other channels.  count = the number of pixels with mask != 0 (given). */
int hs_ws_mse(const float* pred, const float* target, const float* mask, uint16_t* dy, float* loss,
              int32_t n, int32_t h, int32_t w, int32_t c, int32_t nf, double count, void* stream);

/* Adam update of n FP32 parameters (Kingma and Ba, 2015) with bias
 * correction at step t >= 1. */
int hs_ws_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
               float eps, int32_t t, void* stream);

/* Test hook: one MFMA tile on the device.  kind 0: 32x32x16, kind 1:
 * 16x16x32 (bfloat16 in, FP32 out).  a, b: host arrays of the operands as
 * each lane holds them (64 lanes x 8 bf16); d: host output as each lane holds
 * it (64 x 16 or 64 x 4 floats).  Synchronous. */
int hs_ws_mfma_tile(int32_t kind, const uint16_t* a, const uint16_t* b, float* d);

#ifdef __cplusplus
}
#endif
#endif /* HYSTRATH_WS_H */
