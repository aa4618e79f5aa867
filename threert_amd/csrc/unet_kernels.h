// Launch interface of the learned denoiser's kernels (unet.hip): a U-Net of 3x3 convolutions on BF16 MFMA
// (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation). Activations are NHWC bfloat16 with a channel count that is
// a multiple of 32; weights are [cout][9 taps][cin] bfloat16 (tap = 3 * (dy + 1) + (dx + 1)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace trt {

// out = act(conv3x3(concat(in0, up2(in1))) + bias); up2 = nearest-neighbour upsampling of a half-size image.
// The same kernel computes the gradient of a convolution's input: in0 = dY, wt = the flipped, transposed weights.
struct ConvArgs {
    const uint16_t *in0;  // (n, h, w, c0)
    const uint16_t *in1;  // (n, h/2, w/2, c1), or null with c1 = 0
    const uint16_t *wt;   // [cout][9][c0 + c1]
    const float *bias;    // [cout] or null
    uint16_t *out_bf16;   // (n, h, w, cout), or null
    float *out_f32;       // (n, h, w, cout), or null
    uint32_t c0, c1, cout, n, h, w, relu;
    // Inference's last layer (cin = 32): with out_color set, the only output is k_output's colour,
    // expm1(max(y, 0)) of channels 0-2, into this (n, h, w, 3) image.
    float *out_color = nullptr;
};
int launch_conv(const ConvArgs &a, hipStream_t s);
// Levels at least `rows` tall take 16-pixel-tall convolution tiles (default 128; UINT32_MAX: none).
void set_conv_tall_min(uint32_t rows);

// dw[co][tap][ci] += sum over pixels of dy[p][co] * x[p + tap][ci], x = concat(in0, up2(in1)) zero-padded;
// db[co] += sum over pixels of dy[p][co]. FP32 atomics into dw / db (zero them first).
struct WgradArgs {
    const uint16_t *in0;
    const uint16_t *in1;
    const uint16_t *dy;  // (n, h, w, cout)
    float *dw;
    float *db;
    uint32_t c0, c1, cout, n, h, w;
};
int launch_wgrad(const WgradArgs &a, hipStream_t s);

// 2x2 average pooling: (n, h, w, c) -> (n, h/2, w/2, c).
int launch_pool(const uint16_t *in, uint16_t *out, uint32_t n, uint32_t h, uint32_t w, uint32_t c, hipStream_t s);

// Backward glue. dx: a convolution's input gradient, FP32 (n, h, w, ctot); the part [coff, coff + c) is taken.
// out = (act > 0) * dx[.., coff + k]                                        (act: the ReLU output it flows into)
int launch_mask(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out, uint32_t n,
                uint32_t h, uint32_t w, uint32_t c, hipStream_t s);
// out(half size) = (act > 0) * sum of the 2x2 block of dx[.., coff + k]   (the gradient through up2)
int launch_mask_pool_sum(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out,
                         uint32_t n, uint32_t h2, uint32_t w2, uint32_t c, hipStream_t s);
// out = (act > 0) * (skip[.., k] + dp[y/2, x/2, k] / 4)                    (skip: ctot-wide, dp: c-wide, half size)
int launch_mask_skip_up(const float *skip, uint32_t ctot, const float *dp, const uint16_t *act, uint16_t *out,
                        uint32_t n, uint32_t h, uint32_t w, uint32_t c, hipStream_t s);

// Features: 32 channels per pixel: log(1 + max(color, 0)), albedo, normal * 0.5 + 0.5, then zeros.
int launch_features(const float *color, const float *albedo, const float *normal, uint16_t *out, uint64_t pixels,
                    hipStream_t s);
// Output: color = exp(max(y[0..2], 0)) - 1 from the last layer's FP32 output (32 channels).
int launch_output(const float *y, float *color, uint64_t pixels, hipStream_t s);
// L1 loss in log space: loss += sum |y[c] - log(1 + max(target[c], 0))| / (3 pixels) over c < 3; dy = its
// gradient (sign / (3 pixels)) in bfloat16, zero in channels 3..31.
int launch_loss(const float *y, const float *target, uint16_t *dy, float *loss, uint64_t pixels, hipStream_t s);

// Adam over the FP32 master parameters (bias-corrected, step t counted from 1).
int launch_adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr, float b1, float b2, float eps,
                uint32_t t, hipStream_t s);
// Weights of one layer to bfloat16: wf[co][tap][ci] = w, and the input-gradient layout wb[ci][8 - tap][co] = w.
int launch_cast_weights(const float *w, uint16_t *wf, uint16_t *wb, uint32_t cout, uint32_t cin, hipStream_t s);

}  // namespace trt
