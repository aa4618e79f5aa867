// The learned denoiser (the reference's denoise::oidn stage, denoise.rs:81-103, replaced by a network of this
// library's own): a five-level U-Net of 3x3 convolutions, 32 to 128 channels, trained and run on the GPU by the
// kernels of unet.hip. Host side: the layer table, the device buffers, the forward pass, the backward pass and Adam.
//
// Layers (cin = c0 + c1, c1 the half-size input upsampled and concatenated; every convolution but the last has a
// ReLU; the last one's 32 outputs are FP32 and only the first three are used):
//   level 0  e0a  features 32 -> 32   e0b 32 -> 32                         d0a e0b 32 + d1b 64 -> 32   d0b 32 -> 32
//   level 1  e1a  pool(e0b) 32 -> 64  e1b 64 -> 64                         d1a e1b 64 + d2b 64 -> 64   d1b 64 -> 64
//   level 2  e2a  pool(e1b) 64 -> 64  e2b 64 -> 64                         d2a e2b 64 + d3b 96 -> 64   d2b 64 -> 64
//   level 3  e3a  pool(e2b) 64 -> 96  e3b 96 -> 96                         d3a e3b 96 + e4b 128 -> 96  d3b 96 -> 96
//   level 4  e4a  pool(e3b) 96 -> 128 e4b 128 -> 128
//   out      d0b 32 -> 32 (linear): log(1 + color) in its first three channels.
// Parameters (FP32, the order of the table): each layer's weights [cout][9][cin], then its biases [cout].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <new>
#include <string>
#include <vector>

#include "../../include/threert.h"
#include "unet_kernels.h"

// The device's stream, mutex and error reporting live in api.cpp; these are its small accessors.
hipStream_t trt_device_stream(trt_device *d);
int trt_device_ordinal(trt_device *d);
void trt_device_lock(trt_device *d, bool lock);
void trt_device_ref(trt_device *d, int delta);
int trt_fail(int code, const std::string &msg);
int trt_ok();

namespace {

enum Layer { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3A, D3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT, NL };

struct Spec {
    uint32_t c0, c1, cout, level;
};
constexpr Spec kSpec[NL] = {
    {32, 0, 32, 0},  {32, 0, 32, 0},   {32, 0, 64, 1},  {64, 0, 64, 1},  {64, 0, 64, 2},  {64, 0, 64, 2},
    {64, 0, 96, 3},  {96, 0, 96, 3},   {96, 0, 128, 4}, {128, 0, 128, 4}, {96, 128, 96, 3}, {96, 0, 96, 3},
    {64, 96, 64, 2}, {64, 0, 64, 2},   {64, 64, 64, 1}, {64, 0, 64, 1},  {32, 64, 32, 0}, {32, 0, 32, 0},
    {32, 0, 32, 0},
};

#define UNET_HIP(call)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return trt_fail(e_ == hipErrorOutOfMemory ? TRT_E_OOM : TRT_E_DEVICE,                      \
                            std::string(#call) + ": " + hipGetErrorString(e_));                        \
    } while (0)
#define UNET_LAUNCH(call)                                                                              \
    do {                                                                                               \
        if ((call) != 0) return trt_fail(TRT_E_DEVICE, std::string("unet: launch failed: ") + #call); \
    } while (0)

}  // namespace

struct trt_unet {
    trt_device *dev = nullptr;
    uint64_t cap = 0;  // pixels at level 0 (batch x height x width)
    uint64_t nparam = 0, nweight_bf16 = 0;
    uint64_t woff[NL] = {}, boff[NL] = {}, hoff[NL] = {};  // FP32 weight / bias offsets, bf16 weight offsets
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    uint32_t step = 0;
    uint16_t *feat = nullptr, *act[NL] = {}, *pool[5] = {};
    float *y = nullptr;
    float *keep[4] = {};  // input gradients of d0a, d1a, d2a, d3a (their skip parts feed the encoder)
    float *dxt = nullptr; // other layers' input gradients
    uint16_t *g[2] = {};  // bf16 gradients of the current layer's output (ping-pong)
    float *loss = nullptr;
    std::vector<void *> owned;
};

namespace {

uint64_t level_pixels(uint64_t p0, uint32_t level) { return p0 >> (2 * level); }

int alloc(trt_unet *u, void **p, uint64_t bytes) {
    UNET_HIP(hipMalloc(p, bytes ? bytes : 256));
    u->owned.push_back(*p);
    return TRT_OK;
}

// Shapes of one call: n images of h x w, h and w multiples of 16, within the capacity.
int check_shape(trt_unet *u, uint32_t n, uint32_t h, uint32_t w) {
    if (n == 0 || h == 0 || w == 0 || h % 16 || w % 16)
        return trt_fail(TRT_E_INVALID, "unet: sizes must be positive and height and width multiples of 16");
    if (uint64_t(n) * h * w > u->cap) return trt_fail(TRT_E_INVALID, "unet: more pixels than the network was made for");
    return TRT_OK;
}

int cast_all(trt_unet *u, hipStream_t s) {
    for (int l = 0; l < NL; l++) {
        const Spec &sp = kSpec[l];
        UNET_LAUNCH(trt::launch_cast_weights(u->params + u->woff[l], u->wf + u->hoff[l], u->wb + u->hoff[l], sp.cout,
                                             sp.c0 + sp.c1, s));
    }
    return TRT_OK;
}

int conv(trt_unet *u, int l, const uint16_t *in0, const uint16_t *in1, uint16_t *out_bf16, float *out_f32,
         uint32_t n, uint32_t h0, uint32_t w0, hipStream_t s) {
    const Spec &sp = kSpec[l];
    trt::ConvArgs a{};
    a.in0 = in0, a.in1 = in1, a.wt = u->wf + u->hoff[l], a.bias = u->params + u->boff[l];
    a.out_bf16 = out_bf16, a.out_f32 = out_f32;
    a.c0 = sp.c0, a.c1 = sp.c1, a.cout = sp.cout, a.n = n, a.h = h0 >> sp.level, a.w = w0 >> sp.level;
    a.relu = l != OUT;
    UNET_LAUNCH(trt::launch_conv(a, s));
    return TRT_OK;
}

// The input gradient of layer l (FP32, all cin channels) from its output gradient dy (bf16).
int conv_back(trt_unet *u, int l, const uint16_t *dy, float *dx, uint32_t n, uint32_t h0, uint32_t w0,
              hipStream_t s) {
    const Spec &sp = kSpec[l];
    trt::ConvArgs a{};
    a.in0 = dy, a.in1 = nullptr, a.wt = u->wb + u->hoff[l], a.bias = nullptr;
    a.out_bf16 = nullptr, a.out_f32 = dx;
    a.c0 = sp.cout, a.c1 = 0, a.cout = sp.c0 + sp.c1, a.n = n, a.h = h0 >> sp.level, a.w = w0 >> sp.level;
    a.relu = 0;
    UNET_LAUNCH(trt::launch_conv(a, s));
    return TRT_OK;
}

int wgrad(trt_unet *u, int l, const uint16_t *in0, const uint16_t *in1, const uint16_t *dy, uint32_t n, uint32_t h0,
          uint32_t w0, hipStream_t s) {
    const Spec &sp = kSpec[l];
    trt::WgradArgs a{};
    a.in0 = in0, a.in1 = in1, a.dy = dy, a.dw = u->grads + u->woff[l], a.db = u->grads + u->boff[l];
    a.c0 = sp.c0, a.c1 = sp.c1, a.cout = sp.cout, a.n = n, a.h = h0 >> sp.level, a.w = w0 >> sp.level;
    UNET_LAUNCH(trt::launch_wgrad(a, s));
    return TRT_OK;
}

// out_color null: training's forward (the FP32 output kept for the loss); else inference, the last convolution
// writing the colour straight to out_color.
int forward(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, const float *color, const float *albedo,
            const float *normal, hipStream_t s, float *out_color = nullptr) {
    UNET_LAUNCH(trt::launch_features(color, albedo, normal, u->feat, uint64_t(n) * h * w, s));
    int rc = conv(u, E0A, u->feat, nullptr, u->act[E0A], nullptr, n, h, w, s);
    if (!rc) rc = conv(u, E0B, u->act[E0A], nullptr, u->act[E0B], nullptr, n, h, w, s);
    const int enc_a[4] = {E1A, E2A, E3A, E4A}, enc_b[4] = {E1B, E2B, E3B, E4B}, prev_b[4] = {E0B, E1B, E2B, E3B};
    for (int k = 0; k < 4 && !rc; k++) {
        const Spec &sp = kSpec[prev_b[k]];
        UNET_LAUNCH(trt::launch_pool(u->act[prev_b[k]], u->pool[k + 1], n, h >> sp.level, w >> sp.level, sp.cout, s));
        rc = conv(u, enc_a[k], u->pool[k + 1], nullptr, u->act[enc_a[k]], nullptr, n, h, w, s);
        if (!rc) rc = conv(u, enc_b[k], u->act[enc_a[k]], nullptr, u->act[enc_b[k]], nullptr, n, h, w, s);
    }
    const int dec_a[4] = {D3A, D2A, D1A, D0A}, dec_b[4] = {D3B, D2B, D1B, D0B};
    const int skip[4] = {E3B, E2B, E1B, E0B}, below[4] = {E4B, D3B, D2B, D1B};
    for (int k = 0; k < 4 && !rc; k++) {
        rc = conv(u, dec_a[k], u->act[skip[k]], u->act[below[k]], u->act[dec_a[k]], nullptr, n, h, w, s);
        if (!rc) rc = conv(u, dec_b[k], u->act[dec_a[k]], nullptr, u->act[dec_b[k]], nullptr, n, h, w, s);
    }
    if (!rc && out_color) {
        const Spec &sp = kSpec[OUT];
        trt::ConvArgs a{};
        a.in0 = u->act[D0B], a.in1 = nullptr, a.wt = u->wf + u->hoff[OUT], a.bias = u->params + u->boff[OUT];
        a.out_bf16 = nullptr, a.out_f32 = nullptr, a.out_color = out_color;
        a.c0 = sp.c0, a.c1 = 0, a.cout = sp.cout, a.n = n, a.h = h, a.w = w, a.relu = 0;
        UNET_LAUNCH(trt::launch_conv(a, s));
    } else if (!rc) {
        rc = conv(u, OUT, u->act[D0B], nullptr, nullptr, u->y, n, h, w, s);
    }
    return rc;
}

// Gradients of the loss (already in g[0] as dL/dy) into u->grads, which must be zeroed.
int backward(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, hipStream_t s) {
    uint16_t *cur = u->g[0], *nxt = u->g[1];
    auto lv = [&](int l) { return kSpec[l].level; };
    auto dims = [&](int l, uint32_t &hh, uint32_t &ww) { hh = h >> lv(l), ww = w >> lv(l); };
    uint32_t hh, ww;
    // out, then the decoder from level 0 up
    int rc = wgrad(u, OUT, u->act[D0B], nullptr, cur, n, h, w, s);
    if (!rc) rc = conv_back(u, OUT, cur, u->dxt, n, h, w, s);
    if (rc) return rc;
    dims(D0B, hh, ww);
    UNET_LAUNCH(trt::launch_mask(u->dxt, 32, 0, u->act[D0B], nxt, n, hh, ww, 32, s));
    std::swap(cur, nxt);
    const int dec_b[4] = {D0B, D1B, D2B, D3B}, dec_a[4] = {D0A, D1A, D2A, D3A};
    const int skip[4] = {E0B, E1B, E2B, E3B}, below[4] = {D1B, D2B, D3B, E4B};
    for (int k = 0; k < 4; k++) {
        const int b = dec_b[k], a = dec_a[k];
        // cur = dL/d(act b)
        rc = wgrad(u, b, u->act[a], nullptr, cur, n, h, w, s);
        if (!rc) rc = conv_back(u, b, cur, u->dxt, n, h, w, s);
        if (rc) return rc;
        dims(a, hh, ww);
        UNET_LAUNCH(trt::launch_mask(u->dxt, kSpec[b].c0, 0, u->act[a], nxt, n, hh, ww, kSpec[a].cout, s));
        std::swap(cur, nxt);
        // cur = dL/d(act a); a reads skip (c0) and below (c1, half size)
        rc = wgrad(u, a, u->act[skip[k]], u->act[below[k]], cur, n, h, w, s);
        if (!rc) rc = conv_back(u, a, cur, u->keep[k], n, h, w, s);
        if (rc) return rc;
        const uint32_t ctot = kSpec[a].c0 + kSpec[a].c1;
        UNET_LAUNCH(trt::launch_mask_pool_sum(u->keep[k], ctot, kSpec[a].c0, u->act[below[k]], nxt, n, hh / 2, ww / 2,
                                              kSpec[a].c1, s));
        std::swap(cur, nxt);
    }
    // cur = dL/d(act e4b); the encoder from level 4 down
    const int enc_b[5] = {E0B, E1B, E2B, E3B, E4B}, enc_a[5] = {E0A, E1A, E2A, E3A, E4A};
    for (int k = 4; k >= 0; k--) {
        const int b = enc_b[k], a = enc_a[k];
        rc = wgrad(u, b, u->act[a], nullptr, cur, n, h, w, s);
        if (!rc) rc = conv_back(u, b, cur, u->dxt, n, h, w, s);
        if (rc) return rc;
        dims(a, hh, ww);
        UNET_LAUNCH(trt::launch_mask(u->dxt, kSpec[b].c0, 0, u->act[a], nxt, n, hh, ww, kSpec[a].cout, s));
        std::swap(cur, nxt);
        // cur = dL/d(act a); a reads the features (level 0) or the pool of the level above
        const uint16_t *in = k == 0 ? u->feat : u->pool[k];
        rc = wgrad(u, a, in, nullptr, cur, n, h, w, s);
        if (rc) return rc;
        if (k == 0) break;
        rc = conv_back(u, a, cur, u->dxt, n, h, w, s);
        if (rc) return rc;
        // the level above's b output feeds the pool (dxt, half size) and its decoder's a layer (keep[k - 1])
        const int up = enc_b[k - 1];
        uint32_t h1, w1;
        dims(up, h1, w1);
        UNET_LAUNCH(trt::launch_mask_skip_up(u->keep[k - 1], kSpec[dec_a[k - 1]].c0 + kSpec[dec_a[k - 1]].c1, u->dxt,
                                             u->act[up], nxt, n, h1, w1, kSpec[up].cout, s));
        std::swap(cur, nxt);
    }
    return TRT_OK;
}

}  // namespace

extern "C" {

int trt_unet_create(trt_device *dev, uint64_t max_pixels, trt_unet **out) {
    if (!dev || !out || max_pixels == 0) return trt_fail(TRT_E_INVALID, "trt_unet_create: bad argument");
    if (max_pixels >= (uint64_t{1} << 31)) return trt_fail(TRT_E_INVALID, "trt_unet_create: too many pixels");
    *out = nullptr;
    auto *u = new (std::nothrow) trt_unet();
    if (!u) return trt_fail(TRT_E_OOM, "trt_unet_create: out of host memory");
    u->dev = dev;
    u->cap = (max_pixels + 255) / 256 * 256;
    uint64_t p = 0, hw = 0;
    for (int l = 0; l < NL; l++) {
        const Spec &sp = kSpec[l];
        const uint64_t nw = uint64_t(sp.cout) * 9 * (sp.c0 + sp.c1);
        u->woff[l] = p, p += nw;
        u->boff[l] = p, p += sp.cout;
        u->hoff[l] = hw, hw += nw;
    }
    u->nparam = p, u->nweight_bf16 = hw;
    UNET_HIP(hipSetDevice(trt_device_ordinal(dev)));
    int rc = TRT_OK;
    auto A = [&](void **q, uint64_t bytes) {
        if (!rc) rc = alloc(u, q, bytes);
    };
    A(reinterpret_cast<void **>(&u->params), p * 4);
    A(reinterpret_cast<void **>(&u->grads), p * 4);
    A(reinterpret_cast<void **>(&u->m), p * 4);
    A(reinterpret_cast<void **>(&u->v), p * 4);
    A(reinterpret_cast<void **>(&u->wf), hw * 2);
    A(reinterpret_cast<void **>(&u->wb), hw * 2);
    const uint64_t c = u->cap;
    A(reinterpret_cast<void **>(&u->feat), c * 32 * 2);
    for (int l = 0; l < OUT; l++) A(reinterpret_cast<void **>(&u->act[l]), level_pixels(c, kSpec[l].level) * kSpec[l].cout * 2);
    const uint32_t pool_c[5] = {0, 32, 64, 64, 96};
    for (int k = 1; k <= 4; k++) A(reinterpret_cast<void **>(&u->pool[k]), level_pixels(c, k) * pool_c[k] * 2);
    A(reinterpret_cast<void **>(&u->y), c * 32 * 4);
    const int dec_a[4] = {D0A, D1A, D2A, D3A};
    uint64_t dxt = 0, gmax = 0;
    for (int k = 0; k < 4; k++) {
        const Spec &sp = kSpec[dec_a[k]];
        A(reinterpret_cast<void **>(&u->keep[k]), level_pixels(c, sp.level) * (sp.c0 + sp.c1) * 4);
    }
    for (int l = 0; l < NL; l++) {
        const Spec &sp = kSpec[l];
        dxt = std::max(dxt, level_pixels(c, sp.level) * (sp.c0 + sp.c1));
        gmax = std::max(gmax, level_pixels(c, sp.level) * sp.cout);
    }
    A(reinterpret_cast<void **>(&u->dxt), dxt * 4);
    A(reinterpret_cast<void **>(&u->g[0]), gmax * 2);
    A(reinterpret_cast<void **>(&u->g[1]), gmax * 2);
    A(reinterpret_cast<void **>(&u->loss), 256);
    if (!rc) {
        hipError_t e = hipMemset(u->params, 0, p * 4);
        if (e == hipSuccess) e = hipMemset(u->m, 0, p * 4);
        if (e == hipSuccess) e = hipMemset(u->v, 0, p * 4);
        if (e == hipSuccess) e = hipMemset(u->wf, 0, hw * 2);
        if (e == hipSuccess) e = hipMemset(u->wb, 0, hw * 2);
        if (e != hipSuccess) rc = trt_fail(TRT_E_DEVICE, std::string("trt_unet_create: ") + hipGetErrorString(e));
    }
    if (rc) {
        std::string keep = trt_last_error();
        for (void *q : u->owned) hipFree(q);
        delete u;
        return trt_fail(rc, keep);
    }
    trt_device_ref(dev, 1);
    *out = u;
    return trt_ok();
}

int trt_unet_release(trt_unet *u) {
    if (!u) return trt_ok();
    hipSetDevice(trt_device_ordinal(u->dev));
    hipStreamSynchronize(trt_device_stream(u->dev));
    for (void *q : u->owned) hipFree(q);
    trt_device *d = u->dev;
    delete u;
    trt_device_ref(d, -1);
    return trt_ok();
}

int trt_unet_set_tall_tiles(uint32_t min_rows) {
    trt::set_conv_tall_min(min_rows);
    return trt_ok();
}

int trt_unet_param_count(trt_unet *u, uint64_t *count) {
    if (!u || !count) return trt_fail(TRT_E_INVALID, "trt_unet_param_count: null argument");
    *count = u->nparam;
    return trt_ok();
}

int trt_unet_params_set(trt_unet *u, const float *params, uint64_t count) {
    if (!u || !params) return trt_fail(TRT_E_INVALID, "trt_unet_params_set: null argument");
    if (count != u->nparam) return trt_fail(TRT_E_INVALID, "trt_unet_params_set: wrong parameter count");
    trt_device_lock(u->dev, true);
    hipStream_t s = trt_device_stream(u->dev);
    int rc = TRT_OK;
    hipError_t e = hipSetDevice(trt_device_ordinal(u->dev));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(u->params, params, count * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(u->m, 0, count * 4, s);
    if (e == hipSuccess) e = hipMemsetAsync(u->v, 0, count * 4, s);
    if (e != hipSuccess) rc = trt_fail(TRT_E_DEVICE, std::string("trt_unet_params_set: ") + hipGetErrorString(e));
    if (!rc) rc = cast_all(u, s);
    if (!rc) u->step = 0;
    trt_device_lock(u->dev, false);
    return rc ? rc : trt_ok();
}

int trt_unet_params_get(trt_unet *u, float *params, uint64_t count) {
    if (!u || !params) return trt_fail(TRT_E_INVALID, "trt_unet_params_get: null argument");
    if (count != u->nparam) return trt_fail(TRT_E_INVALID, "trt_unet_params_get: wrong parameter count");
    UNET_HIP(hipSetDevice(trt_device_ordinal(u->dev)));
    UNET_HIP(hipStreamSynchronize(trt_device_stream(u->dev)));
    UNET_HIP(hipMemcpy(params, u->params, count * 4, hipMemcpyDeviceToHost));
    return trt_ok();
}

int trt_unet_denoise(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, const float *color, const float *albedo,
                     const float *normal, float *out) {
    if (!u || !color || !albedo || !normal || !out) return trt_fail(TRT_E_INVALID, "trt_unet_denoise: null argument");
    int rc = check_shape(u, n, h, w);
    if (rc) return rc;
    trt_device_lock(u->dev, true);
    hipStream_t s = trt_device_stream(u->dev);
    hipSetDevice(trt_device_ordinal(u->dev));
    rc = forward(u, n, h, w, color, albedo, normal, s, out);  // the colour written by the last convolution
    trt_device_lock(u->dev, false);
    return rc ? rc : trt_ok();
}

static int loss_and_grads(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, const float *color, const float *albedo,
                          const float *normal, const float *target, hipStream_t s) {
    int rc = forward(u, n, h, w, color, albedo, normal, s);
    if (rc) return rc;
    UNET_HIP(hipMemsetAsync(u->loss, 0, 4, s));
    UNET_HIP(hipMemsetAsync(u->grads, 0, u->nparam * 4, s));
    UNET_LAUNCH(trt::launch_loss(u->y, target, u->g[0], u->loss, uint64_t(n) * h * w, s));
    return backward(u, n, h, w, s);
}

int trt_unet_gradients(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, const float *color, const float *albedo,
                       const float *normal, const float *target, float *loss, float *grads) {
    if (!u || !color || !albedo || !normal || !target || !loss || !grads)
        return trt_fail(TRT_E_INVALID, "trt_unet_gradients: null argument");
    int rc = check_shape(u, n, h, w);
    if (rc) return rc;
    trt_device_lock(u->dev, true);
    hipStream_t s = trt_device_stream(u->dev);
    hipSetDevice(trt_device_ordinal(u->dev));
    rc = loss_and_grads(u, n, h, w, color, albedo, normal, target, s);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(loss, u->loss, 4, hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipMemcpyAsync(grads, u->grads, u->nparam * 4, hipMemcpyDeviceToHost, s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = trt_fail(TRT_E_DEVICE, std::string("trt_unet_gradients: ") + hipGetErrorString(e));
    trt_device_lock(u->dev, false);
    return rc ? rc : trt_ok();
}

int trt_unet_train_step(trt_unet *u, uint32_t n, uint32_t h, uint32_t w, const float *color, const float *albedo,
                        const float *normal, const float *target, float lr, float *loss) {
    if (!u || !color || !albedo || !normal || !target)
        return trt_fail(TRT_E_INVALID, "trt_unet_train_step: null argument");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return trt_fail(TRT_E_INVALID, "trt_unet_train_step: bad learning rate");
    int rc = check_shape(u, n, h, w);
    if (rc) return rc;
    trt_device_lock(u->dev, true);
    hipStream_t s = trt_device_stream(u->dev);
    hipSetDevice(trt_device_ordinal(u->dev));
    rc = loss_and_grads(u, n, h, w, color, albedo, normal, target, s);
    if (!rc) {
        u->step++;
        if (trt::launch_adam(u->params, u->grads, u->m, u->v, u->nparam, lr, 0.9f, 0.999f, 1e-8f, u->step, s) != 0)
            rc = trt_fail(TRT_E_DEVICE, "trt_unet_train_step: launch failed");
    }
    if (!rc) rc = cast_all(u, s);
    if (!rc && loss) {
        hipError_t e = hipMemcpyAsync(loss, u->loss, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = trt_fail(TRT_E_DEVICE, std::string("trt_unet_train_step: ") + hipGetErrorString(e));
    }
    trt_device_lock(u->dev, false);
    return rc ? rc : trt_ok();
}

}  // extern "C"
