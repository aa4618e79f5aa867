// The learned denoiser's kernels for gfx950 (wave64): 3x3 convolutions as implicit GEMMs on
// v_mfma_f32_16x16x32_bf16 (A = 16 output channels x 32 inputs, B = 32 inputs x 16 pixels, FP32 accumulators),
// their weight gradients on the same instruction (K = pixels), and the memory-bound glue around them.
//
// Convolution: a block computes an 8x8 pixel tile for 16 * NCO output channels. The tile's 10x10 halo of input
// pixels (all input channels, zero outside the image, the second input read at half resolution) is staged once in
// LDS with a 16-byte pad per pixel; each wave then streams the weights of its 16 channels from L1/L2 and multiplies
// them against NCO of the tile's four 16-pixel strips per K step (one tap, 32 input channels).
// Weight gradient: a block owns 32 input channels x 16 * NCO output channels of dW for a run of tiles; per tile it
// stages dY transposed ([channel][pixel]) and the input halo in three column-shifted copies ([dx][channel][row][8]),
// so every MFMA operand is one aligned 16-byte LDS read, then adds its FP32 partial sums to dW with atomics.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>

#include "unet_kernels.h"

namespace trt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round to nearest even; a NaN stays a quiet NaN.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ---- convolution ------------------------------------------------------------------------------------------------
// The halo's rows in LDS: `width` pixels of `stride` elements (a multiple of 8: 16-byte units, an odd number of them),
// the row padded to a pitch of 8 units modulo 16. A B operand's 16 lanes read 8 pixels of one halo row and the 8
// under them; with an odd pixel stride the 8 land on 8 distinct 16-byte bank groups of the 16 and, a row 8 groups
// further on, the other 8 land on the other 8: each ds_read_b128 is conflict-free (measured before: two thirds of
// the LDS cycles of the full-resolution layers were bank conflicts).
__host__ __device__ constexpr uint32_t row_pitch(uint32_t stride, uint32_t width) {
    return (width * (stride / 8) + (8u + 16u - (width * (stride / 8)) % 16u) % 16u) * 8u;
}

// TH: the tile's height in pixels (8 or 16; the width is 8). A taller tile shares each halo row and each weight load
// over more pixels: the full-resolution layers, whose blocks mostly wait on their halo, take 16.
template <int NCO, int TH>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a, uint32_t tiles_x, uint32_t tiles_y) {
    static_assert(TH == 8 || TH == 16, "tiles are 8 or 16 pixels tall");
    constexpr int S = TH * NCO / 8;  // 16-pixel strips (two tile rows each) a wave takes
    extern __shared__ uint4 smem[];
    uint16_t *halo = reinterpret_cast<uint16_t *>(smem);
    const uint32_t cin = a.c0 + a.c1, stride = cin + 8, rp = row_pitch(stride, 10);
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, n = tile / (tiles_x * tiles_y);
    const int x0 = int(tx * 8) - 1, y0 = int(ty * TH) - 1;
    const uint32_t h2 = a.h / 2, w2 = a.w / 2;
    const uint32_t chunks = cin / 8;
    for (uint32_t q = threadIdx.x; q < (TH + 2) * 10 * chunks; q += 256) {
        const uint32_t hp = q / chunks, ch = (q % chunks) * 8;
        const int gy = y0 + int(hp / 10), gx = x0 + int(hp % 10);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (gy >= 0 && gy < int(a.h) && gx >= 0 && gx < int(a.w)) {
            if (ch < a.c0)
                v = *reinterpret_cast<const uint4 *>(a.in0 + ((uint64_t(n) * a.h + gy) * a.w + gx) * a.c0 + ch);
            else
                v = *reinterpret_cast<const uint4 *>(a.in1 + ((uint64_t(n) * h2 + gy / 2) * w2 + gx / 2) * a.c1 +
                                                     (ch - a.c0));
        }
        *reinterpret_cast<uint4 *>(halo + (hp / 10) * rp + (hp % 10) * stride + ch) = v;
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, li = lane & 15u, g = lane >> 4;
    const uint32_t co_tile = wave % NCO, grp = wave / NCO;
    const uint32_t co_base = blockIdx.y * (16 * NCO) + co_tile * 16;
    const uint16_t *wrow = a.wt + uint64_t(co_base + li) * 9 * cin + g * 8;
    f32x4 acc[S];
#pragma unroll
    for (int j = 0; j < S; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t tap = 0; tap < 9; tap++) {
        const uint32_t dy = tap / 3, dx = tap % 3;
        const uint16_t *hrow = halo + (dy + (li >> 3)) * rp + ((li & 7u) + dx) * stride + g * 8;
        for (uint32_t c = 0; c < cin; c += 32) {
            const bf16x8 A = as_bf16x8(*reinterpret_cast<const uint4 *>(wrow + tap * cin + c));
#pragma unroll
            for (int j = 0; j < S; j++) {
                const uint32_t p = grp * S + j;  // the strip: tile rows 2p and 2p + 1
                const bf16x8 B = as_bf16x8(*reinterpret_cast<const uint4 *>(hrow + 2 * p * rp + c));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
            }
        }
    }
    // D[co][pixel]: this lane holds channels co_base + 4g .. +3 of pixel li of each strip
    const uint32_t co = co_base + 4 * g;
    float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (a.bias)
        for (int r = 0; r < 4; r++) b[r] = a.bias[co + r];
#pragma unroll
    for (int j = 0; j < S; j++) {
        const uint32_t p = grp * S + j;
        const uint32_t y = ty * TH + 2 * p + (li >> 3), x = tx * 8 + (li & 7u);
        if (y >= a.h || x >= a.w) continue;
        float v[4];
        for (int r = 0; r < 4; r++) {
            v[r] = acc[j][r] + b[r];
            if (a.relu) v[r] = fmaxf(v[r], 0.0f);
        }
        const uint64_t o = ((uint64_t(n) * a.h + y) * a.w + x) * a.cout + co;
        if (a.out_f32) *reinterpret_cast<float4 *>(a.out_f32 + o) = make_float4(v[0], v[1], v[2], v[3]);
        if (a.out_bf16) {
            const uint32_t lo = uint32_t(to_bf16(v[0])) | (uint32_t(to_bf16(v[1])) << 16);
            const uint32_t hi = uint32_t(to_bf16(v[2])) | (uint32_t(to_bf16(v[3])) << 16);
            *reinterpret_cast<uint2 *>(a.out_bf16 + o) = make_uint2(lo, hi);
        }
    }
}

// k_conv for cin = 32 * CH (CH = 1 .. 7, not 6) with every size fixed at compile time: the halo's loads all issued
// before its stores, its indices and the operands' LDS offsets constants, the outputs rounded to bf16 by the
// hardware's v_cvt_pk_bf16_f32 (round to nearest even, as to_bf16). k_conv spent its time in index arithmetic and
// software rounding (the full-resolution layers were VALU-bound: 743 VALU instructions a wave against 36 MFMAs).
// Per output the same MFMAs in the same order, so the same bits. Inference's last layer writes k_output's colour
// instead of its 32 FP32 channels (ConvArgs out_color).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t));
}

template <int NCO, int TH, int CH>
__global__ __launch_bounds__(256) void k_conv_c(ConvArgs a, uint32_t tiles_x, uint32_t tiles_y) {
    static_assert(TH == 8 || TH == 16, "tiles are 8 or 16 pixels tall");
    constexpr int S = TH * NCO / 8;  // 16-pixel strips (two tile rows each) a wave takes
    constexpr uint32_t cin = CH * 32, stride = cin + 8, C4 = CH * 4, RP = row_pitch(stride, 10);
    constexpr uint32_t items = (TH + 2) * 10 * C4, per = (items + 255) / 256;
    extern __shared__ uint4 smem[];
    uint16_t *halo = reinterpret_cast<uint16_t *>(smem);
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, n = tile / (tiles_x * tiles_y);
    const int x0 = int(tx * 8) - 1, y0 = int(ty * TH) - 1;
    const uint32_t h2 = a.h / 2, w2 = a.w / 2, c0q = a.c0 / 8;
    const uint16_t *in0 = a.in0 + uint64_t(n) * a.h * a.w * a.c0;
    const uint16_t *in1 = a.in1 ? a.in1 + uint64_t(n) * h2 * w2 * a.c1 : nullptr;
    uint4 v[per];
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t q = threadIdx.x + k * 256;
        const uint32_t hp = q / C4, ck = q % C4;
        const uint32_t hy = hp / 10, hx = hp - hy * 10;
        const int gy = y0 + int(hy), gx = x0 + int(hx);
        v[k] = make_uint4(0, 0, 0, 0);
        if (((k + 1) * 256 <= items || q < items) && uint32_t(gy) < a.h && uint32_t(gx) < a.w) {
            const uint32_t pix = uint32_t(gy) * a.w + uint32_t(gx);
            if (ck < c0q)
                v[k] = *reinterpret_cast<const uint4 *>(in0 + pix * a.c0 + ck * 8);
            else
                v[k] = *reinterpret_cast<const uint4 *>(in1 + (uint32_t(gy / 2) * w2 + uint32_t(gx / 2)) * a.c1 +
                                                        (ck - c0q) * 8);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t q = threadIdx.x + k * 256;
        if ((k + 1) * 256 <= items || q < items)
            *reinterpret_cast<uint4 *>(halo + (q / C4 / 10) * RP + (q / C4 % 10) * stride + (q % C4) * 8) = v[k];
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, li = lane & 15u, g = lane >> 4;
    const uint32_t co_tile = wave % NCO, grp = wave / NCO;
    const uint32_t co_base = blockIdx.y * (16 * NCO) + co_tile * 16;
    const uint16_t *wrow = a.wt + uint64_t(co_base + li) * 9 * cin + g * 8;
    const uint16_t *hbase = halo + ((li >> 3) + 2 * grp * S) * RP + (li & 7u) * stride + g * 8;
    f32x4 acc[S];
#pragma unroll
    for (int j = 0; j < S; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t tap = 0; tap < 9; tap++) {
        const uint32_t dy = tap / 3, dx = tap % 3;
        const uint16_t *hrow = hbase + dy * RP + dx * stride;
#pragma unroll
        for (uint32_t c = 0; c < cin; c += 32) {
            const bf16x8 A = as_bf16x8(*reinterpret_cast<const uint4 *>(wrow + tap * cin + c));
#pragma unroll
            for (int j = 0; j < S; j++) {
                const bf16x8 B = as_bf16x8(*reinterpret_cast<const uint4 *>(hrow + 2 * j * RP + c));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
            }
        }
    }
    // D[co][pixel]: this lane holds channels co .. co + 3 of pixel li of each strip
    const uint32_t co = co_base + 4 * g;
    if (a.out_color && co != 0) return;  // the colour is channels 0-2
    float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (a.bias)
        for (int r = 0; r < 4; r++) b[r] = a.bias[co + r];
#pragma unroll
    for (int j = 0; j < S; j++) {
        const uint32_t p = grp * S + j;
        const uint32_t y = ty * TH + 2 * p + (li >> 3), x = tx * 8 + (li & 7u);
        if (y >= a.h || x >= a.w) continue;
        float r4[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            r4[r] = acc[j][r] + b[r];
            if (a.relu) r4[r] = fmaxf(r4[r], 0.0f);
        }
        const uint64_t pix = (uint64_t(n) * a.h + y) * a.w + x;
        if (a.out_color) {
            for (int r = 0; r < 3; r++) a.out_color[3 * pix + r] = expm1f(fmaxf(r4[r], 0.0f));
            continue;
        }
        const uint64_t o = pix * a.cout + co;
        if (a.out_f32) *reinterpret_cast<float4 *>(a.out_f32 + o) = make_float4(r4[0], r4[1], r4[2], r4[3]);
        if (a.out_bf16)
            *reinterpret_cast<uint2 *>(a.out_bf16 + o) = make_uint2(pack_bf16(r4[0], r4[1]), pack_bf16(r4[2], r4[3]));
    }
}

// k_conv_c for a decoder's concatenation, cin = 32 * (C0Q + C1Q): the second input (half resolution, upsampled by 2)
// is staged in LDS once, at its own resolution ((TH / 2 + 2) x 6 pixels for the tile's (TH + 2) x 10 halo), and each
// lane reads the half-resolution pixel under its full-resolution one. k_conv_c staged the upsampled copy, four times
// the pixels: twice the LDS, fewer blocks a CU, and four reads of every half-resolution pixel. Per output the same
// MFMAs in the same order (a tap's first-input channels, then its second input's), so the same bits.
template <int NCO, int TH, int C0Q, int C1Q>
__global__ __launch_bounds__(256) void k_conv_up(ConvArgs a, uint32_t tiles_x, uint32_t tiles_y) {
    static_assert(TH == 8 || TH == 16, "tiles are 8 or 16 pixels tall");
    constexpr int S = TH * NCO / 8;
    constexpr uint32_t c0 = C0Q * 32, c1 = C1Q * 32, cin = c0 + c1, st0 = c0 + 8, st1 = c1 + 8;
    constexpr uint32_t RP0 = row_pitch(st0, 10), RP1 = row_pitch(st1, 6);
    constexpr uint32_t R1 = TH / 2 + 2, W1 = 6;                 // the half-resolution halo's rows and columns
    constexpr uint32_t q0 = c0 / 8, q1 = c1 / 8;                // 16-byte chunks a pixel
    constexpr uint32_t n0 = (TH + 2) * 10 * q0, items = n0 + R1 * W1 * q1, per = (items + 255) / 256;
    extern __shared__ uint4 smem[];
    uint16_t *halo0 = reinterpret_cast<uint16_t *>(smem);
    uint16_t *halo1 = halo0 + (TH + 2) * RP0;
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, n = tile / (tiles_x * tiles_y);
    const int x0 = int(tx * 8) - 1, y0 = int(ty * TH) - 1;       // the full-resolution halo's origin
    const int hx0 = int(tx * 4) - 1, hy0 = int(ty * TH / 2) - 1;  // the half-resolution halo's
    const uint32_t h2 = a.h / 2, w2 = a.w / 2;
    const uint16_t *in0 = a.in0 + uint64_t(n) * a.h * a.w * c0;
    const uint16_t *in1 = a.in1 + uint64_t(n) * h2 * w2 * c1;
    uint4 v[per];
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t q = threadIdx.x + k * 256;
        v[k] = make_uint4(0, 0, 0, 0);
        if (q < n0) {
            const uint32_t hp = q / q0, ck = q % q0;
            const uint32_t hy = hp / 10, hx = hp - hy * 10;
            const int gy = y0 + int(hy), gx = x0 + int(hx);
            if (uint32_t(gy) < a.h && uint32_t(gx) < a.w)
                v[k] = *reinterpret_cast<const uint4 *>(in0 + (uint32_t(gy) * a.w + uint32_t(gx)) * c0 + ck * 8);
        } else if ((k + 1) * 256 <= items || q < items) {
            const uint32_t r = q - n0, hp = r / q1, ck = r % q1;
            const uint32_t hy = hp / W1, hx = hp - hy * W1;
            const int gy = hy0 + int(hy), gx = hx0 + int(hx);
            if (uint32_t(gy) < h2 && uint32_t(gx) < w2)
                v[k] = *reinterpret_cast<const uint4 *>(in1 + (uint32_t(gy) * w2 + uint32_t(gx)) * c1 + ck * 8);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t q = threadIdx.x + k * 256;
        if (q < n0) {
            *reinterpret_cast<uint4 *>(halo0 + (q / q0 / 10) * RP0 + (q / q0 % 10) * st0 + (q % q0) * 8) = v[k];
        } else if ((k + 1) * 256 <= items || q < items) {
            const uint32_t r = q - n0;
            *reinterpret_cast<uint4 *>(halo1 + (r / q1 / W1) * RP1 + (r / q1 % W1) * st1 + (r % q1) * 8) = v[k];
        }
    }
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, li = lane & 15u, g = lane >> 4;
    const uint32_t co_tile = wave % NCO, grp = wave / NCO;
    const uint32_t co_base = blockIdx.y * (16 * NCO) + co_tile * 16;
    const uint16_t *wrow = a.wt + uint64_t(co_base + li) * 9 * cin + g * 8;
    const uint16_t *hbase0 = halo0 + ((li >> 3) + 2 * grp * S) * RP0 + (li & 7u) * st0 + g * 8;
    f32x4 acc[S];
#pragma unroll
    for (int j = 0; j < S; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t tap = 0; tap < 9; tap++) {
        const uint32_t dy = tap / 3, dx = tap % 3;
        const uint16_t *hrow0 = hbase0 + dy * RP0 + dx * st0;
        // the lane's full-resolution halo pixel (row 2 (grp S + j) + li / 8 + dy, column li % 8 + dx) lies on
        // half-resolution halo pixel (grp S + j + (li / 8 + dy + 1) / 2, (li % 8 + dx + 1) / 2)
        const uint32_t rr = ((li >> 3) + dy + 1) >> 1, cc = ((li & 7u) + dx + 1) >> 1;
        const uint16_t *hrow1 = halo1 + (grp * S + rr) * RP1 + cc * st1 + g * 8;
#pragma unroll
        for (uint32_t c = 0; c < c0; c += 32) {
            const bf16x8 A = as_bf16x8(*reinterpret_cast<const uint4 *>(wrow + tap * cin + c));
#pragma unroll
            for (int j = 0; j < S; j++) {
                const bf16x8 B = as_bf16x8(*reinterpret_cast<const uint4 *>(hrow0 + 2 * j * RP0 + c));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
            }
        }
#pragma unroll
        for (uint32_t c = 0; c < c1; c += 32) {
            const bf16x8 A = as_bf16x8(*reinterpret_cast<const uint4 *>(wrow + tap * cin + c0 + c));
#pragma unroll
            for (int j = 0; j < S; j++) {
                const bf16x8 B = as_bf16x8(*reinterpret_cast<const uint4 *>(hrow1 + j * RP1 + c));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
            }
        }
    }
    const uint32_t co = co_base + 4 * g;
    float b[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (a.bias)
        for (int r = 0; r < 4; r++) b[r] = a.bias[co + r];
#pragma unroll
    for (int j = 0; j < S; j++) {
        const uint32_t p = grp * S + j;
        const uint32_t y = ty * TH + 2 * p + (li >> 3), x = tx * 8 + (li & 7u);
        if (y >= a.h || x >= a.w) continue;
        float r4[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            r4[r] = acc[j][r] + b[r];
            if (a.relu) r4[r] = fmaxf(r4[r], 0.0f);
        }
        const uint64_t o = ((uint64_t(n) * a.h + y) * a.w + x) * a.cout + co;
        if (a.out_f32) *reinterpret_cast<float4 *>(a.out_f32 + o) = make_float4(r4[0], r4[1], r4[2], r4[3]);
        if (a.out_bf16)
            *reinterpret_cast<uint2 *>(a.out_bf16 + o) = make_uint2(pack_bf16(r4[0], r4[1]), pack_bf16(r4[2], r4[3]));
    }
}

// ---- weight gradient --------------------------------------------------------------------------------------------
template <int NCO>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a, uint32_t tiles_x, uint32_t tiles_y, uint32_t tiles,
                                               uint32_t per_block) {
    __shared__ uint16_t dyt[64 * 64];           // [channel of the block][pixel of the tile]
    __shared__ uint16_t xs[3 * 32 * 10 * 8];    // [dx][input channel][halo row][column]
    constexpr uint32_t kParts = 4 / NCO;         // waves sharing one 16-channel output tile
    constexpr int kMax = (18 + kParts - 1) / kParts;
    const uint32_t cin = a.c0 + a.c1, cq = 16 * NCO;
    const uint32_t ci_base = blockIdx.y * 32, co_block = blockIdx.z * cq;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, li = lane & 15u, g = lane >> 4;
    const uint32_t co_tile = wave % NCO, part = wave / NCO;
    const uint32_t h2 = a.h / 2, w2 = a.w / 2;
    f32x4 acc[kMax];
#pragma unroll
    for (int k = 0; k < kMax; k++) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t t0 = blockIdx.x * per_block, t1 = min(t0 + per_block, tiles);
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, n = t / (tiles_x * tiles_y);
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < 64 * (cq / 8); q += 256) {
            const uint32_t px = q / (cq / 8), ch = (q % (cq / 8)) * 8;
            const uint32_t y = ty * 8 + px / 8, x = tx * 8 + px % 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (y < a.h && x < a.w)
                v = *reinterpret_cast<const uint4 *>(a.dy + ((uint64_t(n) * a.h + y) * a.w + x) * a.cout + co_block +
                                                     ch);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            for (int e = 0; e < 8; e++) dyt[(ch + e) * 64 + px] = static_cast<uint16_t>(w4[e / 2] >> (16 * (e & 1)));
        }
        for (uint32_t q = threadIdx.x; q < 400; q += 256) {
            const uint32_t hp = q / 4, ch = (q % 4) * 8, hy = hp / 10, hx = hp % 10;
            const int gy = int(ty * 8) - 1 + int(hy), gx = int(tx * 8) - 1 + int(hx);
            const uint32_t c = ci_base + ch;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gy >= 0 && gy < int(a.h) && gx >= 0 && gx < int(a.w)) {
                if (c < a.c0)
                    v = *reinterpret_cast<const uint4 *>(a.in0 + ((uint64_t(n) * a.h + gy) * a.w + gx) * a.c0 + c);
                else
                    v = *reinterpret_cast<const uint4 *>(a.in1 + ((uint64_t(n) * h2 + gy / 2) * w2 + gx / 2) * a.c1 +
                                                         (c - a.c0));
            }
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            for (int e = 0; e < 8; e++) {
                const uint16_t s = static_cast<uint16_t>(w4[e / 2] >> (16 * (e & 1)));
                for (uint32_t dx = 0; dx < 3; dx++) {
                    const int col = int(hx) - int(dx);
                    if (col >= 0 && col < 8) xs[((dx * 32 + ch + e) * 10 + hy) * 8 + col] = s;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t s = 0; s < 2; s++) {
            // K = the 32 pixels of tile rows 4s .. 4s + 3; lane group g holds row 4s + g
            const bf16x8 A = as_bf16x8(*reinterpret_cast<const uint4 *>(dyt + (co_tile * 16 + li) * 64 + (4 * s + g) * 8));
#pragma unroll
            for (int k = 0; k < kMax; k++) {
                const uint32_t qq = part + k * kParts;  // (tap, input-channel tile) pair
                if (qq >= 18) continue;
                const uint32_t tap = qq / 2, ct = qq % 2, dy = tap / 3, dx = tap % 3;
                const bf16x8 B = as_bf16x8(
                    *reinterpret_cast<const uint4 *>(xs + ((dx * 32 + ct * 16 + li) * 10 + 4 * s + g + dy) * 8));
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[k], 0, 0, 0);
            }
        }
    }
    // D[co][ci]: this lane holds output channels 4g .. 4g + 3 of the tile for input channel li
#pragma unroll
    for (int k = 0; k < kMax; k++) {
        const uint32_t qq = part + k * kParts;
        if (qq >= 18) continue;
        const uint32_t tap = qq / 2, ct = qq % 2;
        const uint32_t ci = ci_base + ct * 16 + li;
        for (int r = 0; r < 4; r++) {
            const uint32_t co = co_block + co_tile * 16 + 4 * g + r;
            atomicAdd(a.dw + (uint64_t(co) * 9 + tap) * cin + ci, acc[k][r]);
        }
    }
}

// db[c] += sum over a run of 64 pixels of dy[p][c]
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *dy, float *db, uint32_t cout, uint64_t pixels) {
    const uint64_t p0 = uint64_t(blockIdx.x) * 64;
    for (uint32_t c = threadIdx.x; c < cout; c += 256) {
        float s = 0.0f;
        for (uint32_t k = 0; k < 64 && p0 + k < pixels; k++) s += from_bf16(dy[(p0 + k) * cout + c]);
        atomicAdd(db + c, s);
    }
}

// ---- glue -------------------------------------------------------------------------------------------------------
__global__ void k_pool(const uint16_t *in, uint16_t *out, uint32_t n, uint32_t h, uint32_t w, uint32_t c) {
    const uint32_t h2 = h / 2, w2 = w / 2, chunks = c / 8;
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(n) * h2 * w2 * chunks) return;
    const uint32_t ch = uint32_t(i % chunks) * 8;
    const uint64_t p = i / chunks;
    const uint32_t x = uint32_t(p % w2), y = uint32_t((p / w2) % h2);
    const uint64_t nn = p / (uint64_t(w2) * h2);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t k = 0; k < 4; k++) {
        const uint64_t q = ((nn * h + 2 * y + k / 2) * w + 2 * x + k % 2) * c + ch;
        const uint4 v = *reinterpret_cast<const uint4 *>(in + q);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        for (int e = 0; e < 8; e++) s[e] += from_bf16(static_cast<uint16_t>(w4[e / 2] >> (16 * (e & 1))));
    }
    uint32_t o[4];
    for (int e = 0; e < 4; e++)
        o[e] = uint32_t(to_bf16(s[2 * e] * 0.25f)) | (uint32_t(to_bf16(s[2 * e + 1] * 0.25f)) << 16);
    *reinterpret_cast<uint4 *>(out + p * c + ch) = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ void k_mask(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out,
                       uint64_t count, uint32_t c) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t p = i / c;
    const uint32_t k = uint32_t(i % c);
    out[i] = from_bf16(act[i]) > 0.0f ? to_bf16(dx[p * ctot + coff + k]) : uint16_t(0);
}

__global__ void k_mask_pool_sum(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out,
                                uint32_t n, uint32_t h2, uint32_t w2, uint32_t c) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(n) * h2 * w2 * c) return;
    const uint32_t k = uint32_t(i % c);
    const uint64_t p = i / c;
    const uint32_t x = uint32_t(p % w2), y = uint32_t((p / w2) % h2);
    const uint64_t nn = p / (uint64_t(w2) * h2);
    const uint32_t h = 2 * h2, w = 2 * w2;
    float s = 0.0f;
    for (uint32_t q = 0; q < 4; q++) s += dx[((nn * h + 2 * y + q / 2) * w + 2 * x + q % 2) * ctot + coff + k];
    out[i] = from_bf16(act[i]) > 0.0f ? to_bf16(s) : uint16_t(0);
}

__global__ void k_mask_skip_up(const float *skip, uint32_t ctot, const float *dp, const uint16_t *act, uint16_t *out,
                               uint32_t n, uint32_t h, uint32_t w, uint32_t c) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(n) * h * w * c) return;
    const uint32_t k = uint32_t(i % c);
    const uint64_t p = i / c;
    const uint32_t x = uint32_t(p % w), y = uint32_t((p / w) % h);
    const uint64_t nn = p / (uint64_t(w) * h);
    const float v = skip[p * ctot + k] + dp[((nn * (h / 2) + y / 2) * (w / 2) + x / 2) * c + k] * 0.25f;
    out[i] = from_bf16(act[i]) > 0.0f ? to_bf16(v) : uint16_t(0);
}

__global__ void k_features(const float *color, const float *albedo, const float *normal, uint16_t *out,
                           uint64_t pixels) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    float f[32];
    for (int k = 0; k < 32; k++) f[k] = 0.0f;
    for (int k = 0; k < 3; k++) {
        f[k] = log1pf(fmaxf(color[3 * p + k], 0.0f));
        f[3 + k] = albedo[3 * p + k];
        f[6 + k] = normal[3 * p + k] * 0.5f + 0.5f;
    }
    uint4 *o = reinterpret_cast<uint4 *>(out + p * 32);
    for (int q = 0; q < 4; q++) {
        uint32_t w4[4];
        for (int e = 0; e < 4; e++)
            w4[e] = uint32_t(to_bf16(f[8 * q + 2 * e])) | (uint32_t(to_bf16(f[8 * q + 2 * e + 1])) << 16);
        o[q] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
}

__global__ void k_output(const float *y, float *color, uint64_t pixels) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    for (int k = 0; k < 3; k++) color[3 * p + k] = expm1f(fmaxf(y[p * 32 + k], 0.0f));
}

__global__ __launch_bounds__(256) void k_loss(const float *y, const float *target, uint16_t *dy, float *loss,
                                              uint64_t pixels) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const float inv = 1.0f / (3.0f * static_cast<float>(pixels));
    float s = 0.0f;
    if (p < pixels) {
        uint32_t g[16];
        for (int k = 0; k < 16; k++) g[k] = 0;
        for (int k = 0; k < 3; k++) {
            const float d = y[p * 32 + k] - log1pf(fmaxf(target[3 * p + k], 0.0f));
            s += fabsf(d);
            const float gk = d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f);
            const uint32_t b = to_bf16(gk);
            g[k / 2] |= b << (16 * (k & 1));
        }
        uint4 *o = reinterpret_cast<uint4 *>(dy + p * 32);
        for (int q = 0; q < 4; q++) o[q] = make_uint4(g[4 * q], g[4 * q + 1], g[4 * q + 2], g[4 * q + 3]);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicAdd(loss, s * inv);
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_cast_weights(const float *w, uint16_t *wf, uint16_t *wb, uint32_t cout, uint32_t cin) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(cout) * 9 * cin) return;
    const uint32_t ci = uint32_t(i % cin), tap = uint32_t((i / cin) % 9), co = uint32_t(i / (uint64_t(cin) * 9));
    const uint16_t b = to_bf16(w[i]);
    wf[i] = b;
    wb[(uint64_t(ci) * 9 + (8 - tap)) * cout + co] = b;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return static_cast<uint32_t>((n + per - 1) / per); }
inline int pick_nco(uint32_t cout) { return cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1); }

}  // namespace

// The level height from which convolutions take 16-pixel-tall tiles (the same sums as 8-pixel tiles, so the same bits).
static std::atomic<uint32_t> g_tall_min{128};
void set_conv_tall_min(uint32_t rows) { g_tall_min = rows; }

int launch_conv(const ConvArgs &a, hipStream_t s) {
    const uint32_t cin = a.c0 + a.c1;
    if (a.cout % 16 || cin % 32 || a.c0 % 32 || (a.c1 && (a.h % 2 || a.w % 2)) || cin > 512) return -1;
    const bool tall = a.h >= g_tall_min.load() && a.h >= 16;  // 16-pixel-tall tiles for the large levels
    const uint32_t th = tall ? 16 : 8;
    const uint32_t tx = (a.w + 7) / 8, ty = (a.h + th - 1) / th;
    const uint64_t tiles = uint64_t(tx) * ty * a.n;
    if (tiles == 0) return 0;
    if (tiles >= (uint64_t{1} << 31)) return -1;
    const size_t lds = size_t(th + 2) * row_pitch(cin + 8, 10) * 2;
    const int nco = pick_nco(a.cout);
    const dim3 grid(static_cast<uint32_t>(tiles), a.cout / (16 * nco));
    const uint32_t ch = cin / 32;
    if (a.out_color && (cin != 32 || a.c1)) return -1;  // the fused colour output takes cin = 32 only
    // the decoder's concatenations (unet.py ARCH: 32 + 64, 64 + 64, 64 + 96, 96 + 128) stage their second input at half
    // resolution
    const int up = a.c1 == 0 ? 0 : (a.c0 == 32 && a.c1 == 64 && nco == 2) ? 1 : (a.c0 == 64 && a.c1 == 64 && nco == 4) ? 2
                 : (a.c0 == 64 && a.c1 == 96 && nco == 4) ? 3 : (a.c0 == 96 && a.c1 == 128 && nco == 2) ? 4 : 0;
    if (up) {
        const size_t lds_up = (size_t(th + 2) * row_pitch(a.c0 + 8, 10) + size_t(th / 2 + 2) * row_pitch(a.c1 + 8, 6)) * 2;
#define TRT_U(T)                                                                                                        \
    switch (up) {                                                                                                       \
    case 1: hipLaunchKernelGGL((k_conv_up<2, T, 1, 2>), grid, dim3(256), lds_up, s, a, tx, ty); break;                \
    case 2: hipLaunchKernelGGL((k_conv_up<4, T, 2, 2>), grid, dim3(256), lds_up, s, a, tx, ty); break;                \
    case 3: hipLaunchKernelGGL((k_conv_up<4, T, 2, 3>), grid, dim3(256), lds_up, s, a, tx, ty); break;                \
    default: hipLaunchKernelGGL((k_conv_up<2, T, 3, 4>), grid, dim3(256), lds_up, s, a, tx, ty); break;               \
    }
        if (tall) {
            TRT_U(16)
        } else {
            TRT_U(8)
        }
#undef TRT_U
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    // (64 output channels over 64 inputs on levels of 256 rows or more stay on k_conv: k_conv_c<4, 16, 2> needs
    // 220 registers, one wave a SIMD, and ran those layers 20 % slower)
    if (ch >= 1 && ch <= 7 && ch != 6 && !(nco == 4 && ch == 2 && a.h >= 256)) {
#define TRT_C(N, T)                                                                                    \
    switch (ch) {                                                                                      \
    case 1: hipLaunchKernelGGL((k_conv_c<N, T, 1>), grid, dim3(256), lds, s, a, tx, ty); break;      \
    case 2: hipLaunchKernelGGL((k_conv_c<N, T, 2>), grid, dim3(256), lds, s, a, tx, ty); break;      \
    case 3: hipLaunchKernelGGL((k_conv_c<N, T, 3>), grid, dim3(256), lds, s, a, tx, ty); break;      \
    case 4: hipLaunchKernelGGL((k_conv_c<N, T, 4>), grid, dim3(256), lds, s, a, tx, ty); break;      \
    case 5: hipLaunchKernelGGL((k_conv_c<N, T, 5>), grid, dim3(256), lds, s, a, tx, ty); break;      \
    default: hipLaunchKernelGGL((k_conv_c<N, T, 7>), grid, dim3(256), lds, s, a, tx, ty); break;     \
    }
        if (tall) {
            if (nco == 4) {
                TRT_C(4, 16)
            } else if (nco == 2) {
                TRT_C(2, 16)
            } else {
                TRT_C(1, 16)
            }
        } else {
            if (nco == 4) {
                TRT_C(4, 8)
            } else if (nco == 2) {
                TRT_C(2, 8)
            } else {
                TRT_C(1, 8)
            }
        }
#undef TRT_C
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    if (tall) {
        if (nco == 4)
            hipLaunchKernelGGL((k_conv<4, 16>), grid, dim3(256), lds, s, a, tx, ty);
        else if (nco == 2)
            hipLaunchKernelGGL((k_conv<2, 16>), grid, dim3(256), lds, s, a, tx, ty);
        else
            hipLaunchKernelGGL((k_conv<1, 16>), grid, dim3(256), lds, s, a, tx, ty);
    } else {
        if (nco == 4)
            hipLaunchKernelGGL((k_conv<4, 8>), grid, dim3(256), lds, s, a, tx, ty);
        else if (nco == 2)
            hipLaunchKernelGGL((k_conv<2, 8>), grid, dim3(256), lds, s, a, tx, ty);
        else
            hipLaunchKernelGGL((k_conv<1, 8>), grid, dim3(256), lds, s, a, tx, ty);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_wgrad(const WgradArgs &a, hipStream_t s) {
    const uint32_t cin = a.c0 + a.c1;
    if (a.cout % 16 || cin % 32 || a.c0 % 32 || (a.c1 && (a.h % 2 || a.w % 2))) return -1;
    const uint32_t tx = (a.w + 7) / 8, ty = (a.h + 7) / 8;
    const uint64_t tiles = uint64_t(tx) * ty * a.n;
    if (tiles == 0) return 0;
    if (tiles >= (uint64_t{1} << 31)) return -1;
    const int nco = pick_nco(a.cout);
    const uint32_t ci_chunks = cin / 32, co_chunks = a.cout / (16 * nco);
    // about 2048 blocks in all, each over a run of tiles
    const uint64_t others = uint64_t(ci_chunks) * co_chunks;
    uint64_t runs = std::max<uint64_t>(1, std::min<uint64_t>(tiles, (2048 + others - 1) / others));
    const uint32_t per = static_cast<uint32_t>((tiles + runs - 1) / runs);
    runs = (tiles + per - 1) / per;
    const dim3 grid(static_cast<uint32_t>(runs), ci_chunks, co_chunks);
    if (nco == 4)
        hipLaunchKernelGGL(k_wgrad<4>, grid, dim3(256), 0, s, a, tx, ty, static_cast<uint32_t>(tiles), per);
    else if (nco == 2)
        hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(256), 0, s, a, tx, ty, static_cast<uint32_t>(tiles), per);
    else
        hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(256), 0, s, a, tx, ty, static_cast<uint32_t>(tiles), per);
    if (a.db) {
        const uint64_t pixels = uint64_t(a.n) * a.h * a.w;
        hipLaunchKernelGGL(k_bias_grad, dim3(blocks_for(pixels, 64)), dim3(256), 0, s, a.dy, a.db, a.cout, pixels);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pool(const uint16_t *in, uint16_t *out, uint32_t n, uint32_t h, uint32_t w, uint32_t c, hipStream_t s) {
    if (c % 8 || h % 2 || w % 2) return -1;
    const uint64_t work = uint64_t(n) * (h / 2) * (w / 2) * (c / 8);
    if (work) hipLaunchKernelGGL(k_pool, dim3(blocks_for(work, 256)), dim3(256), 0, s, in, out, n, h, w, c);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mask(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out, uint32_t n,
                uint32_t h, uint32_t w, uint32_t c, hipStream_t s) {
    const uint64_t count = uint64_t(n) * h * w * c;
    if (count) hipLaunchKernelGGL(k_mask, dim3(blocks_for(count, 256)), dim3(256), 0, s, dx, ctot, coff, act, out, count, c);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mask_pool_sum(const float *dx, uint32_t ctot, uint32_t coff, const uint16_t *act, uint16_t *out,
                         uint32_t n, uint32_t h2, uint32_t w2, uint32_t c, hipStream_t s) {
    const uint64_t count = uint64_t(n) * h2 * w2 * c;
    if (count)
        hipLaunchKernelGGL(k_mask_pool_sum, dim3(blocks_for(count, 256)), dim3(256), 0, s, dx, ctot, coff, act, out,
                           n, h2, w2, c);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mask_skip_up(const float *skip, uint32_t ctot, const float *dp, const uint16_t *act, uint16_t *out,
                        uint32_t n, uint32_t h, uint32_t w, uint32_t c, hipStream_t s) {
    if (h % 2 || w % 2) return -1;
    const uint64_t count = uint64_t(n) * h * w * c;
    if (count)
        hipLaunchKernelGGL(k_mask_skip_up, dim3(blocks_for(count, 256)), dim3(256), 0, s, skip, ctot, dp, act, out, n,
                           h, w, c);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_features(const float *color, const float *albedo, const float *normal, uint16_t *out, uint64_t pixels,
                    hipStream_t s) {
    if (pixels) hipLaunchKernelGGL(k_features, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, color, albedo, normal, out, pixels);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_output(const float *y, float *color, uint64_t pixels, hipStream_t s) {
    if (pixels) hipLaunchKernelGGL(k_output, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, y, color, pixels);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_loss(const float *y, const float *target, uint16_t *dy, float *loss, uint64_t pixels, hipStream_t s) {
    if (pixels) hipLaunchKernelGGL(k_loss, dim3(blocks_for(pixels, 256)), dim3(256), 0, s, y, target, dy, loss, pixels);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_adam(float *p, const float *g, float *m, float *v, uint64_t count, float lr, float b1, float b2, float eps,
                uint32_t t, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, float(t)), c2 = 1.0f - powf(b2, float(t));
    if (count)
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, g, m, v, count, lr, b1, b2, eps,
                           c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cast_weights(const float *w, uint16_t *wf, uint16_t *wb, uint32_t cout, uint32_t cin, hipStream_t s) {
    const uint64_t count = uint64_t(cout) * 9 * cin;
    if (count) hipLaunchKernelGGL(k_cast_weights, dim3(blocks_for(count, 256)), dim3(256), 0, s, w, wf, wb, cout, cin);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
