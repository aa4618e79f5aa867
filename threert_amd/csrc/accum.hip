// The accumulator's resolve on the device (gpu.rs:4218-4231's readback normalisation, for a consumer that stays on
// the GPU, the learned denoiser): color, albedo and normal sums divided by the sample count, non-finite sums read as
// zero, into images of pw x ph >= the accumulator's w x h whose extra rows and columns repeat the last ones (the
// denoiser wants sides that are multiples of 16); and the FP64 sum of converged reference images across launches.
#include "render_kernels.h"

namespace trt {
namespace {

__global__ __launch_bounds__(256) void k_resolve(const float4 *__restrict__ acc, uint32_t w, uint32_t h, uint32_t pw,
                                                 uint32_t ph, float *__restrict__ color, float *__restrict__ albedo,
                                                 float *__restrict__ normal) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(pw) * ph) return;
    const uint32_t x = min(uint32_t(i % pw), w - 1), y = min(uint32_t(i / pw), h - 1);
    const float4 *p = acc + 5 * (uint64_t(y) * w + x);
    const float4 c = p[0], a = p[1], n = p[2];
    const float cnt = fmaxf(isfinite(c.w) ? c.w : 0.0f, 1.0f);
    const float f[9] = {c.x, c.y, c.z, a.x, a.y, a.z, n.x, n.y, n.z};
    float *out[3] = {color, albedo, normal};
    for (int k = 0; k < 9; k++) out[k / 3][3 * i + k % 3] = (isfinite(f[k]) ? f[k] : 0.0f) / cnt;
}

// sum[4 i + k] += the accumulator's color sum k (a non-finite sum adds zero), sum[4 i + 3] += its sample count, in
// FP64: the converged reference images are accumulated across launches here, each launch's FP32 sums added once.
__global__ __launch_bounds__(256) void k_add64(const float4 *__restrict__ acc, uint64_t n, double *__restrict__ sum) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = acc[5 * i];
    const bool fin = isfinite(c.x) && isfinite(c.y) && isfinite(c.z);
    sum[4 * i] += fin ? double(c.x) : 0.0;
    sum[4 * i + 1] += fin ? double(c.y) : 0.0;
    sum[4 * i + 2] += fin ? double(c.z) : 0.0;
    sum[4 * i + 3] += double(c.w);
}

}  // namespace

int launch_add64(const float4 *acc, uint64_t n, double *sum, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_add64, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, acc, n, sum);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_resolve(const float4 *acc, uint32_t w, uint32_t h, uint32_t pw, uint32_t ph, float *color, float *albedo,
                   float *normal, hipStream_t s) {
    const uint64_t n = uint64_t(pw) * ph;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_resolve, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, acc, w, h, pw, ph, color, albedo,
                       normal);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
