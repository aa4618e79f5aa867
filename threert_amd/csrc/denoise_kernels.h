// Launch interface of the a-trous denoiser (denoise.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace trt {

// scratch: 9 floats a pixel (two demodulated images, two variances, the filtered variance).
int launch_atrous(uint32_t w, uint32_t h, const float *color, const float *albedo, const float *normal,
                  const float *variance, uint32_t iterations, float *out, float *scratch, hipStream_t s);

}  // namespace trt
