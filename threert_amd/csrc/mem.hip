// Device-to-device copy on the library's stream: 16-byte vector loads and stores, a grid-stride loop over the
// buffer. It is the plain streaming kernel whose HBM counter readings (bytes fetched and written per byte copied)
// scale the same counters read on the walk's kernels.
#include "trace_kernels.h"

namespace trt {
namespace {

__global__ __launch_bounds__(256) void k_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

int launch_copy(void *dst, const void *src, uint64_t n16, int grid, hipStream_t s) {
    if (n16 == 0) return 0;
    const uint64_t need = (n16 + 255) / 256;
    const int g = int(need < uint64_t(grid) ? need : uint64_t(grid));
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(256), 0, s, static_cast<uint4 *>(dst), static_cast<const uint4 *>(src),
                       n16);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
