// Ray-batch traversal kernels for gfx950 (wave64): the walks of walk.h over a batch of rays in HBM.
// Persistent: a wave takes 64 rays at a time from a counter until the batch is done.
#include <hip/hip_runtime.h>

#include "trace_kernels.h"
#include "walk.h"

namespace trt {

namespace {
using namespace walk;

template <bool kWide, bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st{lds + threadIdx.x, a.spill + gtid, a.spill_stride, 0, kLdsStack + a.spill_entries};
    uint32_t flags = 0;
    Reads cnt;
    for (;;) {
        uint32_t first_ray = 0;
        if (lane == 0) first_ray = atomicAdd(a.counter, 64u);
        first_ray = __shfl(first_ray, 0, 64);
        if (uint64_t(first_ray) >= a.n) break;
        const uint64_t i = uint64_t(first_ray) + lane;
        if (i < a.n) {
            const float4 r0 = a.rays[2 * i], r1 = a.rays[2 * i + 1];
            const F3 o{r0.x, r0.y, r0.z}, d{r1.x, r1.y, r1.z};
            const uint32_t mode = __float_as_uint(r1.w);
            const HitRec h = kWide ? trace_wide<kCount>(a, o, d, r0.w, mode, st, flags, cnt)
                                   : trace_binary<kCount>(a, o, d, r0.w, mode, st, flags, cnt);
            store_hit(a.hits, i, h);
        }
    }
    if (flags) atomicOr(a.guard_flag, flags);
    if (kCount) {
        atomicAdd(a.reads + 0, (unsigned long long)cnt.nodes);
        atomicAdd(a.reads + 1, (unsigned long long)cnt.tris);
        atomicAdd(a.reads + 2, (unsigned long long)cnt.units);
    }
}

__global__ void k_reset_counter(uint32_t *c) { *c = 0; }

}  // namespace

int trace_occupancy(bool wide, int *blocks_per_cu) {
    int b = 0;
    hipError_t e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<true, false>, kTraceBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<false, false>, kTraceBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int launch_trace(const TraceArgs &a, bool wide, bool count, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (wide && count)
        hipLaunchKernelGGL((k_trace<true, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide)
        hipLaunchKernelGGL((k_trace<true, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (count)
        hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
