// Ray-batch traversal kernels for gfx950 (wave64): the walks of walk.h over a batch of rays in HBM.
// Binary walk: persistent waves take 64 rays at a time from a counter and walk them to the end.
// Wide walk: a lane whose ray is finished takes a new one while the rest of its wave walks on. Incoherent rays differ
// several-fold in length (on the 3M-triangle workload a wave's longest ray takes about four times its mean), so a
// wave that waited for its longest ray would leave three lanes in four idle. Each lane holds its next ray already
// loaded and starts it as soon as its walk ends; once kRefill lanes have used theirs (or no lane walks), the wave
// takes that many ray indices with one atomic, hands them out by lane rank and loads the rays while it walks on.
// Between refills the wave alternates the two phases of its walks (walk.h WideWalk). On the
// 3M-triangle workload, refilling every 16 lanes cost 18 % against every 24 (more stops of the whole wave), and
// loading each ray only when its lane needed it cost 3 %.
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
    Stack st{(TRT_LDS uint32_t *)(lds + threadIdx.x), a.spill + gtid, a.spill_stride, 0, kLdsStack + a.spill_entries};
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

constexpr uint32_t kRefill = 24;

template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_trace_dyn(TraceArgs a) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st{(TRT_LDS uint32_t *)(lds + threadIdx.x), a.spill + gtid, a.spill_stride, 0, kLdsStack + a.spill_entries};
    uint32_t flags = 0;
    Reads cnt;
    WideWalk w;
    uint64_t ray = 0;
    // Each lane holds its next ray, loaded ahead (index nidx, data n0 and n1), so a lane whose walk ends starts
    // the next one at once, and the wave never waits on a ray load.
    constexpr uint32_t kNoRay = 0xffffffffu;
    uint32_t nidx = kNoRay;
    float4 n0{}, n1{};
    bool active = false, exhausted = false;
    const uint64_t below = (uint64_t{1} << lane) - 1;
    for (;;) {
        if (!active && nidx != kNoRay) {
            ray = nidx;
            nidx = kNoRay;
            if (wide_begin(a, F3{n0.x, n0.y, n0.z}, F3{n1.x, n1.y, n1.z}, n0.w, __float_as_uint(n1.w), w, st))
                active = true;
            else
                store_hit(a.hits, ray, w.h);
        }
        // lanes that used their loaded ray take new indices, kRefill at a time (or all, when no lane walks)
        const uint64_t need = __ballot(nidx == kNoRay && !exhausted);
        const uint64_t busy = __ballot(active);
        if (need != 0 && (__popcll(need) >= kRefill || busy == 0)) {
            uint32_t first = 0;
            if (lane == 0) first = atomicAdd(a.counter, uint32_t(__popcll(need)));
            first = __builtin_amdgcn_readfirstlane(__shfl(first, 0, 64));
            if (nidx == kNoRay && !exhausted) {
                const uint64_t i = uint64_t(first) + uint64_t(__popcll(need & below));
                if (i < a.n) {
                    nidx = uint32_t(i);
                    n0 = a.rays[2 * i];
                    n1 = a.rays[2 * i + 1];
                } else {
                    exhausted = true;
                }
            }
        }
        if (busy == 0) {
            if (__ballot(nidx != kNoRay || !exhausted) == 0) break;
            continue;
        }
        // Node phase: until a kLeafShareBatch-th of the walking lanes wait, no lane has a node left, or kRefill
        // lanes have finished.
        const uint32_t walking = __popcll(busy);
        for (;;) {
            const bool at_node = active && wide_at_node(w);
            if (__ballot(at_node) == 0 || kLeafShareBatch * __popcll(__ballot(active && wide_waits(w))) >= walking ||
                __popcll(__ballot(!active && nidx != kNoRay)) >= kRefill)
                break;
            if (at_node) {
                wide_node<kCount>(a, w, st, flags, cnt);
                if (wide_over(w)) {
                    store_hit(a.hits, ray, w.h);
                    active = false;
                }
            }
        }
        // The other phase: every waiting lane tests its run of triangles, or enters or leaves a unit.
        if (active && wide_waits(w)) {
            wide_other<kCount>(a, w, st, flags, cnt);
            if (wide_over(w)) {
                store_hit(a.hits, ray, w.h);
                active = false;
            }
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
    hipError_t e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_dyn<false>, kTraceBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<false, false>, kTraceBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int launch_trace(const TraceArgs &a, bool wide, bool count, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (wide && count)
        hipLaunchKernelGGL((k_trace_dyn<true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide)
        hipLaunchKernelGGL((k_trace_dyn<false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (count)
        hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
