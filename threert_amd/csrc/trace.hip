// Ray-batch traversal kernels for gfx950 (wave64): the walks of walk.h over a batch of rays in HBM.
// Binary walk: persistent waves take 64 rays at a time from a counter and walk them to the end.
// Wide walk: a lane whose ray is finished takes a new one while the rest of its wave walks on. Incoherent rays differ
// several-fold in length (on the 3M-triangle workload a wave's longest ray takes about four times its mean), so a
// wave that waited for its longest ray would leave three lanes in four idle; instead, once kRefill lanes are idle
// (or all are), the wave hands out that many ray indices by lane rank from a chunk it reserved ahead with one
// atomic. Between refills the wave alternates the two phases of its walks (walk.h WideWalk). A refill stops the
// whole wave on the new rays' loads, so refills are batched (on the 3M-triangle workload, every 16 or 24 lanes ran
// 4 % faster than every 32 and 1 % faster than every 8); with an atomic per refill as well, the best batch was 32
// and 3.5 % slower. Loading each lane's next ray ahead of time hid the loads too, but took eight more registers, one
// wave per SIMD fewer (four instead of five), and was 6 % slower.
#include <hip/hip_runtime.h>

#include <mutex>

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
    Stack st{(TRT_LDS uint32_t *)(lds + threadIdx.x), a.spill, gtid, a.spill_stride, 0, kLdsStack + a.spill_entries};
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
constexpr uint32_t kChunk = 64;  // at least a wave: one refill switches chunks at most once

template <bool kCount, bool kWatertight>
__global__ __launch_bounds__(kTraceBlock) void k_trace_dyn(TraceArgs a) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
    const uint64_t n = a.n_dev ? uint64_t(*a.n_dev) : a.n;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st{(TRT_LDS uint32_t *)(lds + threadIdx.x), a.spill, gtid, a.spill_stride, 0, kLdsStack + a.spill_entries};
    uint32_t flags = 0;
    Reads cnt;
    WideWalk w;
    uint32_t ray = 0;  // the host keeps batches under 2^31 rays
    bool active = false, exhausted = false;
    const uint64_t below = (uint64_t{1} << lane) - 1;
    // Ray indices come to a wave in chunks of kChunk: [res, res_end) is being handed out, and the chunk after it was
    // requested when this one was taken (its start arrives in lane 0's next_v), so a refill waits on no atomic.
    uint32_t next_v = 0;
    if (lane == 0) next_v = atomicAdd(a.counter, 2 * kChunk);
    uint32_t res = __builtin_amdgcn_readfirstlane(__shfl(next_v, 0, 64));
    uint32_t res_end = res + kChunk;
    next_v = res + kChunk;
    for (;;) {
        // lanes without a walk take new rays, kRefill at a time (or all, when no lane walks)
        const uint64_t idle = __ballot(!active && !exhausted);
        uint64_t busy = __ballot(active);
        if (idle != 0 && (__popcll(idle) >= kRefill || busy == 0)) {
            const uint32_t need = uint32_t(__popcll(idle)), avail = res_end - res;
            const uint32_t rank = uint32_t(__popcll(idle & below));
            uint32_t i = res + rank;
            if (need <= avail) {
                res += need;
            } else {
                const uint32_t nx = __builtin_amdgcn_readfirstlane(__shfl(next_v, 0, 64));
                i = rank < avail ? i : nx + (rank - avail);
                res = nx + (need - avail);
                res_end = nx + kChunk;
                if (lane == 0) next_v = atomicAdd(a.counter, kChunk);
            }
            if (!active && !exhausted) {
                if (i < n) {
                    ray = a.ray_index ? a.ray_index[i] : uint32_t(i);
                    const float4 r0 = a.rays[2 * uint64_t(ray)], r1 = a.rays[2 * uint64_t(ray) + 1];
                    if (wide_begin(a, F3{r0.x, r0.y, r0.z}, F3{r1.x, r1.y, r1.z}, r0.w, __float_as_uint(r1.w), w, st))
                        active = true;
                    else
                        store_hit(a.hits, ray, w.h);
                } else {
                    exhausted = true;
                }
            }
            busy = __ballot(active);
        }
        if (busy == 0) {
            if (__ballot(!exhausted) == 0) break;
            continue;
        }
        // Node phase: until a kLeafShareBatch-th of the walking lanes wait, no lane has a node left, or kRefill
        // lanes have finished.
        const uint32_t walking = __popcll(busy);
        for (;;) {
            const bool at_node = active && wide_at_node(w);
            if (__ballot(at_node) == 0 || kLeafShareBatch * __popcll(__ballot(active && wide_waits(w))) >= walking ||
                __popcll(__ballot(!active && !exhausted)) >= kRefill)
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
            wide_other<kCount, kWatertight>(a, w, st, flags, cnt);
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

// The fixed scratch of the trace entry that allocates nothing (trt_trace_rays_device_noalloc): the ray counter in its
// first 256 bytes, then the stack entries past the LDS ones, [entry][thread]. One copy on every GPU the library runs
// on, held until the process ends. Host code reaches it only through the address hipGetSymbolAddress gives on that
// GPU (the array's name is a host-side stand-in there).
static __device__ __align__(256) uint32_t scratch[kScratchBytes / sizeof(uint32_t)];

// Looked up once per GPU (with that GPU current) and kept: every handle on the GPU shares its one array.
int scratch_address(int ordinal, void **ptr) {
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    *ptr = nullptr;
    if (ordinal < 0 || ordinal >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[ordinal]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        addr[ordinal] = p;
    }
    *ptr = addr[ordinal];
    return 0;
}

int trace_occupancy(bool wide, int *blocks_per_cu) {
    int b = 0;
    hipError_t e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_dyn<false, false>, kTraceBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<false, false>, kTraceBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int launch_trace(const TraceArgs &a, bool wide, bool count, int grid, hipStream_t s, bool watertight) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (wide && watertight && count)
        hipLaunchKernelGGL((k_trace_dyn<true, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide && watertight)
        hipLaunchKernelGGL((k_trace_dyn<false, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide && count)
        hipLaunchKernelGGL((k_trace_dyn<true, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide)
        hipLaunchKernelGGL((k_trace_dyn<false, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (count)
        hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(kTraceBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
