// Ray-batch traversal kernels for gfx950 (wave64): the walks of walk.h over a batch of rays in HBM.
// Binary walk: persistent waves take 64 rays at a time from the shards' counters (shards.h) and walk them to the end.
// Wide walk: a lane whose ray is finished takes a new one while the rest of its wave walks on. Incoherent rays differ
// several-fold in length (on the 3M-triangle workload a wave's longest ray takes about four times its mean), so a
// wave that waited for its longest ray would leave three lanes in four idle; instead, once kRefill lanes are idle
// (or all are), the wave hands out that many ray indices by lane rank from a chunk it reserved ahead with one
// atomic on its shard's counter. Between refills the wave alternates the two phases of its walks (walk.h WideWalk). A refill stops the
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
    const ShardSet sh{a.counter, a.seg_counts, static_cast<uint32_t>(a.n), a.seg_cap};
    uint32_t s = blockIdx.x % kShards, closed = 0;
    for (;;) {
        uint32_t req = 0;
        if (lane == 0) req = atomicAdd(sh.counter(s), 64u);
        uint32_t a0, a1;
        if (!shard_take<false>(sh, s, closed, req, 64u, lane, a0, a1)) break;
        const uint32_t k = a0 + lane;
        if (k < a1) {
            const uint64_t i = a.ray_index ? a.ray_index[k] : k;
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

// kShard: the rays come from the shards' counters (the wavefront's queue segments, shards.h) through the queues' ray
// slot lists; else from one counter, [0, n), which keeps two fewer scalars and a few registers live (93 registers
// without a hint; the sharded form needed 102, four waves a SIMD, 9 % slower on the 3M-triangle batch, whose 16.8M
// rays take a third of one counter's atomic rate). The sharded form is its own kernel, k_trace_queue, held to five
// waves a SIMD (94 registers, no spills).
template <bool kCount, bool kWatertight, bool kShard>
__device__ __forceinline__ void trace_dyn(const TraceArgs &a) {
    __shared__ uint32_t lds[kLdsStack * kTraceBlock];
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
    // requested when this one was taken (lane 0's req), so a refill waits on no atomic (until a shard runs out).
    // `more`: a shard may still hold rays. Every refill that needs more than its chunk holds either hands out rays,
    // closes a shard, or finds every shard spent (its idle lanes then stop): a lane left idle while `more` holds would
    // stop the node phase below (kRefill idle lanes end it) without a new ray.
    const ShardSet sh{a.counter, a.seg_counts, static_cast<uint32_t>(a.n), a.seg_cap};
    uint32_t s = blockIdx.x % kShards, closed = 0, res = 0, res_end = 0;
    uint32_t req = 0;
    bool more = kShard;  // one counter: a lane past n is past every ray
    if (kShard) {
        if (lane == 0) req = atomicAdd(sh.counter(s), kChunk);
        more = shard_take<true>(sh, s, closed, req, kChunk, lane, res, res_end);
        if (more && lane == 0) req = atomicAdd(sh.counter(s), kChunk);
    } else {
        if (lane == 0) req = atomicAdd(a.counter, 2 * kChunk);
        res = __builtin_amdgcn_readfirstlane(__shfl(req, 0, 64));
        res_end = res + kChunk;
        req = res + kChunk;
    }
    for (;;) {
        // lanes without a walk take new rays, kRefill at a time (or all, when no lane walks)
        const uint64_t idle = __ballot(!active && !exhausted);
        uint64_t busy = __ballot(active);
        if (idle != 0 && (__popcll(idle) >= kRefill || busy == 0)) {
            const uint32_t need = uint32_t(__popcll(idle)), avail = res_end - res;
            const uint32_t rank = uint32_t(__popcll(idle & below));
            uint32_t i = res + rank;
            bool got = rank < avail;
            if (!kShard) {
                // one counter: chunks run past n, and the lanes past it are done
                if (need > avail) {
                    const uint32_t nx = __builtin_amdgcn_readfirstlane(__shfl(req, 0, 64));
                    i = got ? i : nx + (rank - avail);
                    res = nx + (need - avail);
                    res_end = nx + kChunk;
                    if (lane == 0) req = atomicAdd(a.counter, kChunk);
                } else {
                    res += need;
                }
                got = i < a.n;
            } else if (need <= avail) {
                res += need;
            } else {
                res = res_end;
                uint32_t b0 = 0, b1 = 0;
                if (more && shard_take<true>(sh, s, closed, req, kChunk, lane, b0, b1)) {
                    // the lanes past the old chunk take from the new one (shorter at a shard's end: the rest wait)
                    const uint32_t extra = need - avail < b1 - b0 ? need - avail : b1 - b0;
                    if (!got && rank - avail < extra) {
                        i = b0 + (rank - avail);
                        got = true;
                    }
                    res = b0 + extra;
                    res_end = b1;
                    if (lane == 0) req = atomicAdd(sh.counter(s), kChunk);
                } else {
                    more = false;
                }
            }
            if (!active && !exhausted) {
                if (got) {
                    ray = kShard ? a.ray_index[i] : i;  // the queues name ray slots
                    const float4 r0 = a.rays[2 * uint64_t(ray)], r1 = a.rays[2 * uint64_t(ray) + 1];
                    if (wide_begin(a, F3{r0.x, r0.y, r0.z}, F3{r1.x, r1.y, r1.z}, r0.w, __float_as_uint(r1.w), w, st))
                        active = true;
                    else
                        store_hit(a.hits, ray, w.h);
                } else if (!more) {
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
        // (Two ballots a step: the lanes at a node and the lanes walking; the waiting and the idle lanes follow.)
        const uint32_t walking = __popcll(busy);
        const uint64_t live = __ballot(!exhausted);
        for (;;) {
            const bool at_node = active && wide_at_node(w);
            const uint64_t m_node = __ballot(at_node), m_walk = __ballot(active);
            if (m_node == 0 || kLeafShareBatch * __popcll(m_walk & ~m_node) >= walking ||
                __popcll(live & ~m_walk) >= kRefill)
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

template <bool kCount, bool kWatertight>
__global__ __launch_bounds__(kTraceBlock) void k_trace_dyn(TraceArgs a) {
    trace_dyn<kCount, kWatertight, false>(a);
}
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_trace_queue(TraceArgs a) {
    trace_dyn<false, false, true>(a);
}

__global__ void k_reset_counter(uint32_t *c) { c[threadIdx.x * kShardGap] = 0; }

}  // namespace

// The fixed scratch of the trace entry that allocates nothing (trt_trace_rays_device_noalloc): the shards' ray counters
// in its first kScratchCounterBytes, then the stack entries past the LDS ones, [entry][thread]. One copy on every GPU the library runs
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

int launch_shards_reset(uint32_t *ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(kShards), 0, s, ctr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int trace_occupancy(bool wide, int *blocks_per_cu) {
    int b = 0;
    hipError_t e = wide ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_dyn<false, false>, kTraceBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace<false, false>, kTraceBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int launch_trace(const TraceArgs &a, bool wide, bool count, int grid, hipStream_t s, bool watertight) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(kShards), 0, s, a.counter);
    if (wide && a.seg_counts)  // the wavefront's queues (no counting or watertight variant)
        hipLaunchKernelGGL(k_trace_queue, dim3(grid), dim3(kTraceBlock), 0, s, a);
    else if (wide && watertight && count)
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
