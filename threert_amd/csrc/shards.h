// Work handed out from kShards counters instead of one (device code).
//
// atomicAdd on one address is serialised across the whole GPU: on MI355X one address takes about 88 million atomics a
// second (11.4 ns each, the same from 256 or 8192 blocks), 8 addresses 570 million, 64 addresses 4,300 million. A
// persistent kernel that takes 64 rays an atomic from one counter cannot pass 5.6 billion rays a second, and one that
// adds its counts with an atomic a wave spends 11 ns of the whole GPU's time on each. So a launch's items are split
// into kShards ranges, each with its own counter in its own 256-byte line. A wave takes chunks from its own shard
// (blockIdx % kShards) and, once that is spent, from the next shard that is not: every lane reads one shard's counter,
// so a wave finds an open shard in one round trip, and each shard it finds spent is closed to it for good.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace trt {

constexpr uint32_t kShards = 32;
constexpr uint32_t kShardGap = 64;  // 32-bit words between two shards' counters (256 bytes)
constexpr uint64_t kShardBytes = uint64_t{kShards} * kShardGap * sizeof(uint32_t);
static_assert(kShards <= 32, "a wave reads every shard's counter at once; `closed` holds one bit a shard");

// The items of a launch by shard: [0, n) split evenly, or segments (segment s: seg_counts[s * kShardGap] items from
// s * seg_cap), and their counters (shard s's at ctr[s * kShardGap], counted from 0 at the shard's first item). Item
// indices stay below 2^31 (the hosts' limits on batches, tiles and queues).
struct ShardSet {
    uint32_t *ctr;
    const uint32_t *seg_counts;
    uint32_t n, seg_cap;
    template <bool kSeg>  // the segments (else [0, n) split evenly)
    __device__ __forceinline__ void range(uint32_t s, uint32_t &lo, uint32_t &hi) const {
        if (kSeg) {
            lo = s * seg_cap;
            hi = lo + seg_counts[s * kShardGap];
        } else {
            lo = static_cast<uint32_t>(uint64_t(n) * s / kShards);
            hi = static_cast<uint32_t>(uint64_t(n) * (s + 1) / kShards);
        }
    }
    __device__ __forceinline__ uint32_t *counter(uint32_t s) const { return ctr + s * kShardGap; }
};

// Called by the whole wave, every lane active: lane 0's `req` is what its atomicAdd of `chunk` on shard s returned;
// the items [a0, a1) it names, at most `chunk` and at least one. When shard s is spent the wave moves on to the next
// open shard (s changes; closed gathers the shards it found spent) and takes a chunk there. False when every shard is
// spent. kSeg: the shards are ShardSet's segments.
template <bool kSeg>
__device__ __forceinline__ bool shard_take(const ShardSet &sh, uint32_t &s, uint32_t &closed, uint32_t req,
                                           uint32_t chunk, uint32_t lane, uint32_t &a0, uint32_t &a1) {
    uint32_t c = __builtin_amdgcn_readfirstlane(req);  // lane 0's: every lane is active
    for (;;) {
        uint32_t lo, hi;
        sh.range<kSeg>(s, lo, hi);
        lo = __builtin_amdgcn_readfirstlane(lo);
        hi = __builtin_amdgcn_readfirstlane(hi);
        if (c < hi - lo) {
            a0 = lo + c;
            a1 = hi - a0 > chunk ? a0 + chunk : hi;
            return true;
        }
        closed |= 1u << s;
        bool open = false;
        if (lane < kShards && !((closed >> lane) & 1u)) {
            uint32_t ql, qh;
            sh.range<kSeg>(lane, ql, qh);
            const uint32_t v = __hip_atomic_load(sh.counter(lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            open = v < qh - ql;
        }
        const uint64_t m = __ballot(open);
        if (m == 0) return false;
        uint32_t t = s;
        for (uint32_t k = 1; k <= kShards; k++) {
            t = (s + k) % kShards;
            if ((m >> t) & 1u) break;
        }
        s = __builtin_amdgcn_readfirstlane(t);
        uint32_t v = 0;
        if (lane == 0) v = atomicAdd(sh.counter(s), chunk);
        c = __builtin_amdgcn_readfirstlane(v);
    }
}

// The kShards counters at ctr zeroed, on stream s (trace.hip)
int launch_shards_reset(uint32_t *ctr, hipStream_t s);

}  // namespace trt
