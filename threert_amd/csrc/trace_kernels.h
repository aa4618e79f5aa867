// Launch interface of the ray-batch traversal kernels (trace.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shards.h"
#include "trt_internal.h"

namespace trt {

constexpr int kTraceBlock = 256;        // threads per block: four wave64s
constexpr uint32_t kLdsStack = 24;      // stack entries per thread held in LDS (24 KiB a block)
constexpr uint32_t kRefStack = 64;      // units.rs STACK: the binary walk drops pushes past it, as the reference
constexpr uint32_t kGuardSteps = 1u << 24;  // a walk's iterations before it stops and reports (gpu.rs loop guard)
// Bits of the device's guard word, each a reason a kernel stopped a walk or a path and reported instead of reading:
constexpr uint32_t kFlagGuard = 1;   // a walk reached kGuardSteps
constexpr uint32_t kFlagWalk = 2;    // a walk met a reference outside its scene's buffers, or a stack past its room
constexpr uint32_t kFlagTable = 4;   // a shading-table index outside its table

struct TraceArgs {
    const uint4 *wnodes;      // WideNode, 4 x uint4 each
    const float4 *wunits;     // WideUnit, 4 x float4 each
    const float4 *leaves;     // Leaf, 3 x float4 each, the reference's order (the binary walk, shading)
    const float4 *wleaves;    // the same leaves in the wide trees' order, e2.w = the reference index
    const float4 *bnodes;     // Node, 2 x float4 each
    const float4 *instances;  // Instance, 10 x float4 each
    const float4 *rays;       // trt_ray, 2 x float4 each
    uint32_t *hits;           // trt_hit, 6 words each
    uint32_t *counter;        // kShards counters (shards.h): the rays each shard has handed out
    uint32_t *guard_flag;     // kFlag* bits of the walks that stopped and reported
    uint32_t *spill;          // stack entries past kLdsStack, [entry][thread]
    unsigned long long *reads;  // counting launches: nodes, triangles, units
    const uint32_t *ray_index;  // NULL: rays[i]; else rays[ray_index[i]] (hits stored at the same index)
    const uint32_t *seg_counts; // NULL: i in [0, n); else kShards segments of ray_index (the wavefront's queues):
    uint32_t seg_cap;           // segment s from s x seg_cap, seg_counts[s x kShardGap] entries
    uint64_t n;
    uint32_t spill_stride;    // threads in the grid
    uint32_t spill_entries;   // stack entries a thread has past kLdsStack
    uint32_t wnode_count, wunit_count, bnode_count, instance_count, leaf_count;  // the buffers' sizes, checked
    uint32_t top_ref;         // the wide top tree's root reference
    uint32_t top_first;       // the binary top tree's root (P.motion.z)
    uint32_t units;           // P.motion.w
    float top_box[6];
};

// The fixed per-GPU scratch of trt_trace_rays_device_noalloc (trace.hip): the shards' counters, then spilled stack
// entries.
constexpr uint64_t kScratchBytes = uint64_t{32} << 20;
constexpr uint64_t kScratchCounterBytes = kShardBytes;
constexpr int kMaxDevices = 64;
// The scratch's address on GPU `ordinal`, which must be the current device; 0 on success.
int scratch_address(int ordinal, void **ptr);

int trace_occupancy(bool wide, int *blocks_per_cu);
// watertight (wide walk only): triangles by the watertight test instead of Moller-Trumbore
int launch_trace(const TraceArgs &a, bool wide, bool count, int grid, hipStream_t s, bool watertight = false);
// n16 16-byte words from src to dst, both 16-byte aligned; at most `grid` blocks of 256
int launch_copy(void *dst, const void *src, uint64_t n16, int grid, hipStream_t s);
// FP64 re-intersection of n hits (check64.hip); at most `grid` blocks of 256
int launch_check64(const float4 *rays, const uint32_t *hits, const float4 *leaves, const float4 *units, uint64_t n,
                   uint64_t leaf_count, uint64_t unit_count, double *t64, int grid, hipStream_t s);

}  // namespace trt
