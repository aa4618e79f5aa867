// The device build of the library's own mesh trees (gbuild.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "trt_internal.h"

namespace trt {
namespace gb {

struct BuildTimes {
    float device_ms = 0;  // first kernel to last, on the stream's clock (the host's waits between levels included)
    uint32_t levels = 0;
    uint64_t nodes = 0;
};

struct DeviceWide {
    // complete: every mesh was built and collapsed on the device; d_wnodes holds the meshes' wide nodes (mesh after
    // mesh, as the host appends them: wnode_count of them; stream-ordered memory of the build's stream, for
    // hipFreeAsync there) and d_wleaves every leaf in the wide trees' order (hipFree); the caller owns both. mesh_ref / mesh_bound: each mesh's root reference and stack bound.
    bool complete = false;
    void *d_wnodes = nullptr;
    uint64_t wnode_count = 0;
    void *d_wleaves = nullptr;
    std::vector<uint32_t> mesh_ref, mesh_bound;
    // not complete: the own trees the device built (trees[i] for meshes[i]); an entry without nodes is a mesh left
    // to the host (a split at the median of distinct centroids, or no finite triangle)
    std::vector<OwnBuilt> trees;
    BuildTimes times;
};

// Builds the library's own tree of each mesh from the scene's leaves in HBM (`d_leaves`, `leaf_count` of the
// reference's 48-byte leaves) on `stream`, the host build's bytes, and when the device builds them all, collapses
// and quantises them and lays out the leaves on the device too. False with a reason on a device error, or on a
// tree the host would refuse too (a box that cannot be quantised, a mesh past 2^26 triangles).
bool device_wide(const Leaf *d_leaves, uint64_t leaf_count, const std::vector<MeshRef> &meshes, hipStream_t stream,
                 DeviceWide &res, std::string *why);

}  // namespace gb
}  // namespace trt
