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

// Builds the own tree of each mesh from the scene's leaves in HBM (`d_leaves`, the reference's 48-byte leaves) on
// `stream`; out[i] is meshes[i]'s tree, the host build's bytes. An entry left without nodes is a mesh the device
// did not finish (a split at the median of distinct centroids, or no finite triangle): the caller builds it on the
// host. False with a reason on a device error.
bool own_trees_device(const Leaf *d_leaves, const std::vector<MeshRef> &meshes, hipStream_t stream,
                      std::vector<OwnBuilt> &out, BuildTimes *times, std::string *why);

}  // namespace gb
}  // namespace trt
