// Launch interface of the path tracer (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "trace_kernels.h"

namespace trt {

constexpr int kRenderBlock = 256;
// The render launches' ray counts (closest, shadow), one pair a shard (shards.h): shard s's pair at
// counts[s x kRayShardGap], 256 bytes apart; the total is their sum.
constexpr uint32_t kRayShardGap = 32;
constexpr uint64_t kRayCountBytes = uint64_t{kShards} * kRayShardGap * sizeof(unsigned long long);
constexpr int kParamWords = 156;  // Params: 39 vec4 (gpu.rs:129-185)

struct RenderArgs {
    TraceArgs geo;              // the wide walk's buffers (rays / hits unused)
    const float4 *instances;    // 160-byte records, 10 x float4
    const float4 *vertices;     // indexed.rs Vertex, 3 x float4
    const uint32_t *indices;    // 3 a triangle, the leaves' order
    const uint32_t *rows;       // scene.rs Row, 4 words
    const float4 *materials;    // scene.rs GpuMaterial, 18 x float4
    const float4 *lights;       // scene.rs GpuLight, 5 x float4
    const float4 *emitters;     // lib.rs emitter_cdf: entries, guide, header
    const float4 *emitter_tris; // lib.rs GpuEmitterTri, 4 x float4
    const float4 *objects;      // 3 x float4 an object
    const float4 *energy;       // lib.rs energy_lut, 512 x float4
    const float4 *light_tree;   // ltree.rs LightTree nodes, one float4 each
    const float4 *glass_energy; // lib.rs glass_energy_lut, 1024 x float4 (a glass scene's; NULL otherwise)
    const uint32_t *tex_pages;  // texpack.rs pages: RGBA8 texels, per layer levels 0..4 (a textured scene's)
    const float4 *tex_info;     // texpack.rs Atlas::info, 2 x float4 a texture
    uint64_t tex_info_count;    // its float4s
    uint32_t tex_side, tex_layers;
    const uint2 *env;           // gpu.rs env_texture: RGBA half floats, env_w x env_h (a scene with an environment map)
    const float *env_cdf;       // lib.rs build_env_cdf: env_h + env_w x env_h floats
    uint64_t env_cdf_count;
    uint32_t env_w, env_h;
    const uint32_t *vcolors;    // lib.rs gpu_vcolors: 3 RGBA8 sRGB corner colours a triangle, in the leaves' order
    uint64_t vcolor_words;
    float4 *accum;              // records.wgsl Pixel, 5 x float4 a pixel
    // the tables' sizes (in their records), against which every index is checked
    uint32_t vertex_count, tri_count, row_count, material_count, light_count, emitter_vec4s, emitter_tri_count,
        object_count, light_tree_count;
    uint32_t first_sample;
    uint32_t spp;
    uint32_t params[kParamWords];
};

// One job of a launch: its arguments and where its 8x8 tiles start among the launch's.
struct RenderJob {
    RenderArgs args;
    uint32_t tile_first;
    uint32_t pad[3];
};

// ---- the wavefront path tracer (render.hip k_wf_*) ----------------------------------------------------------------
// A chunk's paths (one per pixel and sample) live in structure-of-arrays state; each bounce the walk kernel traces
// every path's next closest ray and the previous bounce's shadow rays together, from a compacted list of ray slots,
// and the shade kernel resolves those shadow rays, shades the hit and queues the next rays.
constexpr uint32_t kWfShadow = 5;             // shadow rays one bounce can leave pending (directional lights + 2)
constexpr uint32_t kWfSlots = kWfShadow + 1;  // a path's ray slots: 0 the next closest ray, then its shadow rays
constexpr uint32_t kWfMaxSuns = kWfShadow - 2;  // directional lights the wavefront takes (more: the megakernel)

// A run of a chunk's paths: `rounds` samples (sample_first on) of pixels [pixel_first, pixel_first + count) of one job,
// path path_first + r x count + j for sample r and pixel j; acc_first: the run's first pixel among the chunk's runs'.
struct WfSub {
    uint32_t job, pixel_first, count, sample_first, rounds, path_first, acc_first, width;
};

struct WfArgs {
    const RenderJob *jobs;
    const WfSub *subs;
    uint32_t nsubs, npaths;  // the chunk's runs and paths
    uint32_t stride;         // entries in each state array and each ray slot
    uint32_t cap;            // paths in a queue segment (stride / kShards): path p is in segment p / cap
    float4 *thr;             // xyz throughput, w the last BSDF pdf
    float4 *rad;             // xyz radiance, w the path's flag bits
    float4 *alb, *nrm;       // xyz the camera hit's albedo and normal
    float4 *misx, *misn;     // xyz the vertex direct light last sampled from and its normal (emit_mis)
    float4 *thr_dl;          // xyz the throughput the pending shadow rays' contributions are weighed by
    float4 *contrib;         // [kWfShadow][stride]: xyz a pending direct-light contribution
    float4 *rays;            // [kWfSlots][stride] trt_ray
    uint32_t *hits;          // [kWfSlots][stride] trt_hit
    uint32_t *list[2];       // ray-slot indices (slot x stride + path) by bounce parity, segment s from s x 6 x cap
    uint32_t *active[2];     // the paths the shade kernel takes, by bounce parity, segment s from s x cap
    uint32_t *counts;        // the segments' lengths: [bounce % 3][ray list, active list][kShards][kShardGap]
    uint32_t *guard;
    unsigned long long *rays_traced;  // the sharded ray counts (kRayShardGap)
};
int wf_shade_occupancy(int *blocks_per_cu);
// the chunk's counts set, and every path's camera ray queued (list 0, active 0)
int launch_wf_begin(const WfArgs &a, int grid, hipStream_t s);
// bounce k's shade kernel, after the walk of list k % 2: reads active list k % 2, queues into the lists (k + 1) % 2
int launch_wf_shade(const WfArgs &a, uint32_t k, bool coat, bool tree, int grid, hipStream_t s);
// every path's radiance, albedo and normal added to its pixel's record, sample after sample; pixels: the runs' pixels
int launch_wf_accum(const WfArgs &a, uint32_t pixels, int grid, hipStream_t s);

int render_occupancy(int *blocks_per_cu);
// accum.hip: an accumulator's color sums (non-finite ones as zero) and count added to an FP64 image, 4 doubles a pixel
int launch_add64(const float4 *acc, uint64_t n, double *sum, hipStream_t s);
// accum.hip: the accumulator's color / albedo / normal means into pw x ph images (edge rows and columns repeated)
int launch_resolve(const float4 *acc, uint32_t w, uint32_t h, uint32_t pw, uint32_t ph, float *color, float *albedo,
                   float *normal, hipStream_t s);
// d_jobs: njobs jobs in device memory, tile_first ascending from 0; total_tiles: all their tiles; counter: the kShards
// tile counters. rays: the sharded ray counts the launch adds its traced rays to (closest; shadow and any-hit); coat: some job's scene has a
// material with a clearcoat (the kernel with the coat's terms; the same bits for scenes without one); tree: some
// job's Params ask for the light tree (the kernel with its code); ext: some job's scene has glass, textures or local
// lights (the kernel with their code; a job's glass_energy and tex_info, null without them, and its Params say which it
// has); reads: NULL, or four counters a counting launch adds to (wide nodes, triangles, units its walks read; hits it
// shades)
int launch_render(const RenderJob *d_jobs, uint32_t njobs, uint32_t total_tiles, uint32_t *counter,
                  uint32_t *spill, uint32_t spill_stride, uint32_t spill_entries, uint32_t *guard,
                  unsigned long long *rays, bool coat, bool tree, bool ext, unsigned long long *reads, int grid,
                  hipStream_t s);

}  // namespace trt
