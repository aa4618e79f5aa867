// The C ABI of libthreert (include/threert.h): devices, memory, host tree builders, scene upload, ray batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <new>
#include <string>

#include "../../include/threert.h"
#include "denoise_kernels.h"
#include "render_kernels.h"
#include "trace_kernels.h"
#include "gbuild.h"
#include "trt_internal.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}
int ok() {
    g_error.clear();
    return TRT_OK;
}

#define TRT_HIP(call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? TRT_E_OOM : TRT_E_DEVICE,                          \
                        std::string(#call) + ": " + hipGetErrorString(e_));                            \
    } while (0)

}  // namespace

constexpr int kJobRing = 8;
struct JobRing {
    void *host[kJobRing] = {};
    void *dev[kJobRing] = {};
    uint64_t cap[kJobRing] = {};
    hipEvent_t ev[kJobRing] = {};
    int next = 0;
};

struct trt_device {
    int ordinal = 0;
    JobRing jobs;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t *counter = nullptr;
    uint32_t *guard = nullptr;
    uint32_t *spill = nullptr;
    uint64_t spill_bytes = 0;
    unsigned long long *reads = nullptr;
    unsigned long long *ray_counts = nullptr;  // kShards pairs (closest, shadow), kRayShardGap apart
    void *denoise_scratch = nullptr;  // trt_denoise_atrous' intermediate images
    uint64_t denoise_bytes = 0;
    void *staging = nullptr;  // trt_trace_rays' device copies of the caller's rays and hits
    uint64_t staging_bytes = 0;
    uint32_t *scratch = nullptr;  // this GPU's fixed scratch (trace.hip), trt_trace_rays_device_noalloc's only memory
    int cus = 0;
    int blocks_wide = 1, blocks_binary = 1, blocks_render = 1, blocks_shade = 1;
    std::atomic<uint32_t> build{TRT_BUILD_AUTO};  // where uploads build the library's own mesh trees
    std::atomic<uint32_t> render{TRT_RENDER_AUTO};  // how render launches run (trt_device_set_render)
    void *wf = nullptr;    // the wavefront's path state, rays, hits and queues for wf_stride paths
    uint64_t wf_stride = 0;
    hipDeviceProp_t props{};
    std::mutex mu;
    std::atomic<int> refs{1};
};

// d->reads: [0..2] a counting trace launch's reads, [kRenderReads .. + 3] a counting render launch's reads (nodes,
// triangles, units, shaded hits). The rays the render launches traced are in d->ray_counts, sharded (kRayShardGap).
constexpr int kRenderReads = 8;

struct trt_scene {
    trt_device *dev = nullptr;
    void *nodes = nullptr, *leaves = nullptr, *instances = nullptr, *wnodes = nullptr, *wunits = nullptr,
         *wleaves = nullptr;
    // shading tables (trt_render)
    void *vertices = nullptr, *indices = nullptr, *rows = nullptr, *materials = nullptr, *lights = nullptr,
         *emitters = nullptr, *emitter_tris = nullptr, *objects = nullptr, *energy = nullptr, *light_tree = nullptr,
         *glass_energy = nullptr, *tex_pages = nullptr, *tex_info = nullptr, *env = nullptr, *env_cdf = nullptr,
         *vcolors = nullptr;
    uint64_t light_tree_count = 0, tex_info_vec4s = 0, env_cdf_count = 0, vcolor_words = 0;
    uint32_t tex_side = 0, tex_layers = 0, env_w = 0, env_h = 0;  // env_w 0: no environment map
    bool tex = false;    // the scene has texture pages (the reference's F_TEX)
    bool shading = false;
    bool coat = false;   // a material has a clearcoat (layer.z > 0: the reference's F_COAT, gpu.rs features_of)
    bool glass = false;  // a material has glass (layer.x > 0: the reference's F_GLASS, gpu.rs:2030-2032)
    bool unlit = false;  // a material is unlit (tex1.z MAT_UNLIT: the reference's F_UNLIT, gpu.rs:2037)
    uint64_t light_count = 0, emitter_entries = 0, emitter_tri_count = 0, n_dir = 0;  // n_dir: directional lights
    uint64_t vertex_count = 0, tri_count = 0, row_count = 0, material_count = 0, object_count = 0, emitter_vec4s = 0;
    uint64_t node_count = 0, leaf_count = 0, instance_count = 0, wnode_count = 0, wunit_count = 0;
    uint32_t top_first = 0, units = 0, top_ref = trt::kNone, max_stack = 0;
    float top_box[6] = {0, 0, 0, 0, 0, 0};
    trt_scene_stats stats{};
};

extern "C" {

const char *trt_last_error(void) { return g_error.c_str(); }

int trt_device_create(int ordinal, trt_device **out) {
    if (!out) return fail(TRT_E_INVALID, "trt_device_create: out is null");
    *out = nullptr;
    int count = 0;
    TRT_HIP(hipGetDeviceCount(&count));
    if (ordinal < 0 || ordinal >= count)
        return fail(TRT_E_INVALID, "trt_device_create: no HIP device " + std::to_string(ordinal));
    TRT_HIP(hipSetDevice(ordinal));
    auto *d = new (std::nothrow) trt_device();
    if (!d) return fail(TRT_E_OOM, "trt_device_create: out of host memory");
    d->ordinal = ordinal;
    hipError_t e = hipGetDeviceProperties(&d->props, ordinal);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&d->ev0);
    if (e == hipSuccess) e = hipEventCreate(&d->ev1);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->counter), trt::kShardBytes);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->guard), 256);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->reads), 256);
    if (e == hipSuccess) e = hipMemset(d->guard, 0, 256);
    if (e == hipSuccess) e = hipMemset(d->reads, 0, 256);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->ray_counts), trt::kRayCountBytes);
    if (e == hipSuccess) e = hipMemset(d->ray_counts, 0, trt::kRayCountBytes);
    if (e != hipSuccess) {
        trt_device_release(d);
        return fail(TRT_E_DEVICE, std::string("trt_device_create: ") + hipGetErrorString(e));
    }
    void *scratch = nullptr;
    if (trt::scratch_address(ordinal, &scratch)) {
        trt_device_release(d);
        return fail(TRT_E_DEVICE, "trt_device_create: the fixed scratch's device address could not be found");
    }
    d->scratch = static_cast<uint32_t *>(scratch);
    d->cus = d->props.multiProcessorCount;
    if (trt::trace_occupancy(true, &d->blocks_wide) || d->blocks_wide < 1) d->blocks_wide = 1;
    if (trt::trace_occupancy(false, &d->blocks_binary) || d->blocks_binary < 1) d->blocks_binary = 1;
    if (trt::render_occupancy(&d->blocks_render) || d->blocks_render < 1) d->blocks_render = 1;
    if (trt::wf_shade_occupancy(&d->blocks_shade) || d->blocks_shade < 1) d->blocks_shade = 1;
    *out = d;
    return ok();
}

// A device lives while its handle or any scene or accumulator made on it does: the last release frees it.
static void device_unref(trt_device *d) {
    if (--d->refs > 0) return;
    hipSetDevice(d->ordinal);
    if (d->stream) hipStreamSynchronize(d->stream);
    hipFree(d->counter);
    hipFree(d->guard);
    hipFree(d->reads);
    hipFree(d->ray_counts);
    hipFree(d->denoise_scratch);
    hipFree(d->spill);
    hipFree(d->staging);
    hipFree(d->wf);
    for (int k = 0; k < kJobRing; k++) {
        hipHostFree(d->jobs.host[k]);
        hipFree(d->jobs.dev[k]);
        if (d->jobs.ev[k]) hipEventDestroy(d->jobs.ev[k]);
    }
    if (d->ev0) hipEventDestroy(d->ev0);
    if (d->ev1) hipEventDestroy(d->ev1);
    if (d->stream) hipStreamDestroy(d->stream);
    delete d;
}

int trt_device_release(trt_device *d) {
    if (!d) return ok();
    device_unref(d);
    return ok();
}
}  // extern "C"

// Accessors for the library's other translation units (unet.cpp).
hipStream_t trt_device_stream(trt_device *d) { return d->stream; }
int trt_device_ordinal(trt_device *d) { return d->ordinal; }
void trt_device_lock(trt_device *d, bool lock) {
    if (lock)
        d->mu.lock();
    else
        d->mu.unlock();
}
void trt_device_ref(trt_device *d, int delta) {
    if (delta > 0)
        d->refs += delta;
    else
        device_unref(d);
}
int trt_fail(int code, const std::string &msg) { return fail(code, msg); }
int trt_ok() { return ok(); }

extern "C" {

int trt_device_info_get(trt_device *d, trt_device_info *out) {
    if (!d || !out) return fail(TRT_E_INVALID, "trt_device_info_get: null argument");
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", d->props.gcnArchName);
    out->compute_units = static_cast<uint32_t>(d->props.multiProcessorCount);
    out->wave_size = static_cast<uint32_t>(d->props.warpSize);
    out->hbm_bytes = d->props.totalGlobalMem;
    out->clock_khz = static_cast<uint32_t>(d->props.clockRate);
    out->gcn_arch = static_cast<uint32_t>(std::strtoul(d->props.gcnArchName + 3, nullptr, 10));
    return ok();
}

int trt_device_sync(trt_device *d) {
    if (!d) return fail(TRT_E_INVALID, "trt_device_sync: null device");
    uint32_t g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    {
        // Read and clear the guard words under the device's lock, on its stream: no other thread's launch can land
        // between the read and the clear, so a report reaches the first sync after the launch that raised it.
        std::lock_guard<std::mutex> lock(d->mu);
        TRT_HIP(hipSetDevice(d->ordinal));
        TRT_HIP(hipStreamSynchronize(d->stream));
        TRT_HIP(hipMemcpyAsync(g, d->guard, sizeof(g), hipMemcpyDeviceToHost, d->stream));
        TRT_HIP(hipStreamSynchronize(d->stream));
        if (g[0]) {
            TRT_HIP(hipMemsetAsync(d->guard, 0, sizeof(g), d->stream));
            TRT_HIP(hipStreamSynchronize(d->stream));
        }
    }
    if (g[0]) {
        // the first out-of-range read (walk.h report_range ids)
        static const char *what[] = {"?", "binary top node", "binary node", "binary children", "binary stack",
                                     "unit record", "binary leaf run", "wide node", "wide stack", "wide unit",
                                     "wide leaf run"};
        static const char *table[] = {"instance", "material", "light", "emitter entry", "row", "object",
                                      "triangle indices", "vertex", "energy table", "emitter triangle (RIS candidate)",
                                      "leaf", "emitter triangle (RIS pick)", "emitter triangle (single draw)",
                                      "emitter triangle (a hit's own)", "light tree node", "glass energy table",
                                      "texture page layer", "texture info", "environment texel",
                                      "environment CDF", "vertex colours"};
        std::string where;
        if (g[1]) {
            const char *name = g[1] <= 10 ? what[g[1]] : (g[1] >= 16 && g[1] <= 36 ? table[g[1] - 16] : "?");
            where = std::string(" (first: ") + name + " index " +
                    std::to_string((uint64_t(g[3]) << 32) | g[2]) + " of " + std::to_string(g[4]) + "; context " +
                    std::to_string(g[5]) + " " + std::to_string(g[6]) + " " + std::to_string(g[7]) + ")";
        }
        if (g[0] & trt::kFlagWalk)
            return fail(TRT_E_GUARD,
                        "a walk met a reference outside its scene's buffers or a stack past its room" + where);
        if (g[0] & trt::kFlagTable) return fail(TRT_E_GUARD, "a shading-table index was outside its table" + where);
        return fail(TRT_E_GUARD, "a walk stopped at its iteration guard");
    }
    return ok();
}

int trt_mem_alloc(trt_device *d, uint64_t bytes, void **p) {
    if (!d || !p) return fail(TRT_E_INVALID, "trt_mem_alloc: null argument");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipMalloc(p, bytes ? bytes : 1));
    return ok();
}
int trt_mem_free(trt_device *d, void *p) {
    if (!d) return fail(TRT_E_INVALID, "trt_mem_free: null device");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipFree(p));
    return ok();
}
int trt_mem_upload(trt_device *d, void *dst, const void *src, uint64_t bytes) {
    if (!d || (bytes && (!dst || !src))) return fail(TRT_E_INVALID, "trt_mem_upload: null argument");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipStreamSynchronize(d->stream));
    TRT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ok();
}
int trt_mem_download(trt_device *d, void *dst, const void *src, uint64_t bytes) {
    if (!d || (bytes && (!dst || !src))) return fail(TRT_E_INVALID, "trt_mem_download: null argument");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipStreamSynchronize(d->stream));
    TRT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ok();
}

int trt_mem_copy(trt_device *d, void *dst, const void *src, uint64_t bytes) {
    if (!d || (bytes && (!dst || !src))) return fail(TRT_E_INVALID, "trt_mem_copy: null argument");
    if ((bytes | reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
        return fail(TRT_E_INVALID, "trt_mem_copy: pointers and size must be multiples of 16 bytes");
    TRT_HIP(hipSetDevice(d->ordinal));
    if (trt::launch_copy(dst, src, bytes / 16, d->cus * 8, d->stream) != 0)
        return fail(TRT_E_DEVICE, "trt_mem_copy: launch failed");
    return ok();
}

int trt_timer_start(trt_device *d) {
    if (!d) return fail(TRT_E_INVALID, "trt_timer_start: null device");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipEventRecord(d->ev0, d->stream));
    return ok();
}
int trt_timer_stop(trt_device *d, float *ms) {
    if (!d || !ms) return fail(TRT_E_INVALID, "trt_timer_stop: null argument");
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipEventRecord(d->ev1, d->stream));
    TRT_HIP(hipEventSynchronize(d->ev1));
    TRT_HIP(hipEventElapsedTime(ms, d->ev0, d->ev1));
    return ok();
}

// ---- builders ----------------------------------------------------------------------------------------------------
static int export_tree(trt::Tree &t, trt_tree *out) {
    std::memset(out, 0, sizeof(*out));
    out->node_count = t.nodes.size();
    out->order_count = t.order.size();
    out->nodes = std::malloc(std::max<size_t>(1, t.nodes.size()) * sizeof(trt::Node));
    out->order = static_cast<uint32_t *>(std::malloc(std::max<size_t>(1, t.order.size()) * sizeof(uint32_t)));
    if (!out->nodes || !out->order) {
        std::free(out->nodes);
        std::free(out->order);
        out->nodes = nullptr, out->order = nullptr;
        return fail(TRT_E_OOM, "out of host memory");
    }
    if (!t.nodes.empty()) std::memcpy(out->nodes, t.nodes.data(), t.nodes.size() * sizeof(trt::Node));
    if (!t.order.empty()) std::memcpy(out->order, t.order.data(), t.order.size() * sizeof(uint32_t));
    out->depth = t.depth;
    std::memcpy(out->lo, t.lo, 12);
    std::memcpy(out->hi, t.hi, 12);
    return ok();
}

int trt_build_tris(const float *tris, uint64_t n, uint32_t max_depth, trt_tree *out) {
    if (!out || (n && !tris)) return fail(TRT_E_INVALID, "trt_build_tris: null argument");
    trt::Tree t;
    if (!trt::build_tris(tris, n, max_depth, t))
        return fail(TRT_E_INVALID, std::to_string(n) + " items do not fit a tree " + std::to_string(max_depth) + " deep");
    return export_tree(t, out);
}
int trt_build_mesh_tree(const float *tris, uint64_t n, trt_tree *out) {
    if (!out || (n && !tris)) return fail(TRT_E_INVALID, "trt_build_mesh_tree: null argument");
    trt::Tree t;
    std::string why;
    if (!trt::mesh_tree(tris, n, t, &why)) return fail(TRT_E_INVALID, why);
    return export_tree(t, out);
}
int trt_build_sah(const float *boxes, uint64_t n, uint32_t max_depth, trt_tree *out) {
    if (!out || (n && !boxes)) return fail(TRT_E_INVALID, "trt_build_sah: null argument");
    if (max_depth < 64 && n > (uint64_t{1} << max_depth))
        return fail(TRT_E_INVALID, "more boxes than a tree of that depth holds");
    trt::Tree t;
    trt::build_sah(boxes, n, max_depth, t);
    return export_tree(t, out);
}
int trt_build_median(const float *boxes, uint64_t n, uint32_t max_leaf, trt_tree *out) {
    if (!out || (n && !boxes) || max_leaf == 0) return fail(TRT_E_INVALID, "trt_build_median: bad argument");
    trt::Tree t;
    trt::build_median(boxes, n, max_leaf, t);
    return export_tree(t, out);
}
int trt_build_top_tree(const float *boxes, const uint32_t *records, uint64_t n, trt_tree *out) {
    if (!out || (n && (!boxes || !records))) return fail(TRT_E_INVALID, "trt_build_top_tree: null argument");
    trt::Tree t;
    std::string why;
    if (!trt::top_tree(boxes, records, n, t, &why)) return fail(TRT_E_INVALID, why);
    return export_tree(t, out);
}
void trt_unit_box(const float *world, const float *lo, const float *hi, float *out) {
    trt::unit_box(world, lo, hi, out);
}
void trt_unit_record(void *record, uint32_t root, int cast_shadow) {
    trt::unit_record(*static_cast<trt::Instance *>(record), root, cast_shadow != 0);
}
void trt_tree_free(trt_tree *t) {
    if (!t) return;
    std::free(t->nodes);
    std::free(t->order);
    t->nodes = nullptr, t->order = nullptr;
    t->node_count = t->order_count = 0;
}

// ---- scene ---------------------------------------------------------------------------------------------------------
int trt_scene_release(trt_scene *s) {
    if (!s) return ok();
    hipSetDevice(s->dev->ordinal);
    hipStreamSynchronize(s->dev->stream);
    hipFree(s->nodes);
    hipFree(s->leaves);
    hipFree(s->instances);
    hipFree(s->wnodes);
    hipFree(s->wunits);
    hipFree(s->wleaves);
    for (void *p : {s->vertices, s->indices, s->rows, s->materials, s->lights, s->emitters, s->emitter_tris,
                    s->objects, s->energy, s->light_tree, s->glass_energy, s->tex_pages, s->tex_info, s->env,
                    s->env_cdf, s->vcolors})
        hipFree(p);
    trt_device *d = s->dev;
    delete s;
    device_unref(d);
    return ok();
}

// A material the path loop would render without what it asks for is refused, named: the reference's `surface`
// (pathtrace.wgsl:1320-1425) reads these only under the F_VCOLOR, F_FUNC and F_UNLIT features, and its cutout paths
// (1300-1310), none of which this core has; glass (1542-1640) needs the glass energy table, a texture slot (F_TEX,
// 1360-1425) a texture of the scene's pages. (The occlusion slot, tex1.y, is read by no path of the kernel: it is
// taken and ignored, as the reference's.) Words of the 72-word GpuMaterial (scene.rs:480-491): tex0 12-15 and tex1
// 16-19 (texture slots, -1 none; tex1.z the flags, tex1.w a function + 1), uv.z 22 (transmission texture + 1), uv.w
// 23 (vertex colours), uv2.w 27 (alpha cutoff), layer.x 28 (transmission).
static bool material_rendered(const float *m, uint64_t count, bool glass_table, uint64_t textures, std::string *why) {
    static const char *slot[5] = {"base-colour texture", "normal map", "roughness texture", "metalness texture",
                                  "emissive texture"};
    for (uint64_t k = 0; k < count; k++) {
        const float *f = m + 72 * k;
        int32_t w[8];
        std::memcpy(w, f + 12, sizeof(w));
        // texture slots: each a texture of the pages (the transmission map's slot + 1 in uv.z)
        const int64_t tslot = f[22] > 0.5f ? (f[22] < 2147483648.0f ? int64_t(f[22]) - 1 : -2) : -1;
        for (int j = 0; j < 6; j++) {
            const int64_t id = j < 5 ? w[j] : tslot;
            if (id == -1 || (j < 5 && id < 0) || (id >= 0 && uint64_t(id) < textures)) continue;
            *why = "material " + std::to_string(k) + "'s " + (j < 5 ? slot[j] : "transmission map") + " names texture " +
                   std::to_string(id) + ", and the scene's pages hold " + std::to_string(textures);
            return false;
        }
        const char *what = nullptr;
        if (w[7] > 0) what = "a material function";
        if (!what && f[27] > 0.0f) what = "an alpha cutoff";
        if (what) {
            *why = "material " + std::to_string(k) + " needs " + what + ", which this core does not render";
            return false;
        }
        if (f[28] > 0.0f && !glass_table) {
            *why = "material " + std::to_string(k) + " is glass (transmission, layer.x) and the scene has no glass "
                   "energy table";
            return false;
        }
    }
    return true;
}

// ltree.rs:162-233 LightTree::check, as the kernel reads the tree (pathtrace.wgsl:547-618): from node 0 every
// interior node's children (its word and the next) are after it and inside the tree, every node is reached once and
// no deeper than 26 levels, every leaf's 1-4 entries lie among the CDF's entries; and every emitter triangle whose
// code has depth bits follows them through interior nodes to a leaf.
static bool light_tree_ok(const trt_scene_desc *d, uint64_t entries, std::string *why) {
    const uint64_t n = d->light_tree_vec4s;
    const auto *nd = static_cast<const uint32_t *>(d->light_tree);
    constexpr uint32_t kLeaf = 0x80000000u;
    std::vector<uint8_t> seen(n, 0);
    std::vector<std::pair<uint64_t, uint32_t>> st{{0, 0}};
    while (!st.empty()) {
        const auto [i, depth] = st.back();
        st.pop_back();
        if (seen[i]++) return *why = "light tree node " + std::to_string(i) + " is reached twice", false;
        if (depth > 26) return *why = "light tree deeper than 26 levels", false;
        const uint32_t w = nd[4 * i + 3];
        if (w & kLeaf) {
            if ((w & 0x1fffffffu) + ((w >> 29) & 3u) + 1u > entries)
                return *why = "a light tree leaf names entries past the emitter table", false;
        } else {
            if (w <= i || uint64_t(w) + 1 >= n)
                return *why = "light tree node " + std::to_string(i) + "'s children are not after it in the tree", false;
            st.push_back({w, depth + 1});
            st.push_back({uint64_t(w) + 1, depth + 1});
        }
    }
    for (uint64_t i = 0; i < n; i++)
        if (!seen[i]) return *why = "light tree node " + std::to_string(i) + " is not reached", false;
    const auto *et = static_cast<const uint32_t *>(d->emitter_tris);
    for (uint64_t k = 0; k < d->emitter_tri_count; k++) {
        const uint32_t code = et[16 * k + 11] & 0x7fffffffu;
        if ((code >> 26) == 0) continue;
        uint64_t i = 0;
        for (uint32_t level = 0; level + 1 < (code >> 26); level++) {
            const uint32_t w = nd[4 * i + 3];
            if (w & kLeaf) return *why = "an emitter's light-tree code goes past a leaf", false;
            i = uint64_t(w) + ((code >> level) & 1u);
        }
        if (!(nd[4 * i + 3] & kLeaf)) return *why = "an emitter's light-tree code does not end at a leaf", false;
    }
    return true;
}

// The shading tables' indices, checked once so the kernel can follow them: every unit's triangles' vertices and
// section rows, the rows' materials, the emitter entries' triangles and the emitters' triangles, instances, rows.
// texpack.rs pages as the kernel reads them: a power-of-two side of 16 or more (5 levels), at least one layer, every
// texture's rect on a layer the pages have and inside its page
// binary32 to binary16, rounded to nearest even (half::f16::from_f32)
static uint16_t half_of(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t sign = static_cast<uint16_t>((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u);
    if (ax >= 0x477ff000u) return sign | 0x7c00u;  // 65520 and above round to infinity
    if (ax < 0x38800000u) {                        // below 2^-14: a subnormal half, m x 2^-24
        float a;
        const uint32_t bits = ax;
        std::memcpy(&a, &bits, 4);
        return sign | static_cast<uint16_t>(std::nearbyint(a * 16777216.0f));
    }
    uint32_t h = ((ax >> 23) - 112u) << 10 | ((ax & 0x7fffffu) >> 13);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return static_cast<uint16_t>(sign | h);
}

static uint64_t page_layer_bytes(uint32_t side) {
    uint64_t n = 0;
    for (uint32_t l = 0; l < 5; l++) n += uint64_t(side >> l) * (side >> l) * 4;
    return n;
}
static bool pages_ok(const trt_scene_desc *d, std::string *why) {
    const uint32_t side = d->tex_side;
    if (!d->tex_pages || side < 16 || side > 16384 || (side & (side - 1)) || d->tex_layers == 0 ||
        d->tex_layers > 2048 || d->tex_info_vec4s < 2 || (d->tex_info_vec4s & 1) || d->tex_info_vec4s >= (1u << 30)) {
        *why = "texture pages need a power-of-two side from 16 to 16384, 1 to 2048 layers and 2 info rows a texture";
        return false;
    }
    const float *info = static_cast<const float *>(d->tex_info);
    for (uint64_t i = 0; i < d->tex_info_vec4s / 2; i++) {
        const float *a = info + 8 * i, *b = a + 4;
        bool ok = b[0] >= 0.0f && b[0] < float(d->tex_layers) && b[2] >= 0.0f && b[3] >= 0.0f;
        for (int j = 0; j < 8 && ok; j++) ok = std::isfinite(a[j]);
        if (!ok) {
            *why = "texture " + std::to_string(i) + "'s info names no layer of the pages, or is not finite";
            return false;
        }
    }
    return true;
}

static bool check_shading(const trt_scene_desc *d, std::string *why) {
    constexpr uint64_t kGuideVec4s = 65537 / 4 + 1;
    if (!d->vertices || !d->indices || !d->rows || !d->materials || !d->objects || d->object_count == 0) {
        *why = "shading tables without vertices, indices, rows, materials or objects";
        return false;
    }
    if (d->index_count != d->leaf_count) {
        *why = "the index buffer does not have a triangle for every leaf";
        return false;
    }
    if (d->light_count && !d->lights) {
        *why = "lights without a buffer";
        return false;
    }
    // lib.rs:367-380 light_order: the directional lights first, then point (1), spot (2) and rect (3) lights
    for (uint64_t k = 0, local = 0; k < d->light_count; k++) {
        const float type = static_cast<const float *>(d->lights)[20 * k + 3];
        if (!(type == 0.0f || type == 1.0f || type == 2.0f || type == 3.0f) || (type == 0.0f && local)) {
            *why = "light " + std::to_string(k) + ": a type other than 0-3, or a directional light after a local one";
            return false;
        }
        local = local || type != 0.0f;
    }
    if (d->emitter_vec4s && (d->emitter_vec4s < kGuideVec4s + 1 || !d->emitters)) {
        *why = "an emitter table shorter than its guide and header";
        return false;
    }
    const uint64_t textures = d->tex_info ? d->tex_info_vec4s / 2 : 0;
    if (!material_rendered(static_cast<const float *>(d->materials), d->material_count, d->glass_energy != nullptr,
                           textures, why))
        return false;
    if (d->tex_info && !pages_ok(d, why)) return false;
    if (d->vcolors && d->vcolor_words != 3 * d->index_count) {
        *why = "vcolors needs three words a triangle (vcolor_words = 3 x index_count)";
        return false;
    }
    if (d->env) {
        const uint64_t w = d->env_width, h = d->env_height;
        if (!d->env_cdf || w == 0 || h == 0 || w > 16384 || h > 16384 || d->env_cdf_count != h + w * h) {
            *why = "an environment map needs 1 to 16384 texels a side and env_cdf of env_height + env_width x "
                   "env_height floats";
            return false;
        }
        const float *cdf = static_cast<const float *>(d->env_cdf);
        for (uint64_t i = 0; i < d->env_cdf_count; i++)
            if (!std::isfinite(cdf[i])) {
                *why = "env_cdf holds a value that is not finite";
                return false;
            }
    }
    const auto *inst = static_cast<const trt::Instance *>(d->instances);
    const auto *leaves = static_cast<const trt::Leaf *>(d->leaves);
    const auto *nodes = static_cast<const trt::Node *>(d->nodes);
    // Every record the top tree names: its mesh tree's triangles.
    std::vector<uint32_t> st;
    if (d->unit_count) st.push_back(d->top_first);
    std::vector<uint32_t> recs;
    while (!st.empty()) {
        const uint32_t i = st.back();
        st.pop_back();
        const trt::Node &n = nodes[i];
        if (n.count > 0) {
            recs.push_back(n.left_or_first);
        } else {
            st.push_back(d->top_first + n.left_or_first);
            st.push_back(d->top_first + n.left_or_first + 1);
        }
    }
    auto tri_ok = [&](uint64_t ti, uint64_t base, uint64_t rows_first, std::string *w) {
        const uint32_t *ix = d->indices + 3 * ti;
        for (int k = 0; k < 3; k++)
            if (uint64_t(ix[k]) + base >= d->vertex_count) {
                *w = "a triangle's vertex is outside the vertex buffer";
                return false;
            }
        const uint64_t part = trt::to_bits(leaves[ti].r[0][3]);
        if (rows_first + part >= d->row_count) {
            *w = "a triangle's section has no row";
            return false;
        }
        if ((d->rows[4 * (rows_first + part)] & 0xffffffu) >= d->material_count) {
            *w = "a row names a material that is not there";
            return false;
        }
        return true;
    };
    for (uint32_t r : recs) {
        const trt::Instance &u = inst[r];
        const uint32_t root = trt::to_bits(u.n[0][3]);
        st.assign(1, root);
        while (!st.empty()) {
            const uint32_t i = st.back();
            st.pop_back();
            const trt::Node &n = nodes[i];
            if (n.count == 0) {
                st.push_back(root + n.left_or_first);
                st.push_back(root + n.left_or_first + 1);
                continue;
            }
            for (uint32_t k = 0; k < n.count; k++)
                if (!tri_ok(uint64_t(u.info[0]) + n.left_or_first + k, u.info[2], u.info[3], why)) return false;
        }
        if (u.info[1] >= d->object_count) {
            *why = "a record names an object that is not there";
            return false;
        }
    }
    const uint64_t entries = d->emitter_vec4s ? d->emitter_vec4s - kGuideVec4s - 1 : 0;
    if (d->light_tree_vec4s && (!d->light_tree || !light_tree_ok(d, entries, why))) {
        if (!d->light_tree) *why = "light tree nodes without a buffer";
        return false;
    }
    const auto *em = static_cast<const float *>(d->emitters);
    for (uint64_t k = 0; k < entries; k++)
        if (trt::to_bits(em[4 * k]) >= d->emitter_tri_count) {
            *why = "an emitter entry names a triangle that is not there";
            return false;
        }
    const auto *et = static_cast<const uint32_t *>(d->emitter_tris);
    for (uint64_t k = 0; k < d->emitter_tri_count; k++) {
        const uint32_t tri = et[16 * k + 3], in = et[16 * k + 7], part = et[16 * k + 12], code = et[16 * k + 11];
        const bool tree_code = ((code & 0x7fffffffu) >> 26) != 0 && d->light_tree_vec4s;
        if (tri >= d->leaf_count || in >= d->instance_count ||
            (!tree_code && (code & 0x7fffffffu) >= (entries ? entries : 1))) {
            *why = "an emitter triangle names a triangle, a record or an entry that is not there";
            return false;
        }
        const trt::Instance &u = inst[in];
        if (uint64_t(u.info[3]) + part >= d->row_count || u.info[1] >= d->object_count) {
            *why = "an emitter triangle's section has no row";
            return false;
        }
        const uint32_t *ix = d->indices + 3 * uint64_t(tri);
        for (int j = 0; j < 3; j++)
            if (uint64_t(ix[j]) + u.info[2] >= d->vertex_count) {
                *why = "an emitter triangle's vertex is outside the vertex buffer";
                return false;
            }
        if ((d->rows[4 * (uint64_t(u.info[3]) + part)] & 0xffffffu) >= d->material_count) {
            *why = "an emitter's row names a material that is not there";
            return false;
        }
    }
    // Rows that name emitters: each hit's entry (first + its index in the section) must be inside the list.
    for (uint32_t r : recs) {
        const trt::Instance &u = inst[r];
        const uint32_t root = trt::to_bits(u.n[0][3]);
        st.assign(1, root);
        while (!st.empty()) {
            const uint32_t i = st.back();
            st.pop_back();
            const trt::Node &n = nodes[i];
            if (n.count == 0) {
                st.push_back(root + n.left_or_first);
                st.push_back(root + n.left_or_first + 1);
                continue;
            }
            for (uint32_t k = 0; k < n.count; k++) {
                const uint64_t ti = uint64_t(u.info[0]) + n.left_or_first + k;
                const uint64_t part = trt::to_bits(leaves[ti].r[0][3]);
                const uint32_t first = d->rows[4 * (uint64_t(u.info[3]) + part) + 2];
                if (first != 0xffffffffu && uint64_t(first) + trt::to_bits(leaves[ti].r[1][3]) >= d->emitter_tri_count) {
                    *why = "a row's emitters run past the emitter list";
                    return false;
                }
            }
        }
    }
    return true;
}

// TRT_BUILD_AUTO builds on the device from this many mesh triangles: below it the device's per-level launches and
// waits (a few milliseconds) cost more than one host thread's build.
constexpr uint64_t kDeviceBuildMin = 65536;

static int upload(trt_device *d, void **dst, const void *src, uint64_t bytes, uint64_t *total) {
    TRT_HIP(hipMalloc(dst, bytes ? bytes : 64));
    if (bytes) TRT_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    *total += bytes ? bytes : 64;
    (void)d;
    return TRT_OK;
}

int trt_scene_upload(trt_device *d, const trt_scene_desc *desc, trt_scene **out) {
    if (!d || !desc || !out) return fail(TRT_E_INVALID, "trt_scene_upload: null argument");
    *out = nullptr;
    if ((desc->node_count && !desc->nodes) || (desc->leaf_count && !desc->leaves) ||
        (desc->instance_count && !desc->instances))
        return fail(TRT_E_INVALID, "trt_scene_upload: a buffer with a count has no pointer");
    if (desc->unit_count > 0 && (desc->node_count == 0 || desc->instance_count == 0))
        return fail(TRT_E_INVALID, "trt_scene_upload: units without nodes or records");
    if (desc->instance_count >= (uint64_t{1} << 30) || desc->leaf_count >= (uint64_t{1} << 32) ||
        desc->node_count >= (uint64_t{1} << 32))
        return fail(TRT_E_INVALID, "trt_scene_upload: too many records, leaves or nodes");
    if (desc->energy && (desc->vertex_count >= (uint64_t{1} << 32) || desc->index_count >= (uint64_t{1} << 32) ||
                         desc->row_count >= (uint64_t{1} << 32) || desc->material_count >= (uint64_t{1} << 32) ||
                         desc->light_count >= (uint64_t{1} << 32) || desc->emitter_vec4s >= (uint64_t{1} << 32) ||
                         desc->emitter_tri_count >= (uint64_t{1} << 32) || desc->object_count >= (uint64_t{1} << 32) ||
                         desc->light_tree_vec4s >= (uint64_t{1} << 29)))
        return fail(TRT_E_INVALID, "trt_scene_upload: a shading table of 2^32 records or more");
    const auto *nodes = static_cast<const trt::Node *>(desc->nodes);
    const auto *inst = static_cast<const trt::Instance *>(desc->instances);
    const auto *leaves = static_cast<const trt::Leaf *>(desc->leaves);
    std::string why;
    trt::WidePlan plan;
    if (!trt::wide_plan(nodes, desc->node_count, inst, desc->instance_count, desc->leaf_count, desc->top_first,
                        desc->unit_count, plan, &why))
        return fail(TRT_E_INVALID, "trt_scene_upload: " + why);
    if (desc->energy && !check_shading(desc, &why)) return fail(TRT_E_INVALID, "trt_scene_upload: " + why);
    TRT_HIP(hipSetDevice(d->ordinal));
    auto *s = new (std::nothrow) trt_scene();
    if (!s) return fail(TRT_E_OOM, "trt_scene_upload: out of host memory");
    s->dev = d;
    d->refs++;
    uint64_t total = 0;
    int rc = TRT_OK;
    {
        std::lock_guard<std::mutex> lock(d->mu);
        rc = upload(d, &s->nodes, desc->nodes, desc->node_count * sizeof(trt::Node), &total);
        if (!rc) rc = upload(d, &s->leaves, desc->leaves, desc->leaf_count * sizeof(trt::Leaf), &total);
        if (!rc) rc = upload(d, &s->instances, desc->instances, desc->instance_count * sizeof(trt::Instance), &total);
    }
    // The wide trees: the mesh trees built on the device (on a stream of the call's own, so renders queued by other
    // threads go on meanwhile) unless the device is set to build on the host; collapsed and quantised on the host.
    const auto t_build = std::chrono::steady_clock::now();
    trt::WideScene w;
    trt::gb::DeviceWide dw;
    trt::MeshWide mw;
    uint64_t mesh_tris = 0;
    const std::vector<trt::MeshRef> &meshes = plan.meshes;
    for (const trt::MeshRef &m : meshes) mesh_tris += m.n;
    const uint32_t where = d->build.load();
    const bool on_device = !meshes.empty() && (where == TRT_BUILD_DEVICE ||
                                               (where == TRT_BUILD_AUTO && mesh_tris >= kDeviceBuildMin));
    hipStream_t bs = nullptr;  // the device build's stream, kept until its wide nodes are copied
    if (!rc && on_device) {
        hipError_t e = hipStreamCreateWithFlags(&bs, hipStreamNonBlocking);
        if (e != hipSuccess) {
            bs = nullptr;
            rc = fail(TRT_E_DEVICE, std::string("trt_scene_upload: ") + hipGetErrorString(e));
        } else if (!trt::gb::device_wide(static_cast<const trt::Leaf *>(s->leaves), desc->leaf_count, meshes, bs, dw,
                                         &why)) {
            rc = fail(TRT_E_DEVICE, "trt_scene_upload: " + why);
        }
        if (dw.complete) mw = trt::MeshWide{dw.wnode_count, dw.mesh_ref, dw.mesh_bound};
    }
    auto end_build_stream = [&]() {
        if (!bs) return;
        if (dw.d_wnodes) hipFreeAsync(dw.d_wnodes, bs);
        dw.d_wnodes = nullptr;
        hipStreamSynchronize(bs);
        hipStreamDestroy(bs);
        bs = nullptr;
    };
    if (!rc && !trt::build_wide(nodes, desc->node_count, leaves, inst, desc->instance_count, desc->leaf_count,
                                desc->top_first, desc->unit_count, w, &why, &plan,
                                on_device && !dw.complete ? &dw.trees : nullptr, dw.complete ? &mw : nullptr))
        rc = fail(TRT_E_INVALID, "trt_scene_upload: " + why);
    const uint64_t wide_count = w.node_base + w.nodes.size();
    if (!rc && wide_count >= (uint64_t{1} << 30)) rc = fail(TRT_E_INVALID, "trt_scene_upload: tree too large");
    const double build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count();
    if (rc) {
        std::string keep = g_error;
        end_build_stream();
        hipFree(dw.d_wleaves);
        trt_scene_release(s);
        return fail(rc, keep);
    }
    std::lock_guard<std::mutex> lock(d->mu);
    if (dw.complete) {
        // the meshes' wide nodes and the leaves are on the device already: the top tree's nodes go after them
        const uint64_t mesh_bytes = dw.wnode_count * sizeof(trt::WideNode);
        const uint64_t bytes = wide_count * sizeof(trt::WideNode);
        hipError_t e = hipMalloc(&s->wnodes, bytes ? bytes : 64);
        if (e == hipSuccess && mesh_bytes)
            e = hipMemcpyAsync(s->wnodes, dw.d_wnodes, mesh_bytes, hipMemcpyDeviceToDevice, bs);
        if (e == hipSuccess && bytes > mesh_bytes)
            e = hipMemcpyAsync(static_cast<char *>(s->wnodes) + mesh_bytes, w.nodes.data(), bytes - mesh_bytes,
                               hipMemcpyHostToDevice, bs);
        end_build_stream();  // (waits for the copies; frees the build's copy of the nodes in stream order)
        s->wleaves = dw.d_wleaves;
        total += (bytes ? bytes : 64) + std::max<uint64_t>(desc->leaf_count * sizeof(trt::Leaf), 64);
        if (e != hipSuccess) rc = fail(TRT_E_DEVICE, std::string("trt_scene_upload: ") + hipGetErrorString(e));
        if (!rc) rc = upload(d, &s->wunits, w.units.data(), w.units.size() * sizeof(trt::WideUnit), &total);
    } else {
        end_build_stream();
        rc = upload(d, &s->wnodes, w.nodes.data(), w.nodes.size() * sizeof(trt::WideNode), &total);
        if (!rc) rc = upload(d, &s->wunits, w.units.data(), w.units.size() * sizeof(trt::WideUnit), &total);
        if (!rc) rc = upload(d, &s->wleaves, w.wleaves.data(), w.wleaves.size() * sizeof(trt::Leaf), &total);
    }
    if (!rc && desc->energy) {
        s->shading = true;
        for (uint64_t k = 0; k < desc->material_count; k++) {
            if (static_cast<const float *>(desc->materials)[72 * k + 30] > 0.0f) s->coat = true;
            if (static_cast<const float *>(desc->materials)[72 * k + 28] > 0.0f) s->glass = true;
            int32_t flag;  // tex1.z
            std::memcpy(&flag, static_cast<const float *>(desc->materials) + 72 * k + 18, 4);
            if (flag == 4) s->unlit = true;  // pathtrace.wgsl:54 MAT_UNLIT
        }
        s->light_count = desc->light_count;
        s->n_dir = 0;
        while (s->n_dir < desc->light_count && static_cast<const float *>(desc->lights)[20 * s->n_dir + 3] == 0.0f)
            s->n_dir++;
        s->emitter_tri_count = desc->emitter_tri_count;
        s->emitter_entries = desc->emitter_vec4s ? desc->emitter_vec4s - (65537 / 4 + 1) - 1 : 0;
        s->vertex_count = desc->vertex_count, s->tri_count = desc->index_count, s->row_count = desc->row_count;
        s->material_count = desc->material_count, s->object_count = desc->object_count;
        s->emitter_vec4s = desc->emitter_vec4s;
        rc = upload(d, &s->vertices, desc->vertices, desc->vertex_count * 48, &total);
        if (!rc) rc = upload(d, &s->indices, desc->indices, desc->index_count * 12, &total);
        if (!rc) rc = upload(d, &s->rows, desc->rows, desc->row_count * 16, &total);
        if (!rc) rc = upload(d, &s->materials, desc->materials, desc->material_count * 288, &total);
        if (!rc) rc = upload(d, &s->lights, desc->lights, desc->light_count * 80, &total);
        if (!rc) rc = upload(d, &s->emitters, desc->emitters, desc->emitter_vec4s * 16, &total);
        if (!rc) rc = upload(d, &s->emitter_tris, desc->emitter_tris, desc->emitter_tri_count * 64, &total);
        if (!rc) rc = upload(d, &s->objects, desc->objects, desc->object_count * 48, &total);
        if (!rc) rc = upload(d, &s->energy, desc->energy, 512 * 16, &total);
        if (!rc) rc = upload(d, &s->light_tree, desc->light_tree, desc->light_tree_vec4s * 16, &total);
        s->light_tree_count = desc->light_tree_vec4s;
        if (!rc && s->glass) rc = upload(d, &s->glass_energy, desc->glass_energy, 1024 * 16, &total);
        if (!rc && desc->tex_info) {
            s->tex = true;
            s->tex_side = desc->tex_side, s->tex_layers = desc->tex_layers, s->tex_info_vec4s = desc->tex_info_vec4s;
            rc = upload(d, &s->tex_pages, desc->tex_pages, page_layer_bytes(desc->tex_side) * desc->tex_layers, &total);
            if (!rc) rc = upload(d, &s->tex_info, desc->tex_info, desc->tex_info_vec4s * 16, &total);
        }
        if (!rc && desc->vcolors) {
            s->vcolor_words = desc->vcolor_words;
            rc = upload(d, &s->vcolors, desc->vcolors, desc->vcolor_words * 4, &total);
        }
        if (!rc && desc->env) {
            // gpu.rs:3552-3561: the texels as half floats of min(v, 65000), rounded to nearest even
            const uint64_t n = uint64_t(desc->env_width) * desc->env_height;
            std::vector<uint16_t> half(4 * n);
            const float *src = static_cast<const float *>(desc->env);
            for (uint64_t i = 0; i < 4 * n; i++) half[i] = half_of(std::isnan(src[i]) ? 65000.0f : std::min(src[i], 65000.0f));
            s->env_w = desc->env_width, s->env_h = desc->env_height, s->env_cdf_count = desc->env_cdf_count;
            rc = upload(d, &s->env, half.data(), 8 * n, &total);
            if (!rc) rc = upload(d, &s->env_cdf, desc->env_cdf, desc->env_cdf_count * 4, &total);
        }
    }
    if (rc) {
        std::string keep = g_error;
        trt_scene_release(s);
        return fail(rc, keep);
    }
    s->node_count = desc->node_count;
    s->leaf_count = desc->leaf_count;
    s->instance_count = desc->instance_count;
    s->wnode_count = wide_count;
    s->wunit_count = w.units.size();
    s->top_first = desc->top_first;
    s->units = desc->unit_count;
    s->top_ref = w.top_ref;
    s->max_stack = w.max_stack;
    std::memcpy(s->top_box, w.top_box, sizeof(s->top_box));
    s->stats.wide_nodes = wide_count;
    s->stats.wide_bytes = wide_count * sizeof(trt::WideNode);
    s->stats.binary_bytes = desc->node_count * sizeof(trt::Node);
    s->stats.leaf_bytes = desc->leaf_count * sizeof(trt::Leaf);
    s->stats.unit_bytes = w.units.size() * sizeof(trt::WideUnit);
    s->stats.max_stack = w.max_stack;
    s->stats.units = desc->unit_count;
    s->stats.device_bytes = total;
    s->stats.build_ms = float(build_ms);
    s->stats.device_build_ms = dw.times.device_ms;
    if (dw.complete) s->stats.device_meshes = uint32_t(meshes.size());
    else
        for (const trt::OwnBuilt &o : dw.trees) s->stats.device_meshes += o.nodes.empty() ? 0 : 1;
    s->stats.host_meshes = uint32_t(meshes.size()) - s->stats.device_meshes;
    *out = s;
    return ok();
}

int trt_device_set_render(trt_device *d, uint32_t how) {
    if (!d) return fail(TRT_E_INVALID, "trt_device_set_render: null device");
    if (how != TRT_RENDER_AUTO && how != TRT_RENDER_MEGAKERNEL && how != TRT_RENDER_WAVEFRONT)
        return fail(TRT_E_INVALID, "trt_device_set_render: how must be TRT_RENDER_AUTO, _MEGAKERNEL or _WAVEFRONT");
    d->render = how;
    return ok();
}

int trt_device_set_build(trt_device *d, uint32_t where) {
    if (!d) return fail(TRT_E_INVALID, "trt_device_set_build: null device");
    if (where != TRT_BUILD_DEVICE && where != TRT_BUILD_HOST && where != TRT_BUILD_AUTO)
        return fail(TRT_E_INVALID, "trt_device_set_build: where must be TRT_BUILD_DEVICE, _HOST or _AUTO");
    d->build = where;
    return ok();
}

int trt_scene_wide_read(trt_scene *s, uint32_t which, void *dst, uint64_t bytes) {
    if (!s || !dst) return fail(TRT_E_INVALID, "trt_scene_wide_read: null argument");
    const void *src = which == 0 ? s->wnodes : which == 1 ? s->wleaves : which == 2 ? s->wunits : nullptr;
    const uint64_t size = which == 0 ? s->wnode_count * sizeof(trt::WideNode)
                          : which == 1 ? s->leaf_count * sizeof(trt::Leaf)
                          : which == 2 ? s->wunit_count * sizeof(trt::WideUnit) : 0;
    if (which > 2) return fail(TRT_E_INVALID, "trt_scene_wide_read: which must be 0, 1 or 2");
    if (bytes != size) return fail(TRT_E_INVALID, "trt_scene_wide_read: bytes is not the buffer's size");
    if (!bytes) return ok();
    std::lock_guard<std::mutex> lock(s->dev->mu);
    TRT_HIP(hipSetDevice(s->dev->ordinal));
    TRT_HIP(hipStreamSynchronize(s->dev->stream));
    TRT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ok();
}

int trt_scene_stats_get(trt_scene *s, trt_scene_stats *out) {
    if (!s || !out) return fail(TRT_E_INVALID, "trt_scene_stats_get: null argument");
    *out = s->stats;
    return ok();
}

// ---- ray batches ---------------------------------------------------------------------------------------------------
static int ensure_spill(trt_device *d, uint64_t bytes) {
    if (bytes <= d->spill_bytes) return TRT_OK;
    TRT_HIP(hipStreamSynchronize(d->stream));
    TRT_HIP(hipFree(d->spill));
    d->spill = nullptr;
    d->spill_bytes = 0;
    TRT_HIP(hipMalloc(reinterpret_cast<void **>(&d->spill), bytes));
    d->spill_bytes = bytes;
    return TRT_OK;
}

// noalloc: the counter and the spilled stack entries in the GPU's fixed scratch (d->scratch), no allocation; a grid
// smaller than the full one when a deep tree's entries would not fit it at full occupancy.
static int launch(trt_scene *s, const void *d_rays, uint64_t n, void *d_hits, uint32_t flags, bool count = false,
                  bool noalloc = false) {
    trt_device *d = s->dev;
    const bool wide = flags != TRT_WALK_BINARY;
    const bool watertight = flags == TRT_WALK_WATERTIGHT;
    if (flags != TRT_WALK_WIDE && flags != TRT_WALK_BINARY && flags != TRT_WALK_WATERTIGHT)
        return fail(TRT_E_INVALID, "unknown trace flags (TRT_WALK_WIDE, TRT_WALK_BINARY or TRT_WALK_WATERTIGHT)");
    if (n >= (uint64_t{1} << 31)) return fail(TRT_E_INVALID, "a batch of 2^31 rays or more");
    if (n == 0) return ok();
    int grid = d->cus * (wide ? d->blocks_wide : d->blocks_binary);
    const uint32_t deepest = wide ? s->max_stack : trt::kRefStack;
    const uint64_t entries = deepest > trt::kLdsStack ? deepest - trt::kLdsStack : 1;
    uint32_t *counter = d->counter, *spill = nullptr;
    if (noalloc) {
        const uint64_t room = (trt::kScratchBytes - trt::kScratchCounterBytes) / sizeof(uint32_t);
        const uint64_t blocks = room / (entries * trt::kTraceBlock);
        if (blocks == 0)
            return fail(TRT_E_INVALID, "the scene's walks need " + std::to_string(entries * trt::kTraceBlock * 4) +
                                           " bytes of stack a block past LDS; the fixed scratch holds " +
                                           std::to_string(room * 4));
        grid = static_cast<int>(std::min<uint64_t>(static_cast<uint64_t>(grid), blocks));
        counter = d->scratch;
        spill = d->scratch + trt::kScratchCounterBytes / sizeof(uint32_t);
    }
    const uint32_t threads = static_cast<uint32_t>(grid) * trt::kTraceBlock;
    if (!noalloc) {
        int rc = ensure_spill(d, entries * threads * sizeof(uint32_t));
        if (rc) return rc;
        spill = d->spill;
    }
    trt::TraceArgs a{};
    a.wnodes = static_cast<const uint4 *>(s->wnodes);
    a.wunits = static_cast<const float4 *>(s->wunits);
    a.leaves = static_cast<const float4 *>(s->leaves);
    a.wleaves = static_cast<const float4 *>(s->wleaves);
    a.bnodes = static_cast<const float4 *>(s->nodes);
    a.instances = static_cast<const float4 *>(s->instances);
    a.rays = static_cast<const float4 *>(d_rays);
    a.hits = static_cast<uint32_t *>(d_hits);
    a.counter = counter;
    a.guard_flag = d->guard;
    a.spill = spill;
    a.reads = d->reads;
    a.n = n;
    a.spill_stride = threads;
    a.spill_entries = static_cast<uint32_t>(entries);
    a.wnode_count = static_cast<uint32_t>(s->wnode_count);
    a.wunit_count = static_cast<uint32_t>(s->wunit_count);
    a.bnode_count = static_cast<uint32_t>(s->node_count);
    a.instance_count = static_cast<uint32_t>(s->instance_count);
    a.leaf_count = static_cast<uint32_t>(s->leaf_count);
    a.top_ref = s->top_ref;
    a.top_first = s->top_first;
    a.units = s->units;
    std::memcpy(a.top_box, s->top_box, sizeof(a.top_box));
    if (trt::launch_trace(a, wide, count, grid, d->stream, watertight))
        return fail(TRT_E_DEVICE, "trace kernel launch failed");
    return ok();
}

int trt_trace_reads(trt_scene *s, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags,
                    uint64_t *reads) {
    if (!s || !reads || (n && (!d_rays || !d_hits))) return fail(TRT_E_INVALID, "trt_trace_reads: null argument");
    trt_device *d = s->dev;
    {
        std::lock_guard<std::mutex> lock(d->mu);
        TRT_HIP(hipSetDevice(d->ordinal));
        TRT_HIP(hipMemsetAsync(d->reads, 0, 4 * sizeof(unsigned long long), d->stream));
        int rc = launch(s, d_rays, n, d_hits, flags, true);
        if (rc) return rc;
        unsigned long long r[4] = {0, 0, 0, 0};
        TRT_HIP(hipMemcpyAsync(r, d->reads, sizeof(r), hipMemcpyDeviceToHost, d->stream));
        TRT_HIP(hipStreamSynchronize(d->stream));
        reads[0] = r[0], reads[1] = r[1], reads[2] = r[2];
    }
    return trt_device_sync(d);
}

int trt_trace_check_fp64(trt_scene *s, const trt_ray *d_rays, const trt_hit *d_hits, uint64_t n, double *d_t64) {
    if (!s || (n && (!d_rays || !d_hits || !d_t64))) return fail(TRT_E_INVALID, "trt_trace_check_fp64: null argument");
    if (n == 0) return ok();
    trt_device *d = s->dev;
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    const uint64_t units = s->wunits ? s->wunit_count : 0;
    if (trt::launch_check64(reinterpret_cast<const float4 *>(d_rays), reinterpret_cast<const uint32_t *>(d_hits),
                            static_cast<const float4 *>(s->leaves), static_cast<const float4 *>(s->wunits), n,
                            s->leaf_count, units, d_t64, d->cus * 8, d->stream))
        return fail(TRT_E_DEVICE, "FP64 check kernel launch failed");
    return ok();
}

// ---- path tracing --------------------------------------------------------------------------------------------------
struct trt_accum {
    trt_device *dev = nullptr;
    uint32_t w = 0, h = 0;
    void *buf = nullptr;
};

int trt_accum_create(trt_device *d, uint32_t width, uint32_t height, trt_accum **out) {
    if (!d || !out || width == 0 || height == 0 || uint64_t(width) * height >= (uint64_t{1} << 31))
        return fail(TRT_E_INVALID, "trt_accum_create: bad argument");
    *out = nullptr;
    TRT_HIP(hipSetDevice(d->ordinal));
    auto *a = new (std::nothrow) trt_accum();
    if (!a) return fail(TRT_E_OOM, "trt_accum_create: out of host memory");
    a->dev = d, a->w = width, a->h = height;
    const uint64_t bytes = uint64_t(width) * height * 80;
    hipError_t e = hipMalloc(&a->buf, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a->buf, 0, bytes, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) {
        hipFree(a->buf);
        delete a;
        return fail(e == hipErrorOutOfMemory ? TRT_E_OOM : TRT_E_DEVICE, hipGetErrorString(e));
    }
    d->refs++;
    *out = a;
    return ok();
}
int trt_accum_release(trt_accum *a) {
    if (!a) return ok();
    hipSetDevice(a->dev->ordinal);
    hipStreamSynchronize(a->dev->stream);
    hipFree(a->buf);
    trt_device *d = a->dev;
    delete a;
    device_unref(d);
    return ok();
}
int trt_accum_clear(trt_accum *a) {
    if (!a) return fail(TRT_E_INVALID, "trt_accum_clear: null accumulator");
    TRT_HIP(hipSetDevice(a->dev->ordinal));
    TRT_HIP(hipMemsetAsync(a->buf, 0, uint64_t(a->w) * a->h * 80, a->dev->stream));
    return ok();
}
int trt_accum_read(trt_accum *a, float *pixels) {
    if (!a || !pixels) return fail(TRT_E_INVALID, "trt_accum_read: null argument");
    int rc = trt_device_sync(a->dev);
    if (rc) return rc;
    TRT_HIP(hipMemcpy(pixels, a->buf, uint64_t(a->w) * a->h * 80, hipMemcpyDeviceToHost));
    return ok();
}
int trt_accum_resolve(trt_accum *a, float *color, float *albedo, float *normal, float *variance) {
    if (!a) return fail(TRT_E_INVALID, "trt_accum_resolve: null accumulator");
    const uint64_t n = uint64_t(a->w) * a->h;
    std::vector<float> px;
    try {
        px.resize(n * 20);
    } catch (...) {
        return fail(TRT_E_OOM, "trt_accum_resolve: out of host memory");
    }
    int rc = trt_accum_read(a, px.data());
    if (rc) return rc;
    for (uint64_t i = 0; i < n; i++) {
        float p[20];
        for (int k = 0; k < 20; k++) p[k] = std::isfinite(px[20 * i + k]) ? px[20 * i + k] : 0.0f;
        const float c = std::fmax(px[20 * i + 3], 1.0f);  // f32::max: a NaN count gives 1
        for (int k = 0; k < 3; k++) {
            if (color) color[3 * i + k] = p[k] / c;
            if (albedo) albedo[3 * i + k] = p[4 + k] / c;
            if (normal) normal[3 * i + k] = p[8 + k] / c;
        }
        const float mean = p[7] / c;
        if (variance) variance[i] = std::fmax(p[11] / c - mean * mean, 0.0f) / c;
    }
    return ok();
}

int trt_accum_add_fp64(trt_accum *a, double *d_sum) {
    if (!a || !d_sum) return fail(TRT_E_INVALID, "trt_accum_add_fp64: null argument");
    trt_device *d = a->dev;
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    if (trt::launch_add64(static_cast<const float4 *>(a->buf), uint64_t(a->w) * a->h, d_sum, d->stream))
        return fail(TRT_E_DEVICE, "FP64 accumulation kernel launch failed");
    return ok();
}

int trt_accum_resolve_device(trt_accum *a, uint32_t pw, uint32_t ph, float *d_color, float *d_albedo,
                             float *d_normal) {
    if (!a || !d_color || !d_albedo || !d_normal) return fail(TRT_E_INVALID, "trt_accum_resolve_device: null argument");
    if (pw < a->w || ph < a->h) return fail(TRT_E_INVALID, "trt_accum_resolve_device: images smaller than the accumulator");
    trt_device *d = a->dev;
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    if (trt::launch_resolve(static_cast<const float4 *>(a->buf), a->w, a->h, pw, ph, d_color, d_albedo, d_normal,
                            d->stream))
        return fail(TRT_E_DEVICE, "resolve kernel launch failed");
    return ok();
}

// One job's arguments, checked against its scene and accumulator.
static int job_args(const trt_render_job &j, trt::RenderArgs &a, std::string *why) {
    trt_scene *s = j.scene;
    trt_accum *acc = j.acc;
    if (!s || !j.params || !acc) return *why = "null argument", TRT_E_INVALID;
    if (!s->shading) return *why = "the scene was uploaded without its shading tables", TRT_E_INVALID;
    if (acc->dev != s->dev) return *why = "the accumulator is another device's", TRT_E_INVALID;
    a = trt::RenderArgs{};
    std::memcpy(a.params, j.params, sizeof(a.params));
    const uint32_t *P = a.params;
    if (P[16] != acc->w || P[17] != acc->h) return *why = "P.size is not the accumulator's", TRT_E_INVALID;
    // P.counts2.x the directional lights (the scene's leading ones); past them, local lights with P.counts2.w the
    // first shadowed one (lib.rs:700)
    if (P[36] > s->n_dir || (P[20] > P[36] && P[36] != s->n_dir))
        return *why = "P.counts2.x is not the scene's directional light count", TRT_E_INVALID;
    if (P[20] > P[36] && (P[39] < P[36] || P[39] > P[20]))
        return *why = "P.counts2.w (the first shadowed local light) is not among the local lights", TRT_E_INVALID;
    if (P[20] > s->light_count || P[36] > P[20] || P[21] > s->emitter_entries)
        return *why = "P.counts names more lights or emitters than the scene has", TRT_E_INVALID;
    if (P[21] > 0 && P[21] != s->emitter_entries)
        return *why = "P.counts.y is not the emitter table's entry count", TRT_E_INVALID;
    if (P[22] > 64) return *why = "more than 64 bounces", TRT_E_INVALID;
    // What the offline `main` would read and this core does not render (Params, pathtrace.wgsl:96-140), refused by
    // name rather than rendered without it.
    // pathtrace.wgsl:1895-1941: P.env.x an environment map of P.env.yz texels, the scene's
    if (trt::from_bits(P[24]) >= 0.5f &&
        (!s->env || trt::from_bits(P[25]) != float(s->env_w) || trt::from_bits(P[26]) != float(s->env_h)))
        return *why = s->env ? "P.env.yz is not the size of the scene's environment map"
                             : "P.env.x asks for an environment map and the scene has none",
               TRT_E_INVALID;
    if (trt::from_bits(P[44]) > 0.0f) return *why = "fog (P.fog.x) is not supported", TRT_E_INVALID;
    if (trt::from_bits(P[52]) > 0.0f) return *why = "volumetric scattering (P.vol.x) is not supported", TRT_E_INVALID;
    if (trt::from_bits(P[75]) > 0.0f) return *why = "the live light LOD (P.light_lod.w) is not supported", TRT_E_INVALID;
    if (P[89] > 0) return *why = "procedural geometry (P.procs.y) is not supported", TRT_E_INVALID;
    if (trt::from_bits(P[92]) > 0.0f)
        return *why = "aerial perspective (P.aerial.x) is not supported", TRT_E_INVALID;
    if (trt::from_bits(P[87]) > 0.5f && !s->light_tree_count)
        return *why = "light-tree candidates (P.ltree.w) for a scene uploaded without a light tree", TRT_E_INVALID;
    if (trt::from_bits(P[87]) > 0.5f && (P[85] > s->emitter_entries || trt::from_bits(P[87]) > 64.0f))
        return *why = "P.ltree names more tree entries than the emitter table has, or more than 64 candidates",
               TRT_E_INVALID;
    a.geo.wnodes = static_cast<const uint4 *>(s->wnodes);
    a.geo.wunits = static_cast<const float4 *>(s->wunits);
    a.geo.leaves = static_cast<const float4 *>(s->leaves);
    a.geo.wleaves = static_cast<const float4 *>(s->wleaves);
    a.geo.top_ref = s->top_ref;
    a.geo.top_first = s->top_first;
    a.geo.units = s->units;
    a.geo.wnode_count = static_cast<uint32_t>(s->wnode_count);
    a.geo.wunit_count = static_cast<uint32_t>(s->wunit_count);
    a.geo.bnode_count = static_cast<uint32_t>(s->node_count);
    a.geo.instance_count = static_cast<uint32_t>(s->instance_count);
    a.geo.leaf_count = static_cast<uint32_t>(s->leaf_count);
    std::memcpy(a.geo.top_box, s->top_box, sizeof(a.geo.top_box));
    a.instances = static_cast<const float4 *>(s->instances);
    a.vertices = static_cast<const float4 *>(s->vertices);
    a.indices = static_cast<const uint32_t *>(s->indices);
    a.rows = static_cast<const uint32_t *>(s->rows);
    a.materials = static_cast<const float4 *>(s->materials);
    a.lights = static_cast<const float4 *>(s->lights);
    a.emitters = static_cast<const float4 *>(s->emitters);
    a.emitter_tris = static_cast<const float4 *>(s->emitter_tris);
    a.objects = static_cast<const float4 *>(s->objects);
    a.energy = static_cast<const float4 *>(s->energy);
    a.light_tree = static_cast<const float4 *>(s->light_tree);
    a.light_tree_count = static_cast<uint32_t>(s->light_tree_count);
    a.glass_energy = static_cast<const float4 *>(s->glass_energy);
    a.tex_pages = static_cast<const uint32_t *>(s->tex_pages);
    a.tex_info = static_cast<const float4 *>(s->tex_info);
    a.tex_info_count = s->tex_info_vec4s;
    a.tex_side = s->tex_side, a.tex_layers = s->tex_layers;
    a.env = static_cast<const uint2 *>(s->env);
    a.env_cdf = static_cast<const float *>(s->env_cdf);
    a.env_cdf_count = s->env_cdf_count;
    a.env_w = s->env_w, a.env_h = s->env_h;
    a.vcolors = static_cast<const uint32_t *>(s->vcolors);
    a.vcolor_words = s->vcolor_words;
    a.accum = static_cast<float4 *>(acc->buf);
    a.vertex_count = static_cast<uint32_t>(s->vertex_count);
    a.tri_count = static_cast<uint32_t>(s->tri_count);
    a.row_count = static_cast<uint32_t>(s->row_count);
    a.material_count = static_cast<uint32_t>(s->material_count);
    a.light_count = static_cast<uint32_t>(s->light_count);
    a.emitter_vec4s = static_cast<uint32_t>(s->emitter_vec4s);
    a.emitter_tri_count = static_cast<uint32_t>(s->emitter_tri_count);
    a.object_count = static_cast<uint32_t>(s->object_count);
    a.first_sample = j.first_sample;
    a.spp = j.spp;
    return TRT_OK;
}

// ---- the wavefront path tracer's plan ------------------------------------------------------------------------------
// A batch's paths (pixel x sample) go through the wavefront in chunks of at most kWfMaxPaths, the jobs of one scene
// after another (a chunk holds one scene's jobs: its walk launches take that scene's trees). A job whose image fits
// a chunk goes in whole rounds (every pixel, several samples); a larger one in pixel ranges of one sample. A chunk
// never holds one pixel twice, and a pixel's samples come in sample order, so that the accumulation adds them as the
// megakernel does.
constexpr uint64_t kWfMaxPaths = uint64_t{1} << 22;
// TRT_RENDER_AUTO's choice: on MI355X the wavefront renders scenes of 3,000 to 3M triangles 1.4-4.4x as fast as the
// megakernel (clustered soups at 320x180 to 1280x720, 1-4 spp; the 66M-triangle instanced city 1.4-2.0x); the
// 36-triangle Cornell box, whose bounces cost less than the path state the wavefront moves, 0.7-0.8x as fast at
// 1280x720 and 1-4 spp (up to 3.7M paths) but 1.15x at 512x512 x 16 spp (a full chunk of 4.2M paths).
constexpr uint64_t kWfAutoLeaves = 1024;
constexpr uint64_t kWfBytesPerPath = 24 * 16 + (6 * 6 + 2 * 6 + 2) * 4;  // 24 float4 arrays, hits, lists, actives
constexpr uint64_t kWfCountBytes = 3 * 2 * trt::kShardBytes;               // the queues' segment counts
constexpr uint64_t kWfStrideStep = uint64_t{trt::kShards} * 64;          // a segment holds whole waves of paths
struct WfChunk {
    uint32_t sub_first, nsubs, npaths, pixels;
    uint32_t job;   // a job of the chunk (table index): the scene whose trees its walks take
    uint32_t last;  // the last bounce of its jobs
};

// scene_of[j]: table entry j's scene
static void wf_plan(const std::vector<trt::RenderJob> &table, const std::vector<const trt_scene *> &scene_of,
                    uint64_t stride, std::vector<trt::WfSub> &subs, std::vector<WfChunk> &chunks) {
    WfChunk cur{0, 0, 0, 0, 0, 0};
    auto close = [&]() {
        if (cur.npaths) chunks.push_back(cur);
        cur = WfChunk{static_cast<uint32_t>(subs.size()), 0, 0, 0, 0, 0};
    };
    auto add = [&](uint32_t job, uint32_t pixel_first, uint32_t count, uint32_t sample_first, uint32_t rounds,
                   uint32_t width) {
        subs.push_back(trt::WfSub{job, pixel_first, count, sample_first, rounds, cur.npaths, cur.pixels, width});
        if (cur.nsubs == 0) cur.job = job;
        cur.last = std::max(cur.last, table[job].args.params[22]);  // P.counts.z
        cur.nsubs++;
        cur.npaths += count * rounds;
        cur.pixels += count;
    };
    // the jobs grouped by scene, in the order the scenes first appear
    std::vector<uint32_t> order(table.size());
    for (uint32_t j = 0; j < table.size(); j++) order[j] = j;
    std::vector<const trt_scene *> firsts;
    for (const trt_scene *sc : scene_of)
        if (std::find(firsts.begin(), firsts.end(), sc) == firsts.end()) firsts.push_back(sc);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return std::find(firsts.begin(), firsts.end(), scene_of[x]) < std::find(firsts.begin(), firsts.end(), scene_of[y]);
    });
    for (uint32_t k = 0; k < order.size(); k++) {
        const uint32_t j = order[k];
        if (k > 0 && scene_of[order[k - 1]] != scene_of[j]) close();  // a new scene: a new chunk
        const trt::RenderArgs &a = table[j].args;
        const uint32_t w = a.params[16], spp = a.spp;
        const uint64_t npix = uint64_t(a.params[16]) * a.params[17];
        uint32_t r = 0;
        uint64_t pix = 0;
        while (r < spp) {
            const uint64_t room = stride - cur.npaths;
            if (npix <= stride) {
                if (room < npix) {
                    close();
                    continue;
                }
                const uint32_t rounds = static_cast<uint32_t>(std::min<uint64_t>(spp - r, room / npix));
                add(j, 0, static_cast<uint32_t>(npix), a.first_sample + r, rounds, w);
                r += rounds;
            } else {
                if (room == 0) {
                    close();
                    continue;
                }
                const uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(npix - pix, room));
                add(j, static_cast<uint32_t>(pix), cnt, a.first_sample + r, 1, w);
                pix += cnt;
                if (pix == npix) pix = 0, r++;
            }
        }
    }
    close();
}

static trt::WfArgs wf_layout(void *base, uint64_t stride) {
    trt::WfArgs w{};
    auto *f = static_cast<float4 *>(base);
    w.stride = static_cast<uint32_t>(stride);
    w.cap = static_cast<uint32_t>(stride / trt::kShards);
    w.thr = f;
    w.rad = f + stride;
    w.alb = f + 2 * stride;
    w.nrm = f + 3 * stride;
    w.misx = f + 4 * stride;
    w.misn = f + 5 * stride;
    w.thr_dl = f + 6 * stride;
    w.contrib = f + 7 * stride;                // kWfShadow slots
    w.rays = f + (7 + trt::kWfShadow) * stride;  // kWfSlots slots of 2 float4
    auto *u = reinterpret_cast<uint32_t *>(f + 24 * stride);
    w.hits = u;
    u += 6 * trt::kWfSlots * stride;
    w.list[0] = u;
    u += trt::kWfSlots * stride;
    w.list[1] = u;
    u += trt::kWfSlots * stride;
    w.active[0] = u;
    u += stride;
    w.active[1] = u;
    u += stride;
    w.counts = u;
    return w;
}
static_assert(7 + trt::kWfShadow + 2 * trt::kWfSlots == 24, "the float4 arrays of wf_layout");

static int ensure_wf(trt_device *d, uint64_t stride) {
    if (stride <= d->wf_stride) return TRT_OK;
    TRT_HIP(hipStreamSynchronize(d->stream));
    TRT_HIP(hipFree(d->wf));
    d->wf = nullptr;
    d->wf_stride = 0;
    TRT_HIP(hipMalloc(&d->wf, stride * kWfBytesPerPath + kWfCountBytes));
    d->wf_stride = stride;
    return TRT_OK;
}

// The wavefront's launches for a batch whose jobs share one scene: per chunk, the camera rays, then per bounce the
// walk of the queued rays and the shade kernel, then the accumulation. `last`: the jobs' largest last bounce.
static int wf_launch(trt_device *d, const trt::RenderJob *d_jobs, const trt::WfSub *d_subs,
                     const std::vector<WfChunk> &chunks, uint64_t stride, const std::vector<trt::RenderJob> &table,
                     bool coat, bool tree) {
    trt::WfArgs w = wf_layout(d->wf, stride);
    w.jobs = d_jobs;
    w.guard = d->guard;
    w.rays_traced = d->ray_counts;
    const int trace_grid = d->cus * d->blocks_wide;
    for (const WfChunk &c : chunks) {
        trt::TraceArgs t = table[c.job].args.geo;  // the chunk's scene
        t.rays = w.rays;
        t.hits = w.hits;
        t.n = trt::kWfSlots * stride;
        t.seg_cap = trt::kWfSlots * w.cap;
        const uint32_t last = c.last;
        w.subs = d_subs + c.sub_first;
        w.nsubs = c.nsubs;
        w.npaths = c.npaths;
        const int blocks = static_cast<int>((c.npaths + trt::kRenderBlock - 1) / trt::kRenderBlock);
        const int small = std::min(blocks, d->cus * 8);
        // every queue segment needs a block: at least kShards
        const int shade_grid = std::max(static_cast<int>(trt::kShards), std::min(blocks, d->cus * d->blocks_shade));
        if (trt::launch_wf_begin(w, small, d->stream)) return fail(TRT_E_DEVICE, "wavefront launch failed");
        for (uint32_t k = 0; k <= last; k++) {
            t.ray_index = w.list[k & 1u];
            t.seg_counts = w.counts + (k % 3u) * 2 * trt::kShards * trt::kShardGap;  // ring entry k % 3's ray list
            if (trt::launch_trace(t, true, false, trace_grid, d->stream))
                return fail(TRT_E_DEVICE, "wavefront trace launch failed");
            if (trt::launch_wf_shade(w, k, coat, tree, shade_grid, d->stream))
                return fail(TRT_E_DEVICE, "wavefront shade launch failed");
        }
        const int acc_grid = std::min(static_cast<int>((c.pixels + trt::kRenderBlock - 1) / trt::kRenderBlock),
                                      d->cus * 8);
        if (trt::launch_wf_accum(w, c.pixels, acc_grid, d->stream))
            return fail(TRT_E_DEVICE, "wavefront accumulation launch failed");
    }
    return ok();
}

// reads: NULL, or the counters a counting launch adds to (the kernel with counted walks).
static int render_batch(trt_device *d, const trt_render_job *jobs, uint32_t n, unsigned long long *reads) {
    if (!d || (n && !jobs)) return fail(TRT_E_INVALID, "trt_render_batch: null argument");
    if (n > 4096) return fail(TRT_E_INVALID, "trt_render_batch: more than 4096 jobs in one launch");
    std::vector<trt::RenderJob> table;
    table.reserve(n);
    uint64_t tiles = 0;
    uint32_t deepest = 0;
    bool coat = false, tree = false, glass = false, tex = false, local = false, env = false, other = false;
    for (uint32_t k = 0; k < n; k++) {
        if (jobs[k].spp == 0) continue;
        if (!jobs[k].scene || jobs[k].scene->dev != d)
            return fail(TRT_E_INVALID, "trt_render_batch: job " + std::to_string(k) + "'s scene is not this device's");
        trt::RenderJob j{};
        std::string why;
        int rc = job_args(jobs[k], j.args, &why);
        if (rc) return fail(rc, "trt_render_batch: job " + std::to_string(k) + ": " + why);
        for (uint32_t q = 0; q < k; q++)
            if (jobs[q].spp && jobs[q].acc == jobs[k].acc)
                return fail(TRT_E_INVALID, "trt_render_batch: an accumulator appears twice in one batch");
        j.tile_first = static_cast<uint32_t>(tiles);
        tiles += uint64_t((jobs[k].acc->w + 7) / 8) * ((jobs[k].acc->h + 7) / 8);
        deepest = std::max(deepest, jobs[k].scene->max_stack);
        coat = coat || jobs[k].scene->coat;
        glass = glass || jobs[k].scene->glass;
        tex = tex || jobs[k].scene->tex;
        local = local || j.args.params[20] > j.args.params[36];  // P.counts.x > P.counts2.x: local lights
        env = env || trt::from_bits(j.args.params[24]) >= 0.5f;  // P.env.x: an environment map
        // unlit materials (F_UNLIT) and a background (P.background.w: F_BACKGROUND)
        other = other || jobs[k].scene->unlit || jobs[k].scene->vcolors || trt::from_bits(j.args.params[99]) > 0.5f;
        tree = tree || trt::from_bits(j.args.params[87]) > 0.5f;  // P.ltree.w: the job samples the light tree
        table.push_back(j);
    }
    if (tiles >= (uint64_t{1} << 31)) return fail(TRT_E_INVALID, "trt_render_batch: too many tiles in one launch");
    if (table.empty()) return ok();
    // The wavefront takes batches whose jobs have at most kWfMaxSuns directional lights, no glass, no textures, no
    // local lights, no environment map, no unlit material, no vertex colours and no background; counting launches stay
    // on the megakernel. TRT_RENDER_AUTO takes it when
    // every scene has kWfAutoLeaves triangles or more, or for launches of a whole chunk of paths or more.
    const uint32_t how = d->render.load();
    bool wf = reads == nullptr && how != TRT_RENDER_MEGAKERNEL && !glass && !tex && !local && !env && !other;
    uint64_t paths = 0;
    bool large = true;
    std::vector<const trt_scene *> scene_of;
    for (uint32_t k = 0; k < n; k++) {
        if (jobs[k].spp == 0) continue;
        scene_of.push_back(jobs[k].scene);
        large = large && jobs[k].scene->leaf_count >= kWfAutoLeaves;
    }
    for (const auto &j : table) {
        paths += uint64_t(j.args.params[16]) * j.args.params[17] * j.args.spp;
        wf = wf && j.args.params[36] <= trt::kWfMaxSuns;  // P.counts2.x: the directional lights
    }
    if (wf && how == TRT_RENDER_AUTO) wf = large || paths >= kWfMaxPaths;
    std::vector<trt::WfSub> subs;
    std::vector<WfChunk> chunks;
    const uint64_t stride = (std::min(paths, kWfMaxPaths) + kWfStrideStep - 1) / kWfStrideStep * kWfStrideStep;
    if (wf) wf_plan(table, scene_of, stride, subs, chunks);
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    const int grid = wf ? d->cus * d->blocks_wide : d->cus * d->blocks_render;
    const uint32_t threads = static_cast<uint32_t>(grid) * (wf ? trt::kTraceBlock : trt::kRenderBlock);
    const uint64_t entries = deepest > trt::kLdsStack ? deepest - trt::kLdsStack : 1;
    int rc = ensure_spill(d, entries * threads * sizeof(uint32_t));
    if (rc) return rc;
    if (wf) {
        rc = ensure_wf(d, stride);
        if (rc) return rc;
    }
    for (auto &j : table) {
        j.args.geo.counter = d->counter;
        j.args.geo.guard_flag = d->guard;
        j.args.geo.spill = d->spill;
        j.args.geo.reads = d->reads;
        j.args.geo.spill_stride = threads;
        j.args.geo.spill_entries = static_cast<uint32_t>(entries);
    }
    // The job table goes through a ring of pinned host slots and device slots: a slot is written again only after
    // the launch that read it is done (its event), so launches queue without waiting for each other.
    auto &ring = d->jobs;
    const int k = ring.next;
    ring.next = (ring.next + 1) % kJobRing;
    if (ring.ev[k]) TRT_HIP(hipEventSynchronize(ring.ev[k]));
    const uint64_t job_bytes = (table.size() * sizeof(trt::RenderJob) + 255) / 256 * 256;
    const uint64_t bytes = job_bytes + subs.size() * sizeof(trt::WfSub);
    if (bytes > ring.cap[k]) {
        TRT_HIP(hipHostFree(ring.host[k]));
        TRT_HIP(hipFree(ring.dev[k]));
        ring.host[k] = nullptr, ring.dev[k] = nullptr, ring.cap[k] = 0;
        const uint64_t cap = std::max<uint64_t>(bytes, 64 * sizeof(trt::RenderJob));
        TRT_HIP(hipHostMalloc(&ring.host[k], cap));
        TRT_HIP(hipMalloc(&ring.dev[k], cap));
        ring.cap[k] = cap;
    }
    if (!ring.ev[k]) TRT_HIP(hipEventCreateWithFlags(&ring.ev[k], hipEventDisableTiming));
    std::memcpy(ring.host[k], table.data(), table.size() * sizeof(trt::RenderJob));
    if (!subs.empty())
        std::memcpy(static_cast<char *>(ring.host[k]) + job_bytes, subs.data(), subs.size() * sizeof(trt::WfSub));
    TRT_HIP(hipMemcpyAsync(ring.dev[k], ring.host[k], bytes, hipMemcpyHostToDevice, d->stream));
    if (wf) {
        const auto *d_jobs = static_cast<const trt::RenderJob *>(ring.dev[k]);
        const auto *d_subs = reinterpret_cast<const trt::WfSub *>(static_cast<const char *>(ring.dev[k]) + job_bytes);
        rc = wf_launch(d, d_jobs, d_subs, chunks, stride, table, coat, tree);
        if (rc) return rc;
    } else if (trt::launch_render(static_cast<const trt::RenderJob *>(ring.dev[k]), static_cast<uint32_t>(table.size()),
                           static_cast<uint32_t>(tiles), d->counter, d->spill, threads,
                           static_cast<uint32_t>(entries), d->guard, d->ray_counts, coat, tree, glass || tex || local || env || other, reads, grid,
                           d->stream))
        return fail(TRT_E_DEVICE, "render kernel launch failed");
    TRT_HIP(hipEventRecord(ring.ev[k], d->stream));
    return ok();
}

int trt_render_batch(trt_device *d, const trt_render_job *jobs, uint32_t n) { return render_batch(d, jobs, n, nullptr); }

int trt_render_reads(trt_scene *s, const void *params, uint32_t first_sample, uint32_t spp, trt_accum *acc,
                     uint64_t *reads) {
    if (!s || !reads) return fail(TRT_E_INVALID, "trt_render_reads: null argument");
    trt_device *d = s->dev;
    {
        std::lock_guard<std::mutex> lock(d->mu);
        TRT_HIP(hipSetDevice(d->ordinal));
        TRT_HIP(hipMemsetAsync(d->reads + kRenderReads, 0, 4 * sizeof(unsigned long long), d->stream));
    }
    const trt_render_job j{s, params, first_sample, spp, acc};
    int rc = render_batch(d, &j, 1, d->reads + kRenderReads);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lock(d->mu);
        unsigned long long r[4] = {0, 0, 0, 0};
        TRT_HIP(hipMemcpyAsync(r, d->reads + kRenderReads, sizeof(r), hipMemcpyDeviceToHost, d->stream));
        TRT_HIP(hipStreamSynchronize(d->stream));
        for (int k = 0; k < 4; k++) reads[k] = r[k];
    }
    return trt_device_sync(d);
}

int trt_render_ray_counts(trt_device *d, uint64_t *closest, uint64_t *shadow) {
    if (!d || !closest || !shadow) return fail(TRT_E_INVALID, "trt_render_ray_counts: null argument");
    {
        std::lock_guard<std::mutex> lock(d->mu);
        TRT_HIP(hipSetDevice(d->ordinal));
        std::vector<unsigned long long> r(trt::kRayCountBytes / sizeof(unsigned long long));
        TRT_HIP(hipMemcpyAsync(r.data(), d->ray_counts, trt::kRayCountBytes, hipMemcpyDeviceToHost, d->stream));
        TRT_HIP(hipMemsetAsync(d->ray_counts, 0, trt::kRayCountBytes, d->stream));
        TRT_HIP(hipStreamSynchronize(d->stream));
        uint64_t c = 0, sh = 0;
        for (uint32_t k = 0; k < trt::kShards; k++) c += r[k * trt::kRayShardGap], sh += r[k * trt::kRayShardGap + 1];
        *closest = c, *shadow = sh;
    }
    return trt_device_sync(d);
}

int trt_render(trt_scene *s, const void *params, uint32_t first_sample, uint32_t spp, trt_accum *acc) {
    if (!s) return fail(TRT_E_INVALID, "trt_render: null scene");
    const trt_render_job j{s, params, first_sample, spp, acc};
    trt::RenderArgs a;
    std::string why;
    int rc = job_args(j, a, &why);
    if (rc) return fail(rc, "trt_render: " + why);
    return trt_render_batch(s->dev, &j, 1);
}

int trt_denoise_atrous(trt_device *d, uint32_t w, uint32_t h, const float *color, const float *albedo,
                       const float *normal, const float *variance, uint32_t iterations, float *out) {
    if (!d || !color || !albedo || !normal || !variance || !out || w == 0 || h == 0)
        return fail(TRT_E_INVALID, "trt_denoise_atrous: bad argument");
    if (uint64_t(w) * h >= (uint64_t{1} << 31) || iterations > 16)
        return fail(TRT_E_INVALID, "trt_denoise_atrous: image too large or too many iterations");
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    const uint64_t need = uint64_t(w) * h * 9 * sizeof(float);
    if (need > d->denoise_bytes) {
        TRT_HIP(hipStreamSynchronize(d->stream));
        TRT_HIP(hipFree(d->denoise_scratch));
        d->denoise_scratch = nullptr;
        d->denoise_bytes = 0;
        TRT_HIP(hipMalloc(&d->denoise_scratch, need));
        d->denoise_bytes = need;
    }
    if (trt::launch_atrous(w, h, color, albedo, normal, variance, iterations, out,
                           static_cast<float *>(d->denoise_scratch), d->stream))
        return fail(TRT_E_DEVICE, "denoise kernel launch failed");
    return ok();
}

int trt_trace_rays_device(trt_scene *s, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags) {
    if (!s || (n && (!d_rays || !d_hits))) return fail(TRT_E_INVALID, "trt_trace_rays_device: null argument");
    std::lock_guard<std::mutex> lock(s->dev->mu);
    TRT_HIP(hipSetDevice(s->dev->ordinal));
    return launch(s, d_rays, n, d_hits, flags);
}

int trt_trace_rays_device_noalloc(trt_scene *s, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags) {
    if (!s || (n && (!d_rays || !d_hits)))
        return fail(TRT_E_INVALID, "trt_trace_rays_device_noalloc: null argument");
    std::lock_guard<std::mutex> lock(s->dev->mu);
    TRT_HIP(hipSetDevice(s->dev->ordinal));
    return launch(s, d_rays, n, d_hits, flags, false, true);
}

int trt_trace_rays(trt_scene *s, const trt_ray *rays, uint64_t n, trt_hit *hits, uint32_t flags) {
    if (!s || (n && (!rays || !hits))) return fail(TRT_E_INVALID, "trt_trace_rays: null argument");
    if (n == 0) return ok();
    trt_device *d = s->dev;
    {
        std::lock_guard<std::mutex> lock(d->mu);
        TRT_HIP(hipSetDevice(d->ordinal));
        const uint64_t need = n * (sizeof(trt_ray) + sizeof(trt_hit));
        if (need > d->staging_bytes) {
            TRT_HIP(hipStreamSynchronize(d->stream));
            TRT_HIP(hipFree(d->staging));
            d->staging = nullptr;
            d->staging_bytes = 0;
            TRT_HIP(hipMalloc(&d->staging, need));
            d->staging_bytes = need;
        }
        char *dr = static_cast<char *>(d->staging);
        char *dh = dr + n * sizeof(trt_ray);
        TRT_HIP(hipMemcpyAsync(dr, rays, n * sizeof(trt_ray), hipMemcpyHostToDevice, d->stream));
        int rc = launch(s, dr, n, dh, flags);
        if (rc) return rc;
        TRT_HIP(hipMemcpyAsync(hits, dh, n * sizeof(trt_hit), hipMemcpyDeviceToHost, d->stream));
    }
    return trt_device_sync(d);
}

}  // extern "C"
