// The C ABI of libthreert (include/threert.h): devices, memory, host tree builders, scene upload, ray batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/threert.h"
#include "trace_kernels.h"
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

struct trt_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t *counter = nullptr;
    uint32_t *guard = nullptr;
    uint32_t *spill = nullptr;
    uint64_t spill_bytes = 0;
    unsigned long long *reads = nullptr;
    void *staging = nullptr;  // trt_trace_rays' device copies of the caller's rays and hits
    uint64_t staging_bytes = 0;
    int cus = 0;
    int blocks_wide = 1, blocks_binary = 1;
    hipDeviceProp_t props{};
    std::mutex mu;
};

struct trt_scene {
    trt_device *dev = nullptr;
    void *nodes = nullptr, *leaves = nullptr, *instances = nullptr, *wnodes = nullptr, *wunits = nullptr;
    uint64_t node_count = 0, leaf_count = 0, instance_count = 0;
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
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->counter), 256);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->guard), 256);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&d->reads), 256);
    if (e == hipSuccess) e = hipMemset(d->guard, 0, 256);
    if (e != hipSuccess) {
        trt_device_release(d);
        return fail(TRT_E_DEVICE, std::string("trt_device_create: ") + hipGetErrorString(e));
    }
    d->cus = d->props.multiProcessorCount;
    if (trt::trace_occupancy(true, &d->blocks_wide) || d->blocks_wide < 1) d->blocks_wide = 1;
    if (trt::trace_occupancy(false, &d->blocks_binary) || d->blocks_binary < 1) d->blocks_binary = 1;
    *out = d;
    return ok();
}

int trt_device_release(trt_device *d) {
    if (!d) return ok();
    hipSetDevice(d->ordinal);
    if (d->stream) hipStreamSynchronize(d->stream);
    hipFree(d->counter);
    hipFree(d->guard);
    hipFree(d->reads);
    hipFree(d->spill);
    hipFree(d->staging);
    if (d->ev0) hipEventDestroy(d->ev0);
    if (d->ev1) hipEventDestroy(d->ev1);
    if (d->stream) hipStreamDestroy(d->stream);
    delete d;
    return ok();
}

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
    TRT_HIP(hipSetDevice(d->ordinal));
    TRT_HIP(hipStreamSynchronize(d->stream));
    uint32_t g = 0;
    TRT_HIP(hipMemcpy(&g, d->guard, 4, hipMemcpyDeviceToHost));
    if (g) {
        TRT_HIP(hipMemset(d->guard, 0, 4));
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
    delete s;
    return ok();
}

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
    if (desc->instance_count >= (uint64_t{1} << 30) || desc->leaf_count >= (uint64_t{1} << 32))
        return fail(TRT_E_INVALID, "trt_scene_upload: too many records or leaves");
    const auto *nodes = static_cast<const trt::Node *>(desc->nodes);
    const auto *inst = static_cast<const trt::Instance *>(desc->instances);
    trt::WideScene w;
    std::string why;
    if (!trt::build_wide(nodes, desc->node_count, inst, desc->instance_count, desc->leaf_count, desc->top_first,
                         desc->unit_count, w, &why))
        return fail(TRT_E_INVALID, "trt_scene_upload: " + why);
    if (w.nodes.size() >= (uint64_t{1} << 30)) return fail(TRT_E_INVALID, "trt_scene_upload: tree too large");
    std::lock_guard<std::mutex> lock(d->mu);
    TRT_HIP(hipSetDevice(d->ordinal));
    auto *s = new (std::nothrow) trt_scene();
    if (!s) return fail(TRT_E_OOM, "trt_scene_upload: out of host memory");
    s->dev = d;
    uint64_t total = 0;
    int rc = upload(d, &s->nodes, desc->nodes, desc->node_count * sizeof(trt::Node), &total);
    if (!rc) rc = upload(d, &s->leaves, desc->leaves, desc->leaf_count * sizeof(trt::Leaf), &total);
    if (!rc) rc = upload(d, &s->instances, desc->instances, desc->instance_count * sizeof(trt::Instance), &total);
    if (!rc) rc = upload(d, &s->wnodes, w.nodes.data(), w.nodes.size() * sizeof(trt::WideNode), &total);
    if (!rc) rc = upload(d, &s->wunits, w.units.data(), w.units.size() * sizeof(trt::WideUnit), &total);
    if (rc) {
        std::string keep = g_error;
        trt_scene_release(s);
        return fail(rc, keep);
    }
    s->node_count = desc->node_count;
    s->leaf_count = desc->leaf_count;
    s->instance_count = desc->instance_count;
    s->top_first = desc->top_first;
    s->units = desc->unit_count;
    s->top_ref = w.top_ref;
    s->max_stack = w.max_stack;
    std::memcpy(s->top_box, w.top_box, sizeof(s->top_box));
    s->stats.wide_nodes = w.nodes.size();
    s->stats.wide_bytes = w.nodes.size() * sizeof(trt::WideNode);
    s->stats.binary_bytes = desc->node_count * sizeof(trt::Node);
    s->stats.leaf_bytes = desc->leaf_count * sizeof(trt::Leaf);
    s->stats.unit_bytes = w.units.size() * sizeof(trt::WideUnit);
    s->stats.max_stack = w.max_stack;
    s->stats.units = desc->unit_count;
    s->stats.device_bytes = total;
    *out = s;
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

static int launch(trt_scene *s, const void *d_rays, uint64_t n, void *d_hits, uint32_t flags, bool count = false) {
    trt_device *d = s->dev;
    const bool wide = flags == TRT_WALK_WIDE;
    if (flags != TRT_WALK_WIDE && flags != TRT_WALK_BINARY) return fail(TRT_E_INVALID, "unknown trace flags");
    if (n >= (uint64_t{1} << 32) - 64) return fail(TRT_E_INVALID, "a batch of 2^32 rays or more");
    if (n == 0) return ok();
    const int grid = d->cus * (wide ? d->blocks_wide : d->blocks_binary);
    const uint32_t threads = static_cast<uint32_t>(grid) * trt::kTraceBlock;
    const uint32_t deepest = wide ? s->max_stack : trt::kRefStack;
    const uint64_t entries = deepest > trt::kLdsStack ? deepest - trt::kLdsStack : 1;
    int rc = ensure_spill(d, entries * threads * sizeof(uint32_t));
    if (rc) return rc;
    trt::TraceArgs a{};
    a.wnodes = static_cast<const uint4 *>(s->wnodes);
    a.wunits = static_cast<const float4 *>(s->wunits);
    a.leaves = static_cast<const float4 *>(s->leaves);
    a.bnodes = static_cast<const float4 *>(s->nodes);
    a.instances = static_cast<const float4 *>(s->instances);
    a.rays = static_cast<const float4 *>(d_rays);
    a.hits = static_cast<uint32_t *>(d_hits);
    a.counter = d->counter;
    a.guard_flag = d->guard;
    a.spill = d->spill;
    a.reads = d->reads;
    a.n = n;
    a.spill_stride = threads;
    a.top_ref = s->top_ref;
    a.top_first = s->top_first;
    a.units = s->units;
    std::memcpy(a.top_box, s->top_box, sizeof(a.top_box));
    if (trt::launch_trace(a, wide, count, grid, d->stream)) return fail(TRT_E_DEVICE, "trace kernel launch failed");
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

int trt_trace_rays_device(trt_scene *s, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags) {
    if (!s || (n && (!d_rays || !d_hits))) return fail(TRT_E_INVALID, "trt_trace_rays_device: null argument");
    std::lock_guard<std::mutex> lock(s->dev->mu);
    TRT_HIP(hipSetDevice(s->dev->ordinal));
    return launch(s, d_rays, n, d_hits, flags);
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
