/*
 * threert_amd: an MI355X-native (gfx950, HIP) core for ThreeRT's software path-trace pass.
 *
 * A plain C ABI: opaque handles, pointers and sizes, int status codes. No exceptions cross it; the library never
 * keeps a caller's host pointer after a call returns. Every entry point names the reference interface it stands in
 * for (paths relative to the ThreeRT repository root).
 *
 * Byte formats are the reference's kernel inputs:
 *   node      32 B  bvh.rs:6-15 GpuNode   (bmin[3], left_or_first, bmax[3], count; count 0 = interior)
 *   leaf      48 B  indexed.rs:19-26 Leaf (v0.xyz, section id bits), (e1.xyz, index bits), (e2.xyz, 0)
 *   instance 160 B  scene.rs:229-234 GpuInstance (m rows, n rows, info, p rows); units.rs:329-334 unit_record bits
 */
#ifndef THREERT_H
#define THREERT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference returns anyhow::Result; gpu.rs:2613, gpu.rs:4116). */
#define TRT_OK 0
#define TRT_E_INVALID 1 /* bad argument or scene data (the reference's upload/prepare errors)          */
#define TRT_E_OOM 2     /* device or host allocation failed                                           */
#define TRT_E_DEVICE 3  /* a HIP call failed                                                          */
#define TRT_E_GUARD 4   /* a kernel stopped a walk and reported instead of reading on: its iteration guard
                         * (gpu.rs:4199-4203 loop-guard trip), a tree reference outside the scene's buffers or a
                         * stack past its room, or a shading-table index outside its table (trt_last_error says) */

/* Trace modes (pathtrace.wgsl TRACE_CLOSEST / TRACE_SHADOW / TRACE_ANY). */
#define TRT_TRACE_CLOSEST 0u
#define TRT_TRACE_SHADOW 1u
#define TRT_TRACE_ANY 2u

/* trt_trace_rays flags: which node format the kernel walks. */
#define TRT_WALK_WIDE 0u   /* the library's 64-byte 4-wide quantised tree (default)            */
#define TRT_WALK_BINARY 1u /* the reference's own 32-byte binary node buffer, as uploaded      */
#define TRT_WALK_WATERTIGHT 2u /* the 4-wide walk with the watertight triangle test (Woop, Benthin and Wald 2013)
                                * in place of the reference's Moller-Trumbore: a ray on an edge two triangles share
                                * meets one of them. No counterpart in the reference (pathtrace.wgsl:742-760 is MT) */

#define TRT_NO_HIT 0xffffffffu /* trt_hit.tri / trt_hit.inst of a ray that hit nothing */

typedef struct trt_device trt_device;
typedef struct trt_scene trt_scene;

/* The last error's message on this thread ("" after success). */
const char *trt_last_error(void);

/* --- device (gpu.rs:2113-2248 Gpu::new / new_with) ------------------------------------------------------------ */
typedef struct {
    char name[64];
    uint32_t compute_units;
    uint32_t wave_size;
    uint64_t hbm_bytes;
    uint32_t clock_khz;
    uint32_t gcn_arch; /* e.g. 950 */
} trt_device_info;

int trt_device_create(int ordinal, trt_device **out);
int trt_device_release(trt_device *dev);
int trt_device_info_get(trt_device *dev, trt_device_info *out);
/* Wait for all work the library queued on the device's stream (the reference's device.poll(Wait)). */
int trt_device_sync(trt_device *dev);

/* Device memory for callers that keep their inputs resident in HBM (the reference's wgpu::Buffer role). */
int trt_mem_alloc(trt_device *dev, uint64_t bytes, void **dptr);
int trt_mem_free(trt_device *dev, void *dptr);
int trt_mem_upload(trt_device *dev, void *dst, const void *src, uint64_t bytes);
int trt_mem_download(trt_device *dev, void *dst, const void *src, uint64_t bytes);
/* Device-to-device copy queued on the device's stream (the reference's encoder.copy_buffer_to_buffer role); dst,
 * src and bytes multiples of 16. Its kernel is a plain streaming copy, the yardstick for HBM byte counters. */
int trt_mem_copy(trt_device *dev, void *dst, const void *src, uint64_t bytes);

/* HIP-event timer on the device's stream, the stream every kernel of the library runs on. */
int trt_timer_start(trt_device *dev);
int trt_timer_stop(trt_device *dev, float *ms);

/* --- host tree builders (prepare side: lib.rs:836-913 unit_trees) ------------------------------------------- */
typedef struct {
    void *nodes;         /* 32-byte nodes */
    uint64_t node_count;
    uint32_t *order;     /* order[i]: the item now at position i */
    uint64_t order_count;
    uint32_t depth;
    float lo[3], hi[3];  /* mesh trees: the model-space bounds */
} trt_tree;

/* bvh.rs:81-106 build_tris: tris n x 9 floats; max_depth as the reference's (48 for bvh::build). */
int trt_build_tris(const float *tris, uint64_t n, uint32_t max_depth, trt_tree *out);
/* units.rs:231-246 mesh_tree: binary SAH, median splits when deeper than 45. */
int trt_build_mesh_tree(const float *tris, uint64_t n, trt_tree *out);
/* units.rs:140-223 build_sah over boxes (n x 6: min xyz, max xyz). */
int trt_build_sah(const float *boxes, uint64_t n, uint32_t max_depth, trt_tree *out);
/* units.rs:84-118 build_median over boxes. */
int trt_build_median(const float *boxes, uint64_t n, uint32_t max_leaf, trt_tree *out);
/* units.rs:303-321 top_tree: leaves name records[]. */
int trt_build_top_tree(const float *boxes, const uint32_t *records, uint64_t n, trt_tree *out);
/* units.rs:288-297 UnitBox::new: world is a column-major 4x4; out = lo xyz, hi xyz. */
void trt_unit_box(const float *world, const float *lo, const float *hi, float *out);
/* units.rs:329-342 unit_record on a 160-byte record, in place. */
void trt_unit_record(void *record, uint32_t root, int cast_shadow);
void trt_tree_free(trt_tree *t);

/* ltree.rs:409-459 build (lib.rs:256-264 light_tree_of, without moving emitters or the slot for emitters added in
 * place): the light tree over n emitter triangles, given their world vertices (n x 9, as GpuEmitterTri::vertices),
 * centroids (n x 3) and weights (n; lib.rs:237-241 emitter_power; those not above 0 are left out). node_count 0: fewer
 * than 256 emitters weigh something, and the kernel samples the CDF alone (ltree.rs:31 MIN_STATIC). */
typedef struct {
    float *nodes;         /* node_count vec4s: two packed half-float pairs (centre x, y; centre z, radius), power, word */
    uint64_t node_count;
    uint32_t *order;      /* entry_count: the emitter of each tree entry, in leaf order (the CDF's first entries) */
    uint32_t *code;       /* entry_count: each entry's path code (depth + 1) << 26 | path */
    uint64_t entry_count;
    float power;          /* the root's */
} trt_light_tree;
int trt_build_light_tree(const float *verts, const float *centroids, const float *weights, uint64_t n,
                         trt_light_tree *out);
void trt_light_tree_free(trt_light_tree *t);

/* --- scene upload (gpu.rs:2613 Gpu::upload) ---------------------------------------------------------------- */
typedef struct {
    const void *nodes;       /* the node buffer: mesh trees, then the top tree (units.rs:532-595 Layout) */
    uint64_t node_count;
    const void *leaves;      /* 48-byte leaves, trees' order */
    uint64_t leaf_count;
    const void *instances;   /* 160-byte records */
    uint64_t instance_count;
    uint32_t top_first;      /* P.motion.z: the top tree's root */
    uint32_t unit_count;     /* P.motion.w: units in the top tree */
    /* What shading reads (trt_render); all may be NULL / 0 for a scene that only traces rays. */
    const void *vertices;     /* indexed.rs:10-13 Vertex, 48 B: (position, u), (normal, v), (tangent, hand) */
    uint64_t vertex_count;
    const uint32_t *indices;  /* 3 a triangle, in the leaves' order, counted from the record's first vertex */
    uint64_t index_count;     /* triangles (= leaf_count) */
    const uint32_t *rows;     /* scene.rs:99 Row, 4 words: material | class, section first, first emitter, flags */
    uint64_t row_count;
    const void *materials;    /* scene.rs:480-491 GpuMaterial, 288 B */
    uint64_t material_count;
    const void *lights;       /* scene.rs:495-501 GpuLight, 80 B: lib.rs:367-380's order, directional lights (type
                               * pos.w 0) first, then point (1), spot (2) and rect (3) lights, those that cast no
                               * shadow (ext.w 0) before those that do; other types, or a directional light after a
                               * local one, are refused */
    uint64_t light_count;
    const void *emitters;     /* lib.rs:451-481 emitter_cdf: entries, 16385 guide vec4s, header vec4 */
    uint64_t emitter_vec4s;
    const void *emitter_tris; /* lib.rs:143-157 GpuEmitterTri, 64 B */
    uint64_t emitter_tri_count;
    const void *objects;      /* object table, 3 vec4 an object (scene.rs:25 object_stride without instance colours;
                               * a tinted scene's 4 vec4 an object, the reference's F_TINT, is not taken); the
                               * first's w is its power */
    uint64_t object_count;
    const void *energy;       /* lib.rs:1088-1104 energy_lut: 512 vec4; NULL: no shading tables */
    const void *light_tree;   /* ltree.rs LightTree::nodes, one vec4 a node (gpu.rs:2613 `light_tree`); NULL / 0: none.
                               * Its emitters are the CDF's first P.ltree.y entries, in leaf order; their triangles'
                               * codes name their paths (lib.rs:419-441 emitter_order / emitter_codes) */
    uint64_t light_tree_vec4s;
    const void *glass_energy; /* lib.rs:1254 glass_energy_lut: 1024 vec4, R + T of the rough dielectric over 16 sqrt(cos)
                               * x 16 roughness x 8 sqrt((ior - 1) / (ior + 1)) nodes, entering then leaving. Needed
                               * (and read) only when a material has glass (layer.x > 0, the reference's F_GLASS) */
    /* texpack.rs Packed (gpu.rs:2433-2446 uploads it as an Rgba8Unorm 2D array of 5 levels): tex_layers layers of
     * RGBA8 texels, each layer its levels 0..4 one after the other (level l is tex_side >> l square, rows top to
     * bottom), layer after layer; tex_side a power of two >= 16. tex_info: Atlas::info, two vec4 a texture ((u0, v0,
     * du, dv) of its rect in page UV; (layer, wrap, width, height)). NULL tex_info: no textures (the reference's F_TEX
     * off); then no material may name a texture slot */
    const void *tex_pages;
    uint32_t tex_side;
    uint32_t tex_layers;
    const void *tex_info;
    uint64_t tex_info_vec4s;
    /* The environment map, as Gpu::upload takes it (gpu.rs:2613 `env`, `env_cdf`): env, an equirectangular map of
     * env_width x env_height RGBA float texels (lib.rs SkyBaked's EnvMap; row 0 the top, +y), kept as RGBA half floats
     * of min(v, 65000) (gpu.rs:3552-3561 write_env_texels); env_cdf, lib.rs:916-942 build_env_cdf's env_height +
     * env_width x env_height floats. Rendered where Params' env.x is set (env.yz its size); NULL env: none */
    const void *env;
    uint32_t env_width;
    uint32_t env_height;
    const void *env_cdf;
    uint64_t env_cdf_count;
    /* three.js vertexColors (lib.rs:601 gpu_vcolors, scene.rs:2324-2334 tri_colors): three RGBA8 sRGB words a
     * triangle, one a corner, in the leaves' order (as `indices`); vcolor_words = 3 x index_count. Read for materials
     * whose uv.w is set; NULL: none, and uv.w is ignored, as the reference does without F_VCOLOR */
    const void *vcolors;
    uint64_t vcolor_words;
} trt_scene_desc;

typedef struct {
    uint64_t wide_nodes;     /* 64-byte nodes of the 4-wide tree */
    uint64_t wide_bytes;
    uint64_t binary_bytes;   /* the reference node buffer as uploaded */
    uint64_t leaf_bytes;
    uint64_t unit_bytes;
    uint32_t max_stack;      /* the deepest a walk's stack can get over the 4-wide tree */
    uint32_t units;
    uint64_t device_bytes;   /* every device allocation the scene holds */
    float build_ms;          /* the upload's wide-tree build (own mesh trees, collapse, quantisation), wall clock */
    float device_build_ms;   /* of which the mesh trees' build on the device, first kernel to last */
    uint32_t device_meshes;  /* mesh trees built on the device */
    uint32_t host_meshes;    /* mesh trees built on the host (TRT_BUILD_HOST, or cases the device leaves to it) */
} trt_scene_stats;

/* Copies the host buffers to HBM and builds the 4-wide quantised tree from the node buffer. The caller's memory
 * is borrowed for the call only. Refuses (TRT_E_INVALID) trees the reference's checks refuse (units.rs check_mesh /
 * check_top, ltree.rs LightTree::check), records with cutout bits or with glass among shadow-casting sections (the
 * reference's F_IGLASS; a record whose sections are all glass casts no shadow and is taken), materials that need
 * what this core does not render (a texture slot naming no texture of the pages, alpha cutoffs, material
 * functions), glass without the glass energy table, and vertex colours of another length than 3 x index_count, each
 * named in trt_last_error. */
int trt_scene_upload(trt_device *dev, const trt_scene_desc *desc, trt_scene **out);
int trt_scene_release(trt_scene *scene);
int trt_scene_stats_get(trt_scene *scene, trt_scene_stats *out);

/* Where later uploads on this device build the library's own mesh trees and their 4-wide nodes:
 * TRT_BUILD_DEVICE (binned SAH one tree level at a time over every mesh together, then the collapse, the
 * quantisation and the leaves' layout, on a stream of the upload's own), TRT_BUILD_HOST (one host thread), or
 * TRT_BUILD_AUTO (the default: the device for scenes of 65,536 mesh triangles or more, where it is faster). All give
 * the same bytes; a mesh the device cannot build as the host would (a median split of distinct centroids, past
 * depth 64 or when every split's cost overflows) is built on the host. No reference counterpart: the reference
 * builds its trees on the host in `prepare` (lib.rs:836-913, bvh.rs:130-239); the library's own trees replace only
 * the walk's copy of them. */
#define TRT_BUILD_DEVICE 0u
#define TRT_BUILD_HOST 1u
#define TRT_BUILD_AUTO 2u
int trt_device_set_build(trt_device *dev, uint32_t where);

/* Copies one of the scene's wide-tree buffers to host memory: which = 0 the 64-byte wide nodes
 * (stats.wide_nodes of them), 1 the leaves in the wide trees' order (48 bytes each, as many as the scene's leaves),
 * 2 the 64-byte units (one a record). `bytes` must be the buffer's size. No reference counterpart (for checking
 * that builds agree). */
int trt_scene_wide_read(trt_scene *scene, uint32_t which, void *dst, uint64_t bytes);

/* --- ray batches (trace_units.wgsl:44-135 via pathtrace.wgsl:764 `trace`; units.rs:1019 Kernel::trace) ------- */
typedef struct {
    float o[3];
    float tmax;    /* > 3e38 (or NaN) is taken as 3e38, the reference's INF */
    float d[3];
    uint32_t mode; /* TRT_TRACE_* */
} trt_ray;         /* 32 bytes */

typedef struct {
    float t;       /* tmax when nothing is hit */
    uint32_t tri;  /* leaf position (first + tree position), TRT_NO_HIT when nothing is hit */
    float u, v;
    uint32_t inst; /* record, TRT_NO_HIT when nothing is hit */
    uint32_t part; /* the leaf's section id (leaf_id(tri, 0)) */
} trt_hit;         /* 24 bytes: pathtrace.wgsl Hit */

/* Closest hits are the lowest (t, instance, triangle): the same answer whatever tree is walked. Any-hit modes
 * report some hit exactly when there is one. Host buffers; blocks until the hits are written. */
int trt_trace_rays(trt_scene *scene, const trt_ray *rays, uint64_t n, trt_hit *hits, uint32_t flags);
/* The same with device buffers (trt_mem_alloc), queued on the device's stream; trt_device_sync waits. */
int trt_trace_rays_device(trt_scene *scene, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags);
/* The same as trt_trace_rays_device, allocating nothing: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * walk's ray counter and its stack entries past LDS live in a fixed 32 MiB scratch array, one per GPU, held until the
 * process ends; a tree too deep for it at full occupancy walks on fewer workgroups, and one whose stack does not fit
 * even one workgroup is refused (TRT_E_INVALID). Every call on a GPU shares that GPU's scratch: calls through one
 * trt_device are ordered on its stream, but calls on two handles of the same GPU that overlap race. Queued on the
 * device's stream; trt_device_sync waits and reports the walk's guard. */
int trt_trace_rays_device_noalloc(trt_scene *scene, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits,
                                  uint32_t flags);
/* The same walk, counting what it reads (units.rs:922-927 Reads, units_cost.rs): reads[0] nodes fetched (64-byte
 * wide or 32-byte binary), reads[1] triangles tested (48 bytes each), reads[2] unit records read (64-byte wide
 * units or 160-byte records). Blocks. */
int trt_trace_reads(trt_scene *scene, const trt_ray *d_rays, uint64_t n, trt_hit *d_hits, uint32_t flags,
                    uint64_t *reads);
/* The FP64 check of a batch's hits (no counterpart in the reference, whose tests compare hits and never a distance
 * against double precision; it is the precision workload's oracle on the GPU): each hit's ray is placed in its unit
 * and the hit triangle intersected again, every operation in double. d_t64[i] = the FP64 distance, 0 where ray i hit
 * nothing, -1 where the FP64 test misses the triangle the walk reported (a disputed ray), -2 where the hit names a
 * triangle or unit outside the scene. Device buffers, queued on the device's stream. */
int trt_trace_check_fp64(trt_scene *scene, const trt_ray *d_rays, const trt_hit *d_hits, uint64_t n, double *d_t64);

/* --- path tracing (gpu.rs:4116-4236 Gpu::render; pathtrace.wgsl:4106 `main`, offline) ---------------------- */
typedef struct trt_accum trt_accum;

/* The accumulator the samples add to: 80 bytes a pixel (records.wgsl:24 Pixel: color, albedo, normal, and two
 * vec4 the offline pass leaves at zero), in HBM (gpu.rs:2388 accum_buffer). Created cleared. */
int trt_accum_create(trt_device *dev, uint32_t width, uint32_t height, trt_accum **out);
int trt_accum_release(trt_accum *acc);
int trt_accum_clear(trt_accum *acc);
/* The raw sums, width x height x 20 floats. Waits for the device. */
int trt_accum_read(trt_accum *acc, float *pixels);
/* gpu.rs:4218-4231: color, albedo, normal divided by the sample count (3 floats a pixel each, any may be NULL),
 * variance of the mean luminance (1 float a pixel, may be NULL). Waits for the device. */
int trt_accum_resolve(trt_accum *acc, float *color, float *albedo, float *normal, float *variance);
/* The same means into device images of pw x ph >= the accumulator's width x height (3 floats a pixel each, rows and
 * columns past the accumulator's repeating its last), on the device's stream: the input of trt_unet_denoise without
 * a trip through host memory (the denoiser wants sides that are multiples of 16). No counterpart in the reference,
 * whose denoisers read the frame back first (denoise.rs:10-103). */
int trt_accum_resolve_device(trt_accum *acc, uint32_t pw, uint32_t ph, float *d_color, float *d_albedo, float *d_normal);
/* Converged reference images in FP64 on the device: d_sum (4 doubles a pixel, width x height, zeroed by the caller)
 * += this accumulator's color sums (a non-finite sum adds zero) and its sample count; the caller renders a
 * reference in launches of a few samples, each into a cleared accumulator, and adds each once. No counterpart in the
 * reference, which keeps FP32 sums only (records.wgsl Pixel). */
int trt_accum_add_fp64(trt_accum *acc, double *d_sum);

/* Samples first_sample .. first_sample + spp - 1 of every pixel added to `acc` (gpu.rs:4177-4196: the sample index
 * is P.size.z), on the device's stream. params: the 624-byte Params block (pathtrace.wgsl:96-140); its size.xy
 * must be the accumulator's. Needs a scene uploaded with its shading tables. Refuses (TRT_E_INVALID, named) what the
 * offline `main` would read and this core does not render: an environment map the scene has not (P.env.x without
 * the scene description's env, or P.env.yz another size), fog, volumetric scattering,
 * procedural geometry, the live light LOD, aerial perspective, more than 64 bounces, and
 * light-tree candidates (P.ltree.w) for a scene uploaded without a tree; and light counts the scene does not hold:
 * P.counts.x past its lights, P.counts2.x other than its directional lights when local lights follow, P.counts2.w
 * (the first shadow-casting local light, lib.rs:700) outside the local lights. */
int trt_render(trt_scene *scene, const void *params, uint32_t first_sample, uint32_t spp, trt_accum *acc);
/* The rays the device's render launches traced since the last call, which resets them: closest-hit rays (camera and
 * bounce) and shadow rays (sun and emitter visibility). Waits for the device. A measurement with no counterpart in
 * the reference, whose passes report timings only (gpu.rs:4198-4208). */
int trt_render_ray_counts(trt_device *dev, uint64_t *closest, uint64_t *shadow);

/* trt_render as one counting launch (the same samples into `acc`, the same result), blocking: reads[0] 64-byte wide
 * nodes fetched, reads[1] triangles tested (48 bytes each), reads[2] 64-byte units entered, by every ray the launch
 * traced; reads[3] hits shaded. A measurement with no counterpart in the reference (units.rs:922-927 Reads counts
 * the CPU walk's reads the same way). */
int trt_render_reads(trt_scene *scene, const void *params, uint32_t first_sample, uint32_t spp, trt_accum *acc,
                     uint64_t *reads);

/* Several render jobs in one launch (continuous batching for a render server; the reference's `serve` loop,
 * main.rs:345-398, renders one job at a time and has no counterpart): each job is trt_render's arguments; their
 * tiles share the GPU. A job's accumulator must not appear twice in one batch. Each job's result is the one
 * trt_render gives it alone. Queued on the device's stream. */
typedef struct {
    trt_scene *scene;
    const void *params; /* 624-byte Params */
    uint32_t first_sample;
    uint32_t spp;
    trt_accum *acc;
} trt_render_job;
int trt_render_batch(trt_device *dev, const trt_render_job *jobs, uint32_t n);

/* How the device's later render launches (trt_render, trt_render_batch) run their paths: TRT_RENDER_MEGAKERNEL (a
 * thread follows one pixel's path through every bounce, tracing as it goes), TRT_RENDER_WAVEFRONT (up to 4,194,304
 * paths at a time, bounce by bounce: one walk launch traces every path's next ray and the previous bounce's shadow
 * rays from a compacted queue, then a shade launch resolves the shadow rays, shades the hits and queues the next
 * rays; about 600 bytes of HBM a path), or TRT_RENDER_AUTO (the default: the wavefront for scenes of 1,024 triangles
 * or more, where it renders 1.4-4.4x as fast, and for launches of 4,194,304 paths or more; the megakernel for
 * smaller launches of smaller scenes, where the path state costs more than it saves).
 * Both give the same bits.
 * The wavefront takes a batch one scene at a time (AUTO: when every scene has 1,024 triangles or more). A batch with a
 * job of more than 3 directional lights, and trt_render_reads, run as the megakernel whatever the setting. No
 * reference counterpart: the reference's offline pass is one megakernel dispatch (pathtrace.wgsl:4106 `main`). */
#define TRT_RENDER_AUTO 0u
#define TRT_RENDER_MEGAKERNEL 1u
#define TRT_RENDER_WAVEFRONT 2u
int trt_device_set_render(trt_device *dev, uint32_t how);

/* --- denoising (denoise.rs:10-74 denoise: edge-avoiding a-trous guided by albedo and normal) ---------------- */
/* Device buffers (trt_mem_alloc): color, albedo, normal 3 floats a pixel, variance 1 (trt_accum_resolve's
 * outputs), out 3 floats a pixel. Queued on the device's stream. */
int trt_denoise_atrous(trt_device *dev, uint32_t width, uint32_t height, const float *d_color, const float *d_albedo,
                       const float *d_normal, const float *d_variance, uint32_t iterations, float *d_out);

/* --- learned denoiser (replaces denoise.rs:81-103 denoise::oidn; OIDN's network is not in the reference) ------ */
/* A five-level U-Net of 3x3 convolutions (32 to 128 channels) on BF16 MFMA with FP32 accumulation, trained and
 * run on the device. Inputs per pixel (3 floats each): color, albedo, normal, as trt_accum_resolve gives them; the
 * output is the denoised color. Images: n of h x w, h and w multiples of 16, n * h * w at most max_pixels. The
 * parameters are FP32 (trt_unet_param_count of them: each layer's weights [cout][9][cin] then its biases); every
 * call but create / release / param_count / params_get is queued on the device's stream. */
typedef struct trt_unet trt_unet;
int trt_unet_create(trt_device *dev, uint64_t max_pixels, trt_unet **out);
int trt_unet_release(trt_unet *net);
int trt_unet_param_count(trt_unet *net, uint64_t *count);
/* The level height from which the convolutions take 16-pixel-tall tiles instead of 8 (default 128; UINT32_MAX: never);
 * the same sums either way, so the same bits. Process-wide. A tuning and test knob with no counterpart in the
 * reference. */
int trt_unet_set_tall_tiles(uint32_t min_rows);
/* Sets the parameters (host memory) and restarts the optimizer. */
int trt_unet_params_set(trt_unet *net, const float *params, uint64_t count);
int trt_unet_params_get(trt_unet *net, float *params, uint64_t count);
/* Device buffers in and out (the reference: denoise::oidn(width, height, color, albedo, normal)). */
int trt_unet_denoise(trt_unet *net, uint32_t n, uint32_t h, uint32_t w, const float *d_color, const float *d_albedo,
                     const float *d_normal, float *d_out);
/* The loss (mean |y - log(1 + target)| over the three color channels, y the network's log-space output) and its
 * gradient with respect to every parameter, into host memory; waits for the device. d_target: 3 floats a pixel. */
int trt_unet_gradients(trt_unet *net, uint32_t n, uint32_t h, uint32_t w, const float *d_color, const float *d_albedo,
                       const float *d_normal, const float *d_target, float *loss, float *grads);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; loss (host, may be NULL: then nothing waits) gets the
 * step's loss before the update. */
int trt_unet_train_step(trt_unet *net, uint32_t n, uint32_t h, uint32_t w, const float *d_color,
                        const float *d_albedo, const float *d_normal, const float *d_target, float lr, float *loss);

#ifdef __cplusplus
}
#endif
#endif
