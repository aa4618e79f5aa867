/*
 * Oracle (test infrastructure only): the sampler and the offline path loop (pathtrace.wgsl `main`, LIVE = false)
 * restated in C for the features the workloads use. See render.c for the list and the reference lines.
 */
#ifndef THREERT_ORACLE_RENDER_H
#define THREERT_ORACLE_RENDER_H

#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t seed, index, path, depth; /* RNG_SEED, RNG_INDEX, RNG_PATH, cur_depth */
} or_rng;

uint32_t or_reverse_bits(uint32_t x);
uint32_t or_owen_rev(uint32_t n, uint32_t seed);
uint32_t or_hash_hp(uint32_t i);
float or_sobol_owen(uint32_t rev_index, uint32_t dim, uint32_t seed);
void or_sobol_dirs(uint32_t *out96);
uint32_t or_rng_dim(const or_rng *g, uint32_t slot);
uint32_t or_rng_sub(uint32_t d, uint32_t k);
float or_rng1(const or_rng *g, uint32_t d);
void or_rng2(const or_rng *g, uint32_t d, float *r);
void or_rng3(const or_rng *g, uint32_t d, float *r);
void or_rng4(const or_rng *g, uint32_t d, float *r);
void or_rng_start(or_rng *g, uint32_t px, uint32_t py, uint32_t index, uint32_t seed, uint32_t path);
void or_rng4_at(uint32_t pixel_seed, uint32_t index, uint32_t d, float *r);

/* Cephes-style sin/cos (Moshier's single-precision reduction and polynomials), fmaf evaluation. */
void or_sincos(float x, float *s, float *c);

/* What the offline kernel reads besides the trees (gpu.rs:2613 upload; heap.rs tables and words). */
typedef struct {
    or_scene geo;
    const float *vertices;      /* indexed.rs Vertex: 12 floats */
    uint64_t vertex_count;
    const uint32_t *indices;    /* 3 per triangle, in the leaves' order, counted from the mesh's first vertex */
    uint64_t tri_count;
    const uint32_t *rows;       /* scene.rs Row: 4 words */
    uint64_t row_count;
    const float *materials;     /* scene.rs GpuMaterial: 72 floats */
    uint64_t material_count;
    const float *lights;        /* scene.rs GpuLight: 20 floats */
    uint32_t light_count;
    const float *emitters;      /* lib.rs emitter_cdf: entries, guide, header (vec4s) */
    uint64_t emitter_vec4s;
    const float *emitter_tris;  /* lib.rs GpuEmitterTri: 16 words */
    uint64_t emitter_tri_count;
    const float *objects;       /* vec4s, OBJECT_STRIDE (3) a object; [k*3].w its power */
    uint64_t object_count;
    const float *energy;        /* lib.rs energy_lut: 512 vec4 */
    const float *light_tree;    /* ltree.rs LightTree::nodes: vec4s (NULL: none) */
    uint64_t light_tree_count;
    /* lib.rs glass_energy_lut: 1024 vec4, entering then leaving. Non-NULL: a material has glass (the reference's F_GLASS,
     * gpu.rs features_of: layer.x > 0), and the glass paths run */
    const float *glass_energy;
    /* texpack.rs Packed: the texture pages, RGBA8, tex_layers layers of MIPS = 5 levels (level l is tex_side >> l
     * square), level after level within a layer, layer after layer; and Atlas::info, two vec4 a texture ((u0, v0, du,
     * dv) of its rect in page UV; (layer, wrap, width, height)). tex_info NULL: no textures (the reference's F_TEX
     * off) */
    const uint8_t *tex_pages;
    uint32_t tex_side, tex_layers;
    const float *tex_info;
    uint64_t tex_info_vec4s;
    /* The environment map as gpu.rs:3537-3561 env_texture holds it: RGBA half floats (each value min(v, 65000)
     * rounded to half), env_w x env_h texels, row after row from the top (v = 0, +y); and lib.rs:916-942
     * build_env_cdf, env_h + env_w x env_h floats (the rows' CDF, then each row's). env NULL: none (P.env.x must be
     * 0) */
    const uint16_t *env;
    uint32_t env_w, env_h;
    const float *env_cdf;
    uint64_t env_cdf_count;
    /* lib.rs:601 gpu_vcolors (scene.rs:2324-2334 tri_colors): three RGBA8 sRGB colours a triangle, one a corner, in
     * the leaves' order; NULL: none (the reference's F_VCOLOR off; a material's uv.w flag is then ignored) */
    const uint32_t *vcolors;
    uint64_t vcolor_words;
} or_render_scene;

/* One sample (params word 18 = P.size.z) of every pixel of rows [y0, y1), accumulated into `accum` (80 bytes a
 * pixel: records.wgsl Pixel). params: the 624-byte Params block (pathtrace.wgsl:96-140). Returns the number of
 * walks that failed (read outside their buffers) plus the table reads whose index was outside its table. */
int64_t or_render_sample(const or_render_scene *s, const uint32_t *params, uint32_t y0, uint32_t y1, float *accum,
                         int threads);

/* pathtrace.wgsl:547-618 emitter_draw for given uniforms u[3] at the point x with normal n (*entry: the CDF entry it
 * picks, -1 none; returns its probability p) and emitter_pmf of emitter triangle ei there; -1 when a read fell outside
 * its table. For tests of the emitter sampling. */
float or_emitter_draw(const or_render_scene *s, const uint32_t *params, const float *x, const float *n,
                      const float *u, int64_t *entry);
float or_emitter_pmf(const or_render_scene *s, const uint32_t *params, const float *x, const float *n, uint32_t ei);
/* glass_energy (pathtrace.wgsl:1582-1603) in the scene's table (-1: read outside it); glass_at's share (930-970) of
 * a material at |cos| cos_d under `params` */
float or_glass_energy(const or_render_scene *s, float nov, float roughness, float ior, int leaving);
float or_glass_share(const or_render_scene *s, const uint32_t *params, uint32_t material, float cos_d);
/* The kernel's log2 and exp2 (Cephes log2f / exp2f); tex (pathtrace.wgsl:1107-1117) of texture `id` at uv with LOD
 * `lod` into out[4] (returns -1 when it read outside its tables) */
float or_log2(float x);
/* The kernel's atan2 and acos (Cephes atanf / asinf); env_radiance, env_pdf and sample_env (pathtrace.wgsl:1889-1941)
 * of the scene's environment under `params` (sample_env for the uniforms u[4]); -1 / NaN when a read fell outside
 * its table */
float or_atan2(float y, float x);
float or_acos(float x);
int or_env(const or_render_scene *s, const uint32_t *params, const float *d, const float *u, float *radiance,
           float *pdf, float *sample);
float or_exp2(float x);
int or_tex(const or_render_scene *s, int32_t id, float u, float v, float lod, float *out);

/* denoise.rs:10-74 denoise (a-trous), and the expf it uses here. Images: 3 floats a pixel (variance: 1). */
float or_expf(float x);
void or_denoise_atrous(uint32_t w, uint32_t h, const float *color, const float *albedo, const float *normal,
                       const float *variance, uint32_t iterations, float *out);

#ifdef __cplusplus
}
#endif
#endif
