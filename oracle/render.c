/*
 * Oracle (test infrastructure only): the offline path loop, pathtrace.wgsl `main` (LIVE = false) restated for the
 * features the workloads use: pinhole camera, instanced triangle meshes with per-section materials (Lambert + GGX
 * metal-rough, emission, clearcoat, glass, textures), directional lights (soft cone), point, spot and rect lights,
 * emissive triangles from the CDF or the light tree (RIS at camera hits), a constant sky or an environment map with
 * light sampling, MIS,
 * Russian roulette, the indirect clamp, FP32 accumulation. Not restated: media, fog, procedural geometry, cutouts,
 * live-only paths.
 *
 *   pathtrace.wgsl:4002-4024 camera_ray       4106-4557 main (offline parts)
 *   pathtrace.wgsl:279-283 inst_vec/inst_sign  286-288 inst_normal  310-313 material_of  330-360 tri_data
 *   pathtrace.wgsl:406-413 inst_attr          430-432 seen_directly
 *   pathtrace.wgsl:463-487 emitter totals / pick   547-560 emitter_draw (no tree)   590-593 emitter_pmf (no tree)
 *   pathtrace.wgsl:846-855 shadow_tr          909-915 light_vis   975-977 sky_tr   1066-1070 emitter_tr
 *   pathtrace.wgsl:1312-1317 emission         1320-1425 surface (no functions)
 *   pathtrace.wgsl:1492-1505 energy_lookup    1521-1525 spec_terms   1642-1699 lobes
 *   pathtrace.wgsl:1707-1730 d_ggx / vis_smith_joint_approx / f_schlick   1733-1764 eval_bsdf_k   1766-1789 g1,
 *   pdf_bsdf   1797-1817 light_mis   1819-1844 basis / sample_vndf   1848-1884 sample_bsdf
 *   pathtrace.wgsl:1889-1941 env_uv / env_radiance / env_pdf / lower_bound / sample_env   1943-1960 power_heuristic,
 *   luminance, offset_origin   1963-1988 sky_mis / sun_disks   2009-2024 emit_mis
 *   pathtrace.wgsl:2269-2393 emitter_radiance / sample_emitter_ris / sample_emitter   2431-2442 env_light
 *   pathtrace.wgsl:2042-2104 local_light    2158-2262 local_lights (offline: no light tiles, no volumes)
 *   pathtrace.wgsl:2444-2540 direct_light (no light_pick)
 *   textures, glass, the clearcoat and the light tree: cited where each is restated below
 * WGSL leaves normalize, inverseSqrt, sin and cos to the driver; here normalize(v) = v / sqrt(v.v),
 * inverseSqrt(x) = 1 / sqrt(x), sin/cos by or_sincos. Everything else rounds in the order the reference writes it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_render.h"

#define PI_F 3.14159265358979f
static const float INV_PI = 1.0f / PI_F;
static const float TWO_PI = 2.0f * PI_F;
static const float FOUR_PI = 4.0f * PI_F;

/* ---- log2 / exp2: Cephes log2f and exp2f (S. L. Moshier): the mantissa in [sqrt(1/2), sqrt(2)) and a minimax
 * polynomial in it; exp2 by its integer part and a polynomial on [-1/2, 1/2], both in fmaf. The kernel has the same
 * steps, so the same bits. */
float or_log2(float xx) {
    if (!(xx > 0.0f)) return xx == 0.0f ? -INFINITY : NAN;
    if (xx == INFINITY) return INFINITY;
    int e;
    float x = frexpf(xx, &e);
    if (x < 0.70710678118654752f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    float z = x * x;
    float y = fmaf(7.0376836292e-2f, x, -1.1514610310e-1f);
    y = fmaf(y, x, 1.1676998740e-1f);
    y = fmaf(y, x, -1.2420140846e-1f);
    y = fmaf(y, x, 1.4249322787e-1f);
    y = fmaf(y, x, -1.6668057665e-1f);
    y = fmaf(y, x, 2.0000714765e-1f);
    y = fmaf(y, x, -2.4999993993e-1f);
    y = fmaf(y, x, 3.3333331174e-1f);
    y = y * x * z;
    y = fmaf(-0.5f, z, y);
    const float LOG2EA = 0.44269504088896340736f;
    float r = y * LOG2EA;
    r = fmaf(x, LOG2EA, r);
    r = r + y;
    r = r + x;
    return r + (float)e;
}
float or_exp2(float xx) {
    if (xx != xx) return xx;
    if (xx > 127.0f) return INFINITY;
    if (xx < -127.0f) return 0.0f;
    float i0 = floorf(xx);
    float x = xx - i0;
    if (x > 0.5f) {
        i0 += 1.0f;
        x -= 1.0f;
    }
    float p = fmaf(1.535336188319500e-4f, x, 1.339887440266574e-3f);
    p = fmaf(p, x, 9.618437357674640e-3f);
    p = fmaf(p, x, 5.550332471162809e-2f);
    p = fmaf(p, x, 2.402264791363012e-1f);
    p = fmaf(p, x, 6.931472028550421e-1f);
    p = fmaf(p, x, 1.0f);
    return ldexpf(p, (int)i0);
}

/* ---- sin / cos: Cephes sinf/cosf (S. L. Moshier): octant reduction by three-part pi/4, minimax polynomials. */
void or_sincos(float x, float *s, float *c) {
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    const float FOPI = 1.27323954473516f;
    float sign_s = 1.0f, ax = x;
    if (x < 0.0f) {
        ax = -x;
        sign_s = -1.0f;
    }
    uint32_t j = (uint32_t)(ax * FOPI);
    float y = (float)j;
    if (j & 1u) {
        j += 1u;
        y += 1.0f;
    }
    j &= 7u;
    float sign_c = 1.0f;
    if (j > 3u) {
        j -= 4u;
        sign_s = -sign_s;
        sign_c = -sign_c;
    }
    if (j > 1u) sign_c = -sign_c;
    float z = fmaf(-y, DP1, ax);
    z = fmaf(-y, DP2, z);
    z = fmaf(-y, DP3, z);
    float zz = z * z;
    float ps = fmaf(fmaf(-1.9515295891e-4f, zz, 8.3321608736e-3f), zz, -1.6666654611e-1f);
    ps = fmaf(ps * zz, z, z);
    float pc = fmaf(fmaf(2.443315711809948e-5f, zz, -1.388731625493765e-3f), zz, 4.166664568298827e-2f);
    pc = fmaf(pc * zz, zz, fmaf(-0.5f, zz, 1.0f));
    float sv, cv;
    if (j == 1u || j == 2u) {
        sv = pc;
        cv = ps;
    } else {
        sv = ps;
        cv = pc;
    }
    *s = sign_s * sv;
    *c = sign_c * cv;
}

/* ---- vectors ---------------------------------------------------------------------------------------------- */
typedef struct {
    float x, y, z;
} V;
static inline V v3(float x, float y, float z) { V r = {x, y, z}; return r; }
static inline V vadd(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V vsub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V vmul(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V vdiv(V a, V b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline V vs(V a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V vds(V a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V vneg(V a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V cross(V a, V b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len(V a) { return sqrtf(dot(a, a)); }
static inline V normalize(V a) { return vds(a, sqrtf(dot(a, a))); }
static inline float fmaxs(float a, float b) { return fmaxf(a, b); }
static inline float fmins(float a, float b) { return fminf(a, b); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline int any_pos(V a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }
static inline int all_nonpos(V a) { return a.x <= 0.0f && a.y <= 0.0f && a.z <= 0.0f; }
static inline float luminance(V c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
static inline uint32_t fb(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline int non_finite(V v) {
    return (fb(v.x) & 0x7fffffffu) >= 0x7f800000u || (fb(v.y) & 0x7fffffffu) >= 0x7f800000u ||
           (fb(v.z) & 0x7fffffffu) >= 0x7f800000u;
}
static inline V ld3(const float *p) { return v3(p[0], p[1], p[2]); }

/* mat3 columns (t, b, n) */
typedef struct {
    V t, b, n;
} M3;
static inline V mmul(M3 m, V v) { return vadd(vadd(vs(m.t, v.x), vs(m.b, v.y)), vs(m.n, v.z)); }
static inline V mtmul(M3 m, V v) { return v3(dot(m.t, v), dot(m.b, v), dot(m.n, v)); }

/* ---- params (pathtrace.wgsl:96-140) ------------------------------------------------------------------------ */
typedef struct {
    const uint32_t *w;
} P_;
static inline float pf(const uint32_t *p, int vec, int k) {
    float f;
    memcpy(&f, &p[4 * vec + k], 4);
    return f;
}
static inline uint32_t pu(const uint32_t *p, int vec, int k) { return p[4 * vec + k]; }
enum { CAM_POS = 0, CAM_FWD = 1, CAM_RIGHT = 2, CAM_UP = 3, SIZE = 4, COUNTS = 5, ENV = 6, MISC = 7, SKY = 8,
       COUNTS2 = 9, LIGHT_LOD = 18, LTREE = 21, BACKGROUND = 24 };

enum { D_CAMERA = 0, D_BSDF = 1, D_RR = 2, D_SUN = 4, D_SUN_VIS = 5, D_LOCAL = 6, D_EMIT = 7, D_EMIT_VIS = 8, D_ENV = 9,
       D_ENV_VIS = 10 };
#define EMIT_RIS_M 6u
#define EMITTER_GUIDE 65536u
#define NO_EMITTER 0xffffffffu
#define INST_MIRRORED 4u
#define MAT_EMITTER_CARD 2
#define MAT_UNLIT 4
#define EMITTER_CARD 1u
#define MIRROR_ROUGHNESS 0.3f
#define OBJECT_STRIDE 3u

typedef struct {
    const or_render_scene *s;
    const uint32_t *P;
    or_rng g;
    int mirror_path;
    V mis_x, mis_n;
    int sampled_spec;
    int sampled_trans; /* the last BSDF sample went through the surface (pathtrace.wgsl:449) */
    int failed;
    int64_t *oob; /* table reads outside their table (each reads entry 0 instead) */
} Ctx;

/* Every table index is checked: one outside its table counts in *oob and reads entry 0. */
static inline uint64_t ck(const Ctx *c, uint64_t i, uint64_t n) {
    if (i < n) return i;
    (*c->oob)++;
    return 0;
}

/* ---- tables --------------------------------------------------------------------------------------------------- */
static inline const or_instance *inst_at(const Ctx *c, uint32_t i) {
    return &c->s->geo.instances[ck(c, i, c->s->geo.instance_count)];
}
static inline const float *material_at(const Ctx *c, uint32_t i) {
    return c->s->materials + 72 * ck(c, i, c->s->material_count);
}
static inline const float *light_at(const Ctx *c, uint32_t i) { return c->s->lights + 20 * ck(c, i, c->s->light_count); }
static inline const float *emitter_at(const Ctx *c, uint32_t i) {
    return c->s->emitters + 4 * ck(c, i, c->s->emitter_vec4s);
}
static inline uint32_t row_word(const Ctx *c, uint32_t first, uint32_t section, uint32_t k) {
    return c->s->rows[4 * ck(c, (uint64_t)first + section, c->s->row_count) + k];
}
static inline float obj_power(const Ctx *c, uint32_t obj) {
    return c->s->objects[4 * ck(c, obj, c->s->object_count) * OBJECT_STRIDE + 3];
}

typedef struct {
    float n0[4], n1[4], n2[4], uv[4], x[4];
} TriAttr;

typedef struct {
    TriAttr a;
    V v0, e1, e2;
    uint64_t vi[3]; /* its vertices (the record's first added) */
} TriData;

/* pathtrace.wgsl:330-360 tri_data (no particles, no deforming meshes). */
static TriData tri_data(const Ctx *c, uint32_t ti, uint32_t inst, uint32_t part) {
    uint32_t base = 0;
    if (inst != OR_NO_INST) base = inst_at(c, inst)->info[2];
    const uint32_t *ix = c->s->indices + 3 * ck(c, ti, c->s->tri_count);
    const float *p0 = c->s->vertices + 12 * ck(c, (uint64_t)ix[0] + base, c->s->vertex_count);
    const float *p1 = c->s->vertices + 12 * ck(c, (uint64_t)ix[1] + base, c->s->vertex_count);
    const float *p2 = c->s->vertices + 12 * ck(c, (uint64_t)ix[2] + base, c->s->vertex_count);
    const float *n0 = p0 + 4, *n1 = p1 + 4, *n2 = p2 + 4;
    TriData t;
    t.vi[0] = (uint64_t)ix[0] + base, t.vi[1] = (uint64_t)ix[1] + base, t.vi[2] = (uint64_t)ix[2] + base;
    t.v0 = ld3(p0);
    t.e1 = vsub(ld3(p1), ld3(p0));
    t.e2 = vsub(ld3(p2), ld3(p0));
    float a0[4] = {n0[0], n0[1], n0[2], p0[3]}, a1[4] = {n1[0], n1[1], n1[2], n0[3]},
          a2[4] = {n2[0], n2[1], n2[2], p1[3]};
    float fpart;
    memcpy(&fpart, &part, 4);
    float uv[4] = {n1[3], p2[3], n2[3], fpart}, x[4] = {0, 0, 0, 0};
    memcpy(t.a.n0, a0, 16);
    memcpy(t.a.n1, a1, 16);
    memcpy(t.a.n2, a2, 16);
    memcpy(t.a.uv, uv, 16);
    memcpy(t.a.x, x, 16);
    return t;
}

static inline V inst_vec(const or_instance *r, V v) {
    return v3(dot(ld3(r->m[0]), v), dot(ld3(r->m[1]), v), dot(ld3(r->m[2]), v));
}
static inline float inst_sign(const or_instance *r) { return (fb(r->n[1][3]) & INST_MIRRORED) ? -1.0f : 1.0f; }
static void inst_normal(const or_instance *r, float *n) {
    V v = ld3(n);
    V m = normalize(v3(dot(ld3(r->n[0]), v), dot(ld3(r->n[1]), v), dot(ld3(r->n[2]), v)));
    n[0] = m.x, n[1] = m.y, n[2] = m.z;
}
/* inst_attr: normals placed, the object from the record (inst_x). */
static void inst_attr(const or_instance *r, TriAttr *a) {
    inst_normal(r, a->n0);
    inst_normal(r, a->n1);
    inst_normal(r, a->n2);
    uint32_t obj = r->info[1] + fb(a->x[1]);
    memcpy(&a->x[1], &obj, 4);
}
static inline uint32_t object_of(const TriAttr *a) { return fb(a->x[1]); }
static inline uint32_t material_of(const Ctx *c, const TriAttr *a, uint32_t inst) {
    if (inst == OR_NO_INST) return 0;
    return row_word(c, inst_at(c, inst)->info[3], fb(a->uv[3]), 0) & 0xffffffu;
}

/* ---- textures (pathtrace.wgsl:1102-1140, texpack.rs) ---------------------------------------------------------- */
#define TEX_MIPS 5u
#define TEX_MAX_LEVEL 4.0f
#define LOD_BIAS 0.5f
static inline int tex_on(const Ctx *c) { return c->s->tex_info != NULL; }
static inline float fractf(float x) { return x - floorf(x); }
static inline int32_t as_i32(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i;
}
/* 1102-1104: each channel (c + 0.055) / 1.055 to the power 2.4, or c / 12.92 at or below 0.04045 */
static inline float srgb1(float c) {
    return c <= 0.04045f ? c / 12.92f : or_exp2(2.4f * or_log2((c + 0.055f) / 1.055f));
}
static inline V srgb_to_linear(V c) { return v3(srgb1(c.x), srgb1(c.y), srgb1(c.z)); }
/* vcolor (269-275): the triangle's three sRGB 8-bit corner colours (unpack4x8unorm), decoded and interpolated */
static V vcolor(const Ctx *c, uint32_t t, float u, float v) {
    V k[3];
    for (uint32_t j = 0; j < 3; j++) {
        const uint32_t w = c->s->vcolors[ck(c, (uint64_t)t * 3u + j, c->s->vcolor_words)];
        k[j] = srgb_to_linear(v3((float)(w & 0xffu) / 255.0f, (float)((w >> 8) & 0xffu) / 255.0f,
                                 (float)((w >> 16) & 0xffu) / 255.0f));
    }
    return vadd(vadd(vs(k[0], 1.0f - u - v), vs(k[1], u)), vs(k[2], v));
}
/* The pages' level `level` of layer `layer`, bilinear at page uv (pu, pv) with the reference's sampler (gpu.rs:2690:
 * repeat, linear): texel centres at (i + 0.5) / n, four texels weighed by the fractions, RGBA8 as unorm. */
static void tex_bilinear(const Ctx *c, uint32_t layer, uint32_t level, float pu, float pv, float *out) {
    const uint64_t side = c->s->tex_side;
    uint64_t layer_bytes = 0;
    for (uint32_t l = 0; l < TEX_MIPS; l++) layer_bytes += (side >> l) * (side >> l) * 4u;
    uint64_t off = (uint64_t)ck(c, layer, c->s->tex_layers) * layer_bytes;
    for (uint32_t l = 0; l < level; l++) off += (side >> l) * (side >> l) * 4u;
    const uint32_t n = (uint32_t)(side >> level);
    float x = pu * (float)n - 0.5f, y = pv * (float)n - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    float fx = x - x0f, fy = y - y0f;
    if (!(x0f >= -1048576.0f && x0f <= 1048576.0f)) x0f = 0.0f;
    if (!(y0f >= -1048576.0f && y0f <= 1048576.0f)) y0f = 0.0f;
    int32_t xi = (int32_t)x0f, yi = (int32_t)y0f, ni = (int32_t)n;
    uint32_t xa = (uint32_t)(((xi % ni) + ni) % ni), ya = (uint32_t)(((yi % ni) + ni) % ni);
    uint32_t xb = xa + 1u == n ? 0u : xa + 1u, yb = ya + 1u == n ? 0u : ya + 1u;
    const uint8_t *pg = c->s->tex_pages + off;
    const uint8_t *t00 = pg + ((uint64_t)ya * n + xa) * 4u, *t10 = pg + ((uint64_t)ya * n + xb) * 4u;
    const uint8_t *t01 = pg + ((uint64_t)yb * n + xa) * 4u, *t11 = pg + ((uint64_t)yb * n + xb) * 4u;
    const float k = 1.0f / 255.0f;
    for (int j = 0; j < 4; j++)
        out[j] = mixf(mixf((float)t00[j] * k, (float)t10[j] * k, fx), mixf((float)t01[j] * k, (float)t11[j] * k, fx), fy);
}
/* tex_level (1107-1112): the texture's rect in its page (wrapped or clamped uv), trilinear between the two levels
 * about `level` (clamped to the pages' 0..4) */
static void tex_level(const Ctx *c, int32_t id, float u, float v, float level, float *out) {
    const float *a = c->s->tex_info + 4 * ck(c, (uint64_t)(uint32_t)id * 2u, c->s->tex_info_vec4s);
    const float *b = c->s->tex_info + 4 * ck(c, (uint64_t)(uint32_t)id * 2u + 1u, c->s->tex_info_vec4s);
    float fu, fv;
    if (b[1] > 0.5f) {
        fu = fractf(u), fv = fractf(v);
    } else {
        fu = clampf(u, 0.0f, 1.0f), fv = clampf(v, 0.0f, 1.0f);
    }
    const float pu = a[0] + fu * a[2], pv = a[1] + fv * a[3];
    const uint32_t layer = (uint32_t)(int32_t)b[0];
    const float lv = clampf(level, 0.0f, TEX_MAX_LEVEL);
    const uint32_t l0 = (uint32_t)lv;
    const float fl = lv - (float)l0;
    tex_bilinear(c, layer, l0, pu, pv, out);
    if (fl > 0.0f) {
        float o1[4];
        tex_bilinear(c, layer, l0 + 1u, pu, pv, o1);
        for (int j = 0; j < 4; j++) out[j] = mixf(out[j], o1[j], fl);
    }
}
/* tex (1114-1117): `lod` is log2 of the footprint in texture-UV units; the level adds the texture's own size */
static void tex(const Ctx *c, int32_t id, float u, float v, float lod, float *out) {
    const float *b = c->s->tex_info + 4 * ck(c, (uint64_t)(uint32_t)id * 2u + 1u, c->s->tex_info_vec4s);
    tex_level(c, id, u, v, lod + 0.5f * or_log2(b[2] * b[3]), out);
}
/* mesh_uv (1119-1122): the hit's uv x the material's repeat + its offset */
static void mesh_uv(const TriAttr *a, const float *m, float u, float v, float *uv) {
    const float w = 1.0f - u - v;
    uv[0] = (a->n0[3] * w + a->n2[3] * u + a->uv[1] * v) * m[20] + m[24];
    uv[1] = (a->n1[3] * w + a->uv[0] * u + a->uv[2] * v) * m[21] + m[25];
}
/* mesh_density (1127-1137): (UV area x repeat, area) of the triangle as it stands; (repeat, 1) where either has none */
static void mesh_density(V e1, V e2, const TriAttr *a, const float *m, float *den) {
    const float d1x = a->n2[3] - a->n0[3], d1y = a->uv[0] - a->n1[3];
    const float d2x = a->uv[1] - a->n0[3], d2y = a->uv[2] - a->n1[3];
    const float uv_area = fabsf(d1x * d2y - d2x * d1y);
    const float area = len(cross(e1, e2));
    const float rep = fmaxf(fabsf(m[20] * m[21]), 1e-12f);
    const int sized = area > 2e-12f && uv_area > 2e-14f;
    den[0] = sized ? uv_area * rep : rep;
    den[1] = sized ? area : 1.0f;
}
/* mesh_lod (1140) */
static inline float mesh_lod(const float *den, float fp2, float over) {
    return 0.5f * or_log2(fmaxf(den[0] * fp2 / (den[1] * over), 1e-30f));
}
/* emission (1312-1317): the factor x the emissive map */
static V emission(const Ctx *c, const float *m, const float *uv, float lod) {
    V em = ld3(m + 4);
    if (all_nonpos(em)) return em;
    if (tex_on(c) && as_i32(m[16]) >= 0) {
        float tx[4];
        tex(c, as_i32(m[16]), uv[0], uv[1], lod, tx);
        em = vmul(em, srgb_to_linear(v3(tx[0], tx[1], tx[2])));
    }
    return em;
}
static inline float lod_height(const Ctx *c) {
    return (float)(pu(c->P, COUNTS2, 2) > 0u ? pu(c->P, COUNTS2, 2) : pu(c->P, SIZE, 1));
}

/* ---- surface -------------------------------------------------------------------------------------------------- */
typedef struct {
    V p, n, ng, base, emissive;
    float roughness, metallic, specular;
    float coat, coat_rough; /* pathtrace.wgsl:1420-1423 (F_COAT: on when any material has one; 0 otherwise) */
    float transmission, ior; /* pathtrace.wgsl:1413-1419 (F_GLASS: a material has glass) */
    int front, card;
} Surface;

/* hit_tangent (366-394): the vertices' tangents (hand of the first vertex; 0: none, then one from the uv edges, or
 * any vector at right angles to the normal where the uvs have no area), placed by the record */
static inline V unit_or_zero(V v) {
    float dd = dot(v, v);
    if (dd > 0.0f && isfinite(dd)) return normalize(v);
    return v3(0, 0, 0);
}
static void hit_tangent(const Ctx *c, const TriData *td, uint32_t inst, float u, float v, float *tw) {
    const float *t0 = c->s->vertices + 12 * ck(c, td->vi[0], c->s->vertex_count) + 8;
    const float *t1 = c->s->vertices + 12 * ck(c, td->vi[1], c->s->vertex_count) + 8;
    const float *t2 = c->s->vertices + 12 * ck(c, td->vi[2], c->s->vertex_count) + 8;
    V t = vadd(vadd(vs(ld3(t0), 1.0f - u - v), vs(ld3(t1), u)), vs(ld3(t2), v));
    float hand = t0[3];
    if (t0[3] == 0.0f) {
        const TriAttr *a = &td->a;
        const float d1x = a->n2[3] - a->n0[3], d1y = a->uv[0] - a->n1[3];
        const float d2x = a->uv[1] - a->n0[3], d2y = a->uv[2] - a->n1[3];
        const float det = d1x * d2y - d2x * d1y;
        if (fabsf(det) < 1e-12f) {
            V any = v3(0.0f, 1.0f, 0.0f);
            if (fabsf(a->n0[0]) < 0.9f) any = v3(1.0f, 0.0f, 0.0f);
            t = normalize(cross(ld3(a->n0), any));
            hand = 1.0f;
        } else {
            const float r = 1.0f / det;
            const V tu = vs(vsub(vs(td->e1, d2y), vs(td->e2, d1y)), r);
            const float l = len(tu);
            V tn = v3(0, 0, 0);
            if (l > 0.0f && isfinite(l)) tn = vds(tu, l);
            const V bt = vs(vsub(vs(td->e2, d1x), vs(td->e1, d2x)), r);
            hand = dot(cross(ld3(a->n0), tn), bt) < 0.0f ? 1.0f : -1.0f;
            t = tn;
        }
    }
    if (inst != OR_NO_INST) {
        const or_instance *r = inst_at(c, inst);
        t = unit_or_zero(inst_vec(r, t));
        hand = hand * inst_sign(r);
    }
    tw[0] = t.x, tw[1] = t.y, tw[2] = t.z, tw[3] = hand;
}

/* surface (1320-1425): `width` is the ray cone's footprint (m) at the hit, for the textures' level */
static Surface surface(Ctx *c, V o, V d, const or_hit *h, float width) {
    TriData td = tri_data(c, h->tri, h->inst, h->part);
    V e1 = td.e1, e2 = td.e2;
    TriAttr a = td.a;
    float side = 1.0f;
    if (h->inst != OR_NO_INST) {
        const or_instance *r = inst_at(c, h->inst);
        e1 = inst_vec(r, e1);
        e2 = inst_vec(r, e2);
        inst_attr(r, &a);
        side = inst_sign(r);
    }
    float w = 1.0f - h->u - h->v;
    V ng = vs(normalize(cross(e1, e2)), side);
    V n = normalize(vadd(vadd(vs(ld3(a.n0), w), vs(ld3(a.n1), h->u)), vs(ld3(a.n2), h->v)));
    const float *m = material_at(c, material_of(c, &a, h->inst));
    int front = dot(ng, d) < 0.0f;
    if (!front) ng = vneg(ng);
    if (dot(n, ng) < 0.0f) n = vneg(n);
    Surface s;
    s.p = vadd(o, vs(d, h->t));
    s.front = front;
    float uv[2] = {0.0f, 0.0f}, lod = 0.0f;
    if (tex_on(c)) { /* 1344-1348 */
        float den[2];
        mesh_uv(&a, m, h->u, h->v, uv);
        mesh_density(e1, e2, &a, m, den);
        const float fp = fmaxf(width, 1e-7f);
        lod = mesh_lod(den, fp * fp, fmaxf(fabsf(dot(ng, d)), 1e-3f)) - LOD_BIAS;
    }
    s.card = as_i32(m[4 * 4 + 2]) == MAT_EMITTER_CARD;
    V base = ld3(m);
    if (c->s->vcolors && m[23] > 0.0f) base = vmul(base, vcolor(c, h->tri, h->u, h->v)); /* 1356 */
    float rough = m[8 + 1], metal = m[8 + 0], spec = m[8 + 2];
    if (tex_on(c)) { /* 1360-1364 */
        float tx[4];
        if (as_i32(m[12]) >= 0) {
            tex(c, as_i32(m[12]), uv[0], uv[1], lod, tx);
            base = vmul(base, srgb_to_linear(v3(tx[0], tx[1], tx[2])));
        }
        if (as_i32(m[14]) >= 0) {
            tex(c, as_i32(m[14]), uv[0], uv[1], lod, tx);
            rough *= tx[1];
        }
        if (as_i32(m[15]) >= 0) {
            tex(c, as_i32(m[15]), uv[0], uv[1], lod, tx);
            metal *= tx[2];
        }
    }
    s.emissive = vs(emission(c, m, uv, lod), obj_power(c, object_of(&a)));
    /* 1366-1372: three.js MeshBasicMaterial (tex1.z MAT_UNLIT): its colour to camera rays and mirror-like paths, lit
     * by nothing, lighting nothing */
    if (as_i32(m[18]) == MAT_UNLIT) {
        s.emissive = (c->g.depth == 0u || c->mirror_path) ? base : v3(0, 0, 0);
        base = v3(0, 0, 0);
        spec = 0.0f;
        metal = 0.0f;
    }
    /* 1376-1387: the normal map where the surface is seen directly or in a mirror-like reflection */
    const int near = c->g.depth == 0u || c->mirror_path;
    if (tex_on(c) && as_i32(m[13]) >= 0 && near) {
        float tx[4], tw[4];
        tex(c, as_i32(m[13]), uv[0], uv[1], lod, tx);
        const V tn = v3(tx[0] * 2.0f - 1.0f, tx[1] * 2.0f - 1.0f, tx[2] * 2.0f - 1.0f);
        hit_tangent(c, &td, h->inst, h->u, h->v, tw);
        const V twv = ld3(tw);
        const V tp = vsub(twv, vs(n, dot(n, twv)));
        if (dot(tp, tp) > 1e-12f) {
            const V tt = normalize(tp);
            const V b = vs(cross(n, tt), tw[3]);
            const V pn = vadd(vadd(vs(tt, tn.x * m[11]), vs(b, tn.y * m[11])), vs(n, tn.z));
            if (dot(pn, pn) > 1e-12f && dot(pn, ng) > 0.0f) n = normalize(pn);
        }
    }
    s.base = base;
    s.roughness = clampf(rough, 0.02f, 1.0f);
    s.metallic = clampf(metal, 0.0f, 1.0f);
    s.specular = spec;
    s.ng = ng;
    s.n = n;
    s.coat = clampf(m[28 + 2], 0.0f, 1.0f); /* layer.z, layer.w */
    s.coat_rough = clampf(m[28 + 3], 0.02f, 1.0f);
    s.ior = 1.5f;
    s.transmission = 0.0f;
    if (c->s->glass_energy) { /* layer.x, layer.y; the transmission map (uv.z: its texture + 1) */
        s.transmission = clampf(m[28 + 0], 0.0f, 1.0f);
        if (tex_on(c) && m[22] > 0.5f) {
            float tx[4];
            tex(c, (int32_t)m[22] - 1, uv[0], uv[1], lod, tx);
            s.transmission *= tx[0];
        }
        s.ior = fmaxf(m[28 + 1], 1.0001f);
    }
    return s;
}

/* ---- BSDF ----------------------------------------------------------------------------------------------------- */
typedef struct {
    V diffuse, f0, f90, w, e_spec, kd;
    float a2, p_spec, p_coat, p_trans;
    float coat, coat_alpha, e_coat;
    /* the dielectric (pathtrace.wgsl:1535-1538): its share x compensation, tint, n(far) / n(view), rough or a delta
     * transmission, its directional reflection albedo */
    V tint;
    float t, eta, e_glass;
    int rough_t;
} Lobes;

/* ---- the dielectric (pathtrace.wgsl:1542-1640) ---------------------------------------------------------------- */
#define GLASS_DELTA 0.05f
#define GLASS_E_MU 16u
#define GLASS_E_ROUGH 16u
#define GLASS_E_IOR 8u
#define GLASS_E_ZMAX 0.70710678f
static inline int glass_on(const Ctx *c) { return c->s->glass_energy != NULL; }
/* 1559-1566 */
static float fresnel_dielectric(float c, float eta) {
    float g2 = eta * eta - 1.0f + c * c;
    if (g2 <= 0.0f) return 1.0f;
    float g = sqrtf(g2);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * a * a * (1.0f + b * b);
}
/* 1569-1572 */
static inline float lambda_ggx(float a2, float c) {
    float c2 = c * c;
    return 0.5f * (sqrtf(1.0f + a2 * (1.0f - c2) / fmaxf(c2, 1e-8f)) - 1.0f);
}
/* 1582-1603 */
static float glass_e_at(const Ctx *c, uint32_t i) {
    const float *v = c->s->glass_energy + 4 * ck(c, i / 4u, 1024);
    return v[i & 3u];
}
static float glass_e_row(const Ctx *c, uint32_t row, float x) {
    uint32_t x0 = (uint32_t)x;
    uint32_t x1 = x0 + 1u < GLASS_E_MU - 1u ? x0 + 1u : GLASS_E_MU - 1u;
    return mixf(glass_e_at(c, row * GLASS_E_MU + x0), glass_e_at(c, row * GLASS_E_MU + x1), x - (float)x0);
}
static float glass_energy(const Ctx *c, float nov, float roughness, float ior, int leaving) {
    float x = clampf(sqrtf(nov), 0.0f, 1.0f) * (float)(GLASS_E_MU - 1u);
    float y = clampf(roughness, 0.0f, 1.0f) * (float)(GLASS_E_ROUGH - 1u);
    float z = clampf(sqrtf(fmaxf(ior - 1.0f, 0.0f) / (ior + 1.0f)) / GLASS_E_ZMAX, 0.0f, 1.0f) * (float)(GLASS_E_IOR - 1u);
    uint32_t y0 = (uint32_t)y;
    uint32_t y1 = y0 + 1u < GLASS_E_ROUGH - 1u ? y0 + 1u : GLASS_E_ROUGH - 1u;
    uint32_t side = leaving ? GLASS_E_IOR : 0u, zi = (uint32_t)z;
    uint32_t z0 = side + zi;
    uint32_t z1 = side + (zi + 1u < GLASS_E_IOR - 1u ? zi + 1u : GLASS_E_IOR - 1u);
    float e0 = mixf(glass_e_row(c, z0 * GLASS_E_ROUGH + y0, x), glass_e_row(c, z0 * GLASS_E_ROUGH + y1, x), y - (float)y0);
    float e1 = mixf(glass_e_row(c, z1 * GLASS_E_ROUGH + y0, x), glass_e_row(c, z1 * GLASS_E_ROUGH + y1, x), y - (float)y0);
    return mixf(e0, e1, z - floorf(z));
}
/* 1606 */
static inline int two_sided(const Ctx *c, const Lobes *lb) { return glass_on(c) && lb->rough_t; }

static void energy_node(const Ctx *c, uint32_t i, float *out) {
    const float *v = c->s->energy + 4 * ck(c, i >> 1, 512);
    out[0] = (i & 1u) ? v[2] : v[0];
    out[1] = (i & 1u) ? v[3] : v[1];
}
static void energy_lookup(const Ctx *c, float roughness, float nov, float *e) {
    const uint32_t N = 32;
    float x = sqrtf(clampf(nov, 0.0f, 1.0f)) * (float)(N - 1);
    float y = clampf(roughness, 0.0f, 1.0f) * (float)(N - 1);
    uint32_t x0 = (uint32_t)x < N - 2 ? (uint32_t)x : N - 2;
    uint32_t y0 = (uint32_t)y < N - 2 ? (uint32_t)y : N - 2;
    uint32_t i = y0 * N + x0;
    float fx = x - (float)x0, fy = y - (float)y0;
    float a[2], b[2], cc[2], d[2];
    energy_node(c, i, a);
    energy_node(c, i + 1, b);
    energy_node(c, i + N, cc);
    energy_node(c, i + N + 1, d);
    for (int k = 0; k < 2; k++) e[k] = mixf(mixf(a[k], b[k], fx), mixf(cc[k], d[k], fx), fy);
}

static Lobes lobes(const Ctx *c, const Surface *s, float nov) {
    Lobes l;
    float t = s->transmission * (1.0f - s->metallic);
    l.diffuse = vs(vs(s->base, 1.0f - s->metallic), 1.0f - t);
    float f0 = 0.08f * s->specular;
    if (t > 0.0f) { /* glass: F0 from the index of refraction (pathtrace.wgsl:1654-1658) */
        float ri = (s->ior - 1.0f) / (s->ior + 1.0f);
        f0 = ri * ri;
    }
    V f0v = v3(f0, f0, f0);
    l.f0 = v3(mixf(f0v.x, s->base.x, s->metallic), mixf(f0v.y, s->base.y, s->metallic),
              mixf(f0v.z, s->base.z, s->metallic));
    float e[2];
    energy_lookup(c, s->roughness, nov, e);
    float comp = pf(c->P, MISC, 2);
    const float k503 = 50.0f / 3.0f;
    float f90 = clampf(dot(l.f0, v3(k503, k503, k503)), 0.0f, 1.0f);
    float inv = comp * (1.0f / fmaxf(e[0] + e[1], 1e-3f) - 1.0f);
    l.w = v3(1.0f + l.f0.x * inv, 1.0f + l.f0.y * inv, 1.0f + l.f0.z * inv);
    l.e_spec = vmul(vadd(vs(l.f0, e[0]), v3(f90 * e[1], f90 * e[1], f90 * e[1])), l.w);
    l.f90 = v3(f90, f90, f90);
    l.kd = vsub(v3(1, 1, 1), l.e_spec);
    if (glass_on(c) && t > 0.0f) { /* 1669-1673: the share that is not the dielectric keeps the specular lobe */
        l.w = vs(l.w, 1.0f - t);
        l.e_spec = vs(l.e_spec, 1.0f - t);
    }
    float r = s->roughness;
    l.a2 = r * r * r * r;
    V lw = v3(0.3f, 0.6f, 0.1f);
    float ws = dot(l.e_spec, lw);
    float wd = dot(vmul(l.diffuse, l.kd), lw);
    l.p_spec = clampf(ws / fmaxf(ws + wd, 1e-6f), 0.05f, 1.0f);
    if (wd <= 0.0f) l.p_spec = 1.0f;
    /* pathtrace.wgsl:1675-1683: the coat's directional albedo, a single-scattering lobe of F0 0.04, F90 1 */
    l.coat = 0.0f;
    l.coat_alpha = 1.0f;
    l.e_coat = 0.0f;
    if (s->coat > 0.0f) {
        float ec[2];
        l.coat = s->coat;
        energy_lookup(c, s->coat_rough, nov, ec);
        l.e_coat = s->coat * (0.04f * ec[0] + ec[1]);
        l.coat_alpha = s->coat_rough * s->coat_rough;
    }
    float e_coat = l.e_coat, wt = 0.0f, under = 1.0f - e_coat;
    l.t = 0.0f;
    l.eta = 0.0f;
    l.rough_t = 0;
    l.tint = v3(0, 0, 0);
    l.e_glass = 0.0f;
    if (glass_on(c) && t > 0.0f) { /* 1688-1697 */
        l.eta = s->front ? s->ior : 1.0f / s->ior;
        l.rough_t = s->roughness > GLASS_DELTA && fabsf(s->ior - 1.0f) > 1e-3f;
        float gc = 1.0f;
        if (l.rough_t) gc = mixf(1.0f, 1.0f / glass_energy(c, nov, s->roughness, s->ior, !s->front), comp);
        l.t = t * under * gc;
        l.tint = vs(s->base, l.t);
        l.e_glass = t * fresnel_dielectric(nov, l.eta);
        wt = under * (l.e_glass + (t - l.e_glass) * dot(s->base, lw));
    }
    float total = (ws + wd) * under + e_coat + wt;
    l.p_coat = e_coat / fmaxf(total, 1e-6f);
    l.p_trans = wt / fmaxf(total, 1e-6f);
    return l;
}

static inline float d_ggx(float a2, float noh) {
    float d = (noh * a2 - noh) * noh + 1.0f;
    return a2 / (PI_F * d * d);
}
static inline float vis_smith(float a2, float nov, float nol) {
    return 0.5f / fmaxf(mixf(2.0f * nol * nov, nol + nov, sqrtf(a2)), 1e-7f);
}
static inline V f_schlick(V f0, V f90, float voh) {
    float c = 1.0f - voh, c2 = c * c, k = c2 * c2 * c;
    return vadd(f0, vs(vsub(f90, f0), k));
}

static inline float g1(float a2, float nov) { return 2.0f * nov / (nov + sqrtf(a2 + (1.0f - a2) * nov * nov)); }
/* The dielectric's transmission toward l (the far side) x |n.l|, Walter 2007 eq. 21 (pathtrace.wgsl:1610-1622) */
static V dielectric_t(const Surface *s, const Lobes *lb, V v, V l, float nov, float nol) {
    if (!lb->rough_t) return v3(0, 0, 0);
    V h = normalize(vadd(v, vs(l, lb->eta)));
    if (dot(h, s->n) < 0.0f) h = vneg(h);
    float voh = dot(v, h), loh = dot(l, h);
    if (voh <= 0.0f || loh >= 0.0f) return v3(0, 0, 0);
    float den = voh + lb->eta * loh;
    float g2 = 1.0f / (1.0f + lambda_ggx(lb->a2, nov) + lambda_ggx(lb->a2, nol));
    float f = (1.0f - fresnel_dielectric(voh, lb->eta)) * d_ggx(lb->a2, fmaxf(dot(s->n, h), 0.0f)) * g2 * voh * lb->eta *
              lb->eta * -loh / (nov * fmaxf(den * den, 1e-12f));
    return vs(lb->tint, f);
}
/* The density of l under the dielectric's sampling x its share of the samples (1626-1640) */
static float pdf_dielectric(const Surface *s, const Lobes *lb, V v, V l) {
    float nov = fmaxf(dot(s->n, v), 1e-4f);
    if (dot(s->n, l) > 0.0f) {
        V h = normalize(vadd(v, l));
        float voh = fmaxf(dot(v, h), 1e-6f);
        float f = fresnel_dielectric(lb->rough_t ? voh : nov, lb->eta);
        return lb->p_trans * f * g1(lb->a2, nov) * d_ggx(lb->a2, fmaxf(dot(s->n, h), 0.0f)) / (4.0f * nov);
    }
    if (!lb->rough_t) return 0.0f;
    V h = normalize(vadd(v, vs(l, lb->eta)));
    if (dot(h, s->n) < 0.0f) h = vneg(h);
    float voh = dot(v, h), loh = dot(l, h);
    if (voh <= 0.0f || loh >= 0.0f) return 0.0f;
    float den = voh + lb->eta * loh;
    return lb->p_trans * (1.0f - fresnel_dielectric(voh, lb->eta)) * g1(lb->a2, nov) * voh *
           d_ggx(lb->a2, fmaxf(dot(s->n, h), 0.0f)) * lb->eta * lb->eta * -loh / (nov * fmaxf(den * den, 1e-12f));
}

static V eval_bsdf_k(const Ctx *c, const Surface *s, const Lobes *lb, V v, V l, float k) {
    float nol = dot(s->n, l);
    float nov = fmaxf(dot(s->n, v), 1e-4f);
    if (two_sided(c, lb) && nol < 0.0f && dot(s->ng, l) < 0.0f) return dielectric_t(s, lb, v, l, nov, -nol);
    if (nol <= 0.0f || dot(s->ng, l) <= 0.0f) return v3(0, 0, 0);
    V h = normalize(vadd(v, l));
    float noh = fmaxf(dot(s->n, h), 0.0f);
    float voh = fmaxf(dot(v, h), 0.0f);
    float dv = d_ggx(lb->a2, noh) * vis_smith(lb->a2, nov, nol);
    V spec = vs(vmul(vs(f_schlick(lb->f0, lb->f90, voh), dv), lb->w), k);
    /* the dielectric's reflection (1747-1748): lb->t holds the coat's share already, so it joins after the coat */
    float glass = 0.0f;
    if (glass_on(c) && lb->t > 0.0f) glass = lb->t * fresnel_dielectric(voh, lb->eta) * dv * k;
    V diff = vmul(vs(lb->diffuse, INV_PI), lb->kd);
    if (lb->coat > 0.0f) { /* pathtrace.wgsl:1750-1755 */
        float ca2 = lb->coat_alpha * lb->coat_alpha;
        float cf = f_schlick(v3(0.04f, 0.04f, 0.04f), v3(1, 1, 1), voh).x;
        float coat = d_ggx(ca2, noh) * vis_smith(ca2, nov, nol) * cf * lb->coat * k;
        diff = vs(diff, 1.0f - lb->e_coat);
        spec = vadd(vs(spec, 1.0f - lb->e_coat), v3(coat, coat, coat));
    }
    spec = vadd(spec, v3(glass, glass, glass));
    return vs(vadd(diff, spec), nol);
}
static V eval_bsdf(const Ctx *c, const Surface *s, const Lobes *lb, V v, V l) { return eval_bsdf_k(c, s, lb, v, l, 1.0f); }


static float pdf_bsdf(const Ctx *c, const Surface *s, const Lobes *lb, V v, V l) {
    float nol = dot(s->n, l);
    float pg = 0.0f;
    if (glass_on(c) && lb->p_trans > 0.0f) pg = pdf_dielectric(s, lb, v, l);
    if (nol <= 0.0f) return pg;
    float nov = fmaxf(dot(s->n, v), 1e-4f);
    V h = normalize(vadd(v, l));
    float noh = fmaxf(dot(s->n, h), 0.0f);
    float voh = fmaxf(dot(v, h), 1e-6f);
    float dv = g1(lb->a2, nov) * voh * d_ggx(lb->a2, noh) / nov;
    float ps = dv / (4.0f * voh);
    float pd = nol / PI_F;
    float pc = 0.0f;
    if (lb->coat > 0.0f) { /* pathtrace.wgsl:1783-1786 */
        float ca2 = lb->coat_alpha * lb->coat_alpha;
        pc = g1(ca2, nov) * d_ggx(ca2, noh) / (4.0f * nov);
    }
    return pg + lb->p_coat * pc + (1.0f - lb->p_coat - lb->p_trans) * (lb->p_spec * ps + (1.0f - lb->p_spec) * pd);
}

static inline float power_heuristic(float a, float b) {
    float a2 = a * a;
    return a2 / fmaxf(a2 + b * b, 1e-20f);
}
static float light_mis(const Ctx *c, float pdf_l, const Surface *s, const Lobes *lb, V v, V l) {
    return power_heuristic(pdf_l, pdf_bsdf(c, s, lb, v, l));
}

static M3 basis(V n) {
    float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sgn + n.z);
    float b = n.x * n.y * a;
    M3 m;
    m.t = v3(1.0f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x);
    m.b = v3(b, sgn + n.y * n.y * a, -n.y);
    m.n = n;
    return m;
}

static V sample_vndf(V v, M3 tbn, float alpha, float u1, float u2) {
    V vl = mtmul(tbn, v);
    V vh = normalize(v3(alpha * vl.x, alpha * vl.y, vl.z));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    V t1 = lensq > 0.0f ? vds(v3(-vh.y, vh.x, 0.0f), sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    V t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = TWO_PI * u2;
    float sp, cp;
    or_sincos(phi, &sp, &cp);
    float p1 = r * cp;
    float p2 = r * sp;
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * sqrtf(1.0f - p1 * p1) + s * p2;
    V nh = vadd(vadd(vs(t1, p1), vs(t2, p2)), vs(vh, sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2))));
    V hl = normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)));
    return mmul(tbn, hl);
}
static inline V reflect(V e1, V e2) { return vsub(e1, vs(e2, 2.0f * dot(e2, e1))); }
/* WGSL refract: the zero vector at total internal reflection */
static inline V refract(V e1, V e2, float eta) {
    float d = dot(e2, e1);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3(0, 0, 0);
    return vsub(vs(e1, eta), vs(e2, eta * d + sqrtf(k)));
}

static V sample_bsdf(Ctx *c, const Surface *s, const Lobes *lb, V v) {
    M3 tbn = basis(s->n);
    float r[4];
    or_rng4(&c->g, or_rng_dim(&c->g, D_BSDF), r);
    c->sampled_trans = 0;
    if (r[0] < lb->p_coat) { /* pathtrace.wgsl:1853-1857: the coat's lobe (p_coat is 0 without one) */
        c->sampled_spec = 1;
        return reflect(vneg(v), sample_vndf(v, tbn, lb->coat_alpha, r[2], r[3]));
    }
    if (r[0] < lb->p_coat + lb->p_trans) {
        /* 1858-1874, the dielectric: a visible microfacet normal (smooth glass refracts through the shading normal),
         * reflected or refracted by the Fresnel reflectance there; a reflection under the surface or a refraction
         * above it is no sample */
        V h = sample_vndf(v, tbn, s->roughness * s->roughness, r[2], r[3]);
        V m = lb->rough_t ? h : s->n;
        V tr = refract(vneg(v), m, 1.0f / lb->eta);
        c->sampled_spec = r[0] - lb->p_coat < fresnel_dielectric(fmaxf(dot(v, m), 1e-4f), lb->eta) * lb->p_trans ||
                          (tr.x == 0.0f && tr.y == 0.0f && tr.z == 0.0f);
        if (c->sampled_spec) {
            V rl = reflect(vneg(v), h);
            return dot(rl, s->n) <= 0.0f ? v3(0, 0, 0) : rl;
        }
        c->sampled_trans = 1;
        return dot(tr, s->n) >= 0.0f ? v3(0, 0, 0) : tr;
    }
    c->sampled_spec = r[1] < lb->p_spec;
    if (c->sampled_spec) {
        V h = sample_vndf(v, tbn, s->roughness * s->roughness, r[2], r[3]);
        return reflect(vneg(v), h);
    }
    float rr = sqrtf(r[2]);
    float phi = TWO_PI * r[3];
    float sp, cp;
    or_sincos(phi, &sp, &cp);
    return mmul(tbn, v3(rr * cp, rr * sp, sqrtf(fmaxf(0.0f, 1.0f - rr * rr))));
}

/* ---- light ---------------------------------------------------------------------------------------------------- */
static inline V offset_origin(V p, V n) {
    float m = fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
    return vadd(p, vs(n, 1e-4f * m));
}

static int occluded(Ctx *c, V o, V d, float tmax, uint32_t mode) {
    or_hit h;
    float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    if (or_trace(&c->s->geo, of, df, tmax, mode, OR_WALK_EXACT, 0, &h, NULL)) c->failed++;
    return h.tri != OR_NO_TRI;
}

/* ---- glass along shadow rays (pathtrace.wgsl:846-977) ------------------------------------------------------- */
typedef struct {
    V ng, tint;
    float share;
} GlassAt;
/* the share of a ray a glass material lets through at |cos| cs (glass_at, pathtrace.wgsl:955-969) */
static float glass_share(const Ctx *c, const float *m, float cs, float tr_map) {
    float tr = clampf(m[28 + 0], 0.0f, 1.0f) * (1.0f - clampf(m[8 + 0], 0.0f, 1.0f));
    if (tr <= 0.0f) return 0.0f;
    tr = tr * tr_map; /* the transmission map's red, 1 without one */
    float ior = fmaxf(m[28 + 1], 1.0001f);
    float ri = (ior - 1.0f) / (ior + 1.0f);
    float f0 = ri * ri;
    float e[2];
    energy_lookup(c, clampf(m[8 + 1], 0.02f, 1.0f), cs, e);
    /* spec_terms (1521-1525) at this F0: the red channel of its directional albedo */
    const float k503 = 50.0f / 3.0f;
    float f90 = clampf(dot(v3(f0, f0, f0), v3(k503, k503, k503)), 0.0f, 1.0f);
    float w = 1.0f + f0 * (pf(c->P, MISC, 2) * (1.0f / fmaxf(e[0] + e[1], 1e-3f) - 1.0f));
    return tr * (1.0f - (f0 * e[0] + f90 * e[1]) * w);
}
/* glass_at (930-970): a hit's geometric normal, the share of a ray its transmission lets through at that angle
 * (transmission x (1 - the specular lobe's directional albedo there); 0: not glass) and its tint; `t` the distance
 * from the shadow ray's start (the maps' level, as a camera ray's footprint there) */
static GlassAt glass_at(const Ctx *c, const or_hit *h, V d, float t) {
    TriData td = tri_data(c, h->tri, h->inst, h->part);
    V e1 = td.e1, e2 = td.e2;
    TriAttr a = td.a;
    float side = 1.0f;
    if (h->inst != OR_NO_INST) {
        const or_instance *r = inst_at(c, h->inst);
        e1 = inst_vec(r, e1);
        e2 = inst_vec(r, e2);
        inst_attr(r, &a);
        side = inst_sign(r);
    }
    const float *m = material_at(c, material_of(c, &a, h->inst));
    GlassAt g;
    g.ng = vs(normalize(cross(e1, e2)), side);
    g.share = 0.0f;
    g.tint = v3(1, 1, 1);
    if (clampf(m[28 + 0], 0.0f, 1.0f) * (1.0f - clampf(m[8 + 0], 0.0f, 1.0f)) <= 0.0f) return g;
    g.tint = ld3(m);
    if (c->s->vcolors && m[23] > 0.0f) g.tint = vmul(g.tint, vcolor(c, h->tri, h->u, h->v)); /* 949 */
    float tr_map = 1.0f;
    if (tex_on(c) && (as_i32(m[12]) >= 0 || m[22] > 0.5f)) { /* 950-956 */
        float uv[2], den[2], tx[4];
        mesh_uv(&a, m, h->u, h->v, uv);
        const float fp = fmaxf(t * 2.0f * pf(c->P, CAM_POS, 3) / lod_height(c), 1e-7f);
        mesh_density(e1, e2, &a, m, den);
        const float lod = mesh_lod(den, fp * fp, 1.0f) - LOD_BIAS;
        if (as_i32(m[12]) >= 0) {
            tex(c, as_i32(m[12]), uv[0], uv[1], lod, tx);
            g.tint = vmul(g.tint, srgb_to_linear(v3(tx[0], tx[1], tx[2])));
        }
        if (m[22] > 0.5f) {
            tex(c, (int32_t)m[22] - 1, uv[0], uv[1], lod, tx);
            tr_map = tx[0];
        }
    }
    g.share = glass_share(c, m, fabsf(dot(g.ng, d)), tr_map);
    return g;
}
/* glass_tr (885-906): past every opaque caster, what the glass a shadow ray crosses lets through, the ray kept
 * straight; toward the sky an odd number of glass surfaces lets nothing through */
static V glass_tr(Ctx *c, V o, V d, float tmax, int sky) {
    float left = tmax;
    V tr = v3(1, 1, 1);
    int odd = 0;
    float t0 = 0.0f;
    for (uint32_t k = 0; k < 8u; k++) {
        or_hit h;
        float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
        if (or_trace(&c->s->geo, of, df, left, OR_TRACE_CLOSEST, OR_WALK_EXACT, 0, &h, NULL)) c->failed++;
        if (h.tri == OR_NO_TRI) break;
        t0 += h.t;
        GlassAt g = glass_at(c, &h, d, t0);
        if (g.share > 0.0f) {
            tr = vmul(tr, vs(g.tint, g.share));
            odd = !odd;
        }
        o = offset_origin(vadd(o, vs(d, h.t)), dot(g.ng, d) < 0.0f ? vneg(g.ng) : g.ng);
        left -= h.t;
        if (left <= 0.0f || all_nonpos(tr)) break;
    }
    if (sky && odd) return v3(0, 0, 0);
    return tr;
}
/* shadow_tr (846-855): 0 past a shadow caster; else the glass's share (objects that cast no shadow are glass) */
static V shadow_tr(Ctx *c, V o, V d, float tmax, int sky) {
    if (occluded(c, o, d, tmax, OR_TRACE_SHADOW)) return v3(0, 0, 0);
    if (!glass_on(c)) return v3(1, 1, 1);
    return glass_tr(c, o, d, tmax, sky);
}
/* light_origin (2153-2156): a shadow ray's origin on the side l leaves the surface from */
static inline V light_origin(const Ctx *c, const Surface *s, V o, V l) {
    if (glass_on(c) && dot(l, s->ng) < 0.0f) return offset_origin(s->p, vneg(s->ng));
    return o;
}

/* unpack2x16float: IEEE binary16 to binary32, exactly */
static float half_to_float(uint32_t h) {
    uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, u;
    if (e == 0x1fu)
        u = sign | 0x7f800000u | (m << 13);
    else if (e == 0u) {
        float f = (float)m * 5.9604644775390625e-8f; /* m x 2^-24, exact */
        memcpy(&u, &f, 4);
        u |= sign;
    } else
        u = sign | ((e + 112u) << 23) | (m << 13);
    float f;
    memcpy(&f, &u, 4);
    return f;
}
/* ---- the environment map (pathtrace.wgsl:1889-1941) ------------------------------------------------------------
 * atan2 and acos by Cephes atanf / asinf (S. L. Moshier): the argument reduced to [0, tan(pi/8)] or [0, 1/2] and a
 * minimax polynomial there, in fmaf; the kernel has the same steps, so the same bits. */
static float atan_red(float x) { /* atan of x >= 0 */
    float y = 0.0f;
    if (x > 2.414213562373095f) {
        y = 1.5707963267948966f;
        x = -(1.0f / x);
    } else if (x > 0.4142135623730950f) {
        y = 0.7853981633974483f;
        x = (x - 1.0f) / (x + 1.0f);
    }
    const float z = x * x;
    float p = fmaf(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = fmaf(p, z, 1.99777106478e-1f);
    p = fmaf(p, z, -3.33329491539e-1f);
    return y + fmaf(p * z, x, x);
}
float or_atan2(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 1.5707963267948966f : (y < 0.0f ? -1.5707963267948966f : 0.0f);
    float w = 0.0f;
    if (x < 0.0f) w = y < 0.0f ? -3.141592653589793f : 3.141592653589793f;
    const float q = y / x;
    const float a = q < 0.0f ? -atan_red(-q) : atan_red(q);
    return w + a;
}
static float asin_red(float a) { /* asin of a in [0, 1] */
    float x, z;
    int flag = 0;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
    }
    float r = x;
    if (!(a < 1e-4f)) {
        float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
        p = fmaf(p, z, 4.5470025998e-2f);
        p = fmaf(p, z, 7.4953002686e-2f);
        p = fmaf(p, z, 1.6666752422e-1f);
        r = fmaf(p * z, x, x);
    }
    if (flag) r = 1.5707963267948966f - (r + r);
    return r;
}
float or_acos(float x) {
    if (!(x >= -1.0f && x <= 1.0f)) return NAN;
    if (x < -0.5f) return 3.141592653589793f - 2.0f * asin_red(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin_red(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - (x < 0.0f ? -asin_red(-x) : asin_red(x));
}
/* WGSL's float to u32: saturating, NaN to 0 */
static inline uint32_t u32_sat(float x) { return x > 0.0f ? (x < 4294967040.0f ? (uint32_t)x : 0xffffff00u) : 0u; }
/* env_uv (1889-1893) */
static void env_uv(V d, float *u, float *v) {
    *u = or_atan2(d.x, -d.z) / TWO_PI + 0.5f;
    *v = or_acos(clampf(d.y, -1.0f, 1.0f)) / PI_F;
}
static inline float env_cdf_at(const Ctx *c, uint64_t i) { return c->s->env_cdf[ck(c, i, c->s->env_cdf_count)]; }
/* env_radiance (1895-1898): the map's texture read at level 0 as the reference's sampler defines it (repeat in u
 * and v, linear), with float weights */
static V env_radiance(const Ctx *c, V d) {
    if (pf(c->P, ENV, 0) < 0.5f) return v3(pf(c->P, SKY, 0), pf(c->P, SKY, 1), pf(c->P, SKY, 2));
    float u, v;
    env_uv(d, &u, &v);
    const uint32_t w = c->s->env_w, h = c->s->env_h;
    float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    const float fx = x - x0f, fy = y - y0f;
    if (!(x0f >= -1048576.0f && x0f <= 1048576.0f)) x0f = 0.0f;
    if (!(y0f >= -1048576.0f && y0f <= 1048576.0f)) y0f = 0.0f;
    const int32_t xi = (int32_t)x0f, yi = (int32_t)y0f, wi = (int32_t)w, hi = (int32_t)h;
    const uint32_t xa = (uint32_t)(((xi % wi) + wi) % wi), ya = (uint32_t)(((yi % hi) + hi) % hi);
    const uint32_t xb = xa + 1u == w ? 0u : xa + 1u, yb = ya + 1u == h ? 0u : ya + 1u;
    const uint64_t n = (uint64_t)w * h;
    const uint16_t *t00 = c->s->env + 4u * ck(c, (uint64_t)ya * w + xa, n);
    const uint16_t *t10 = c->s->env + 4u * ck(c, (uint64_t)ya * w + xb, n);
    const uint16_t *t01 = c->s->env + 4u * ck(c, (uint64_t)yb * w + xa, n);
    const uint16_t *t11 = c->s->env + 4u * ck(c, (uint64_t)yb * w + xb, n);
    float o[3];
    for (int j = 0; j < 3; j++)
        o[j] = mixf(mixf(half_to_float(t00[j]), half_to_float(t10[j]), fx),
                    mixf(half_to_float(t01[j]), half_to_float(t11[j]), fx), fy);
    return v3(o[0], o[1], o[2]);
}
/* env_pdf (1900-1913): the rows' and the row's CDF steps of the texel d falls in, per solid angle */
static float env_pdf(const Ctx *c, V d) {
    if (pf(c->P, ENV, 0) < 0.5f) return 1.0f / FOUR_PI;
    const uint32_t w = c->s->env_w, h = c->s->env_h;
    float u, v;
    env_uv(d, &u, &v);
    const uint32_t x = umin(u32_sat(u * (float)w), w - 1u), y = umin(u32_sat(v * (float)h), h - 1u);
    const float row_p = env_cdf_at(c, y) - (y > 0u ? env_cdf_at(c, y - 1u) : 0.0f);
    const uint64_t base = (uint64_t)h + (uint64_t)y * w;
    const float col_p = env_cdf_at(c, base + x) - (x > 0u ? env_cdf_at(c, base + x - 1u) : 0.0f);
    float sin_t, cos_t;
    or_sincos(((float)y + 0.5f) / (float)h * PI_F, &sin_t, &cos_t);
    return row_p * col_p * (float)w * (float)h / (2.0f * PI_F * PI_F * fmaxf(sin_t, 1e-4f));
}
/* lower_bound (1915-1923) */
static uint32_t env_lower_bound(const Ctx *c, uint64_t start, uint32_t count, float x) {
    uint32_t lo = 0u, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (env_cdf_at(c, start + mid) < x)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return umin(lo, count - 1u);
}
static int env_sampled(const Ctx *c) {
    return pu(c->P, COUNTS2, 1) != 0u &&
           (pf(c->P, ENV, 0) > 0.5f || luminance(v3(pf(c->P, SKY, 0), pf(c->P, SKY, 1), pf(c->P, SKY, 2))) > 0.0f);
}
/* sample_env (1925-1941) for the uniforms u[4]: the uniform sphere for a constant sky; the map's row by its CDF,
 * the texel in it by the row's, a uniform point of the texel in (u, v) */
static V sample_env_u(const Ctx *c, const float *u) {
    if (pf(c->P, ENV, 0) < 0.5f) {
        float z = 1.0f - 2.0f * u[0];
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float phi = TWO_PI * u[1];
        float sp, cp;
        or_sincos(phi, &sp, &cp);
        return v3(r * cp, z, r * sp);
    }
    const uint32_t w = c->s->env_w, h = c->s->env_h;
    const uint32_t y = env_lower_bound(c, 0u, h, u[0]);
    const uint32_t x = env_lower_bound(c, (uint64_t)h + (uint64_t)y * w, w, u[1]);
    const float uu = ((float)x + u[2]) / (float)w, vv = ((float)y + u[3]) / (float)h;
    const float theta = vv * PI_F, phi = (uu - 0.5f) * 2.0f * PI_F;
    float st, ct, sp, cp;
    or_sincos(theta, &st, &ct);
    or_sincos(phi, &sp, &cp);
    return v3(st * sp, ct, -st * cp);
}
static V sample_env(Ctx *c, uint32_t rd) {
    float u[4];
    or_rng4(&c->g, rd, u);
    return sample_env_u(c, u);
}

static V sun_disks(const Ctx *c, V d, float pdf) {
    V sum = v3(0, 0, 0);
    for (uint32_t i = 0; i < pu(c->P, COUNTS2, 0); i++) {
        const float *li = light_at(c, i);
        if (dot(d, vneg(ld3(li + 4))) >= li[7]) {
            float omega = TWO_PI * (1.0f - li[7]);
            sum = vadd(sum, vs(vds(ld3(li + 8), omega), power_heuristic(pdf, 1.0f / omega)));
        }
    }
    return sum;
}
/* sky_mis (1963-1974); a camera ray (or a path bent straight through glass) that leaves the scene sees three.js's
 * scene.background where P.background.w says there is one (the reference's F_BACKGROUND) */
static V sky_mis(const Ctx *c, V d, float pdf) {
    if (pdf <= 0.0f && pf(c->P, BACKGROUND, 3) > 0.5f)
        return v3(pf(c->P, BACKGROUND, 0), pf(c->P, BACKGROUND, 1), pf(c->P, BACKGROUND, 2));
    float wgt = 1.0f;
    if (pdf > 0.0f && pu(c->P, COUNTS2, 1) != 0u) wgt = power_heuristic(pdf, env_pdf(c, d));
    V col = vs(env_radiance(c, d), wgt);
    if (pdf > 0.0f) col = vadd(col, sun_disks(c, d, pdf));
    return col;
}

static inline uint32_t emitter_count(const Ctx *c) { return pu(c->P, COUNTS, 1); }
static inline const float *emitter_totals(const Ctx *c) {
    return emitter_at(c, emitter_count(c) + EMITTER_GUIDE / 4u + 1u);
}
static int emitters_on(const Ctx *c) { return emitter_count(c) != 0u && emitter_totals(c)[0] > 0.0f; }
static inline uint32_t emitter_guide(const Ctx *c, uint32_t k) {
    return fb(emitter_at(c, emitter_count(c) + k / 4u)[k % 4u]);
}
static const float *emitter_pick_u(const Ctx *c, float u) {
    uint32_t k = (uint32_t)(u * (float)EMITTER_GUIDE);
    if (k > EMITTER_GUIDE - 1u) k = EMITTER_GUIDE - 1u;
    uint32_t lo = emitter_guide(c, k), hi = emitter_guide(c, k + 1u);
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (emitter_at(c, mid)[1] < u)
            lo = mid + 1u;
        else
            hi = mid;
    }
    uint32_t n = emitter_count(c);
    return emitter_at(c, lo < n - 1u ? lo : n - 1u);
}

/* ---- the light tree (pathtrace.wgsl:489-618) ------------------------------------------------------------------ */
#define LT_LEAF 0x80000000u
#define LT_PATH_BITS 26u
static inline int lt_on(const Ctx *c) { return pf(c->P, LTREE, 3) > 0.5f; }
static inline const float *lt_node(const Ctx *c, uint32_t i) {
    return c->s->light_tree + 4 * ck(c, i, c->s->light_tree_count);
}
/* lt_imp (504-524): power x the best cosine any point of the node's sphere has at x / max(d^2, r^2); the offline
 * kernel has no approximations, so no node is cut for distance */
static float lt_imp(const float *nd, V x, V n) {
    uint32_t a = fb(nd[0]), b = fb(nd[1]);
    V dv = vsub(v3(half_to_float(a & 0xffffu), half_to_float(a >> 16), half_to_float(b & 0xffffu)), x);
    float r = half_to_float(b >> 16);
    float d2 = dot(dv, dv);
    float r2 = r * r;
    float cb = 1.0f;
    if (d2 > r2 && dot(n, n) > 0.0f) {
        float inv = 1.0f / sqrtf(d2);
        float cos_t = dot(n, dv) * inv;
        float sin_b = r * inv;
        float cos_b = sqrtf(fmaxf(1.0f - sin_b * sin_b, 0.0f));
        if (cos_t < cos_b) cb = fmaxf(cos_t * cos_b + sqrtf(fmaxf(1.0f - cos_t * cos_t, 0.0f)) * sin_b, 0.0f);
    }
    return nd[2] * cb / fmaxf(d2, r2);
}
/* emitter_pick_range (527-540) */
static const float *emitter_pick_range(const Ctx *c, uint32_t lo, uint32_t hi, float r) {
    float p0 = lo > 0u ? emitter_at(c, (lo > 1u ? lo : 1u) - 1u)[1] : 0.0f;
    float p1 = emitter_at(c, hi - 1u)[1];
    float u = fminf(p0 + r * (p1 - p0), p1);
    uint32_t k = (uint32_t)(u * (float)EMITTER_GUIDE);
    if (k > EMITTER_GUIDE - 1u) k = EMITTER_GUIDE - 1u;
    uint32_t l = emitter_guide(c, k), h = emitter_guide(c, k + 1u);
    while (l < h) {
        uint32_t mid = (l + h) / 2u;
        if (emitter_at(c, mid)[1] < u)
            l = mid + 1u;
        else
            h = mid;
    }
    if (l < lo) l = lo;
    if (l > hi - 1u) l = hi - 1u;
    return emitter_at(c, l);
}

typedef struct {
    const float *e;
    float p, w;
} EPick;
/* emitter_draw (547-585): without the tree the CDF; with it the tail by its share, else the tree descended by
 * importance at (x, n) to a leaf and one of its entries by weight */
static EPick emitter_draw_u(const Ctx *c, V x, V n, const float *r3) {
    const float *tot = emitter_totals(c);
    EPick pk = {NULL, 0.0f, 0.0f};
    if (!(tot[0] > 0.0f)) return pk;
    if (!lt_on(c)) {
        pk.e = emitter_pick_u(c, r3[0]);
        pk.p = pk.e[2] / tot[0];
        pk.w = pk.e[2];
        return pk;
    }
    if (r3[0] < tot[1]) {
        pk.e = emitter_pick_range(c, pu(c->P, LTREE, 1), emitter_count(c), r3[1]);
        pk.p = pk.e[2] / tot[0];
        pk.w = pk.e[2];
        return pk;
    }
    float p = 1.0f - tot[1];
    float u = r3[1];
    uint32_t w = fb(lt_node(c, 0u)[3]);
    for (uint32_t k = 0u; k <= LT_PATH_BITS && (w & LT_LEAF) == 0u; k++) {
        const float *l = lt_node(c, w), *r = lt_node(c, w + 1u);
        float il = lt_imp(l, x, n), ir = lt_imp(r, x, n);
        if (!(il + ir > 0.0f)) return pk;
        float pl = il / (il + ir);
        if (u < pl) {
            w = fb(l[3]);
            p *= pl;
            u = u / pl;
        } else {
            w = fb(r[3]);
            p *= 1.0f - pl;
            u = (u - pl) / (1.0f - pl);
        }
        u = fminf(fmaxf(u, 0.0f), 0.99999994f);
    }
    if ((w & LT_LEAF) == 0u) return pk;
    uint32_t lo = w & 0x1fffffffu, hi = lo + ((w >> 29) & 3u) + 1u;
    float wsum = 0.0f;
    for (uint32_t j = lo; j < hi; j++) wsum += emitter_at(c, j)[2];
    float t = r3[2] * wsum;
    const float *e = emitter_at(c, hi - 1u);
    for (uint32_t j = lo; j < hi; j++) {
        const float *cc = emitter_at(c, j);
        if (t < cc[2]) {
            e = cc;
            break;
        }
        t -= cc[2];
    }
    pk.e = e;
    pk.p = p * e[2] / wsum;
    pk.w = e[2];
    return pk;
}
static EPick emitter_draw(Ctx *c, V x, V n, uint32_t rd) {
    float r3[3];
    or_rng3(&c->g, rd, r3);
    return emitter_draw_u(c, x, n, r3);
}

typedef struct {
    V v0, e1, e2;
    uint32_t tri, inst, code, part, flags;
} Em;
static Em emitter_tri_at(const Ctx *c, uint32_t i) {
    const float *r = c->s->emitter_tris + 16 * ck(c, i, c->s->emitter_tri_count);
    Em e;
    e.v0 = ld3(r);
    e.tri = fb(r[3]);
    e.e1 = ld3(r + 4);
    e.inst = fb(r[7]);
    e.e2 = ld3(r + 8);
    e.code = fb(r[11]);
    e.part = fb(r[12]);
    e.flags = fb(r[13]);
    return e;
}
/* emitter_pmf (590-618): the probability emitter_draw at (x, n) picks emitter `ei`: one of the tree's (its code names
 * its path) by the same choices down the path and its weight in its leaf; any other by its entry's weight */
static float emitter_pmf(const Ctx *c, uint32_t ei, const Em *em, V x, V n) {
    const float *tot = emitter_totals(c);
    if (!(tot[0] > 0.0f)) return 0.0f;
    uint32_t code = em->code & ~0x80000000u;
    if ((code >> LT_PATH_BITS) == 0u) return emitter_at(c, code)[2] / tot[0];
    uint32_t depth = (code >> LT_PATH_BITS) - 1u;
    float p = 1.0f - tot[1];
    uint32_t w = fb(lt_node(c, 0u)[3]);
    for (uint32_t k = 0u; k < depth; k++) {
        const float *l = lt_node(c, w), *r = lt_node(c, w + 1u);
        float il = lt_imp(l, x, n), ir = lt_imp(r, x, n);
        if (!(il + ir > 0.0f)) return 0.0f;
        uint32_t right = (code >> k) & 1u;
        p *= (right == 1u ? ir : il) / (il + ir);
        w = fb(right == 1u ? r[3] : l[3]);
    }
    uint32_t lo = w & 0x1fffffffu, hi = lo + ((w >> 29) & 3u) + 1u;
    float wsum = 0.0f, we = 0.0f;
    for (uint32_t j = lo; j < hi; j++) {
        const float *cc = emitter_at(c, j);
        wsum += cc[2];
        if (fb(cc[0]) == ei) we = cc[2];
    }
    return p * we / wsum;
}

/* emitter_radiance (2269-2283): the emission at the sample (b1, b2), its map read at a 5 cm footprint */
static V emitter_radiance(const Ctx *c, const Em *em, float b1, float b2) {
    TriAttr a = tri_data(c, em->tri, em->inst, em->part).a;
    if (em->inst != OR_NO_INST) {
        uint32_t obj = inst_at(c, em->inst)->info[1] + fb(a.x[1]);
        memcpy(&a.x[1], &obj, 4);
    }
    const float *m = material_at(c, material_of(c, &a, em->inst));
    float uv[2] = {0.0f, 0.0f}, lod = 0.0f;
    if (tex_on(c)) {
        float den[2];
        mesh_uv(&a, m, b1, b2, uv);
        mesh_density(em->e1, em->e2, &a, m, den);
        lod = mesh_lod(den, 0.0025f, 1.0f);
    }
    return vs(emission(c, m, uv, lod), obj_power(c, object_of(&a)));
}

typedef struct {
    V le, l;
    float dist, pdf_w, w;
} ES;

static ES sample_emitter_ris(Ctx *c, V x, V n, uint32_t m, const Surface *sf, const Lobes *lb, V v, uint32_t rd) {
    ES r = {v3(0, 0, 0), v3(0, 1, 0), 0.0f, 0.0f, 0.0f};
    float wsum = 0.0f, pick_b1 = 0.0f, pick_b2 = 0.0f, pick_ph = 0.0f, pick_pdf = 0.0f;
    uint32_t pick_i = 0;
    for (uint32_t k = 0; k < m; k++) {
        uint32_t dk = or_rng_sub(rd, k);
        EPick pk = emitter_draw(c, x, n, dk);
        float bb[3];
        or_rng3(&c->g, or_rng_sub(dk, 1u), bb);
        float b1 = bb[0], b2 = bb[1];
        if (pk.p <= 0.0f) continue;
        if (b1 + b2 > 1.0f) {
            b1 = 1.0f - b1;
            b2 = 1.0f - b2;
        }
        uint32_t ei = fb(pk.e[0]);
        Em em = emitter_tri_at(c, ei);
        V lp = vadd(vadd(em.v0, vs(em.e1, b1)), vs(em.e2, b2));
        V cr = cross(em.e1, em.e2);
        float a2 = len(cr);
        if (!(a2 > 0.0f)) continue;
        V ln = vds(cr, a2);
        V dl = vsub(lp, x);
        float d2 = fmaxf(dot(dl, dl), 1e-8f);
        V l = vs(dl, 1.0f / sqrtf(d2));
        float g = fabsf(dot(ln, l)) * (dot(n, n) == 0.0f ? 1.0f : fmaxf(dot(n, l), 0.0f)) / d2;
        if (m > 1u) g = fabsf(dot(ln, l)) * luminance(eval_bsdf(c, sf, lb, v, l)) / d2;
        float ph = pk.w / (pk.e[3] * 1.0f) * g;
        float pdf_a = pk.p / (0.5f * a2);
        float w = ph / pdf_a;
        if (!(w > 0.0f)) continue;
        wsum += w;
        if (bb[2] * wsum < w) {
            pick_b1 = b1;
            pick_b2 = b2;
            pick_ph = ph;
            pick_i = ei;
            pick_pdf = pdf_a;
        }
    }
    if (pick_ph <= 0.0f) return r;
    Em em = emitter_tri_at(c, pick_i);
    V lp = vadd(vadd(em.v0, vs(em.e1, pick_b1)), vs(em.e2, pick_b2));
    V ln = normalize(cross(em.e1, em.e2));
    V dl = vsub(lp, x);
    float d2 = dot(dl, dl);
    float dist = sqrtf(d2);
    V l = vds(dl, dist);
    float cos_l = dot(ln, vneg(l));
    if (!(em.flags & EMITTER_CARD)) cos_l = fabsf(cos_l);
    if (cos_l <= 0.0f) return r;
    r.le = emitter_radiance(c, &em, pick_b1, pick_b2);
    r.l = l;
    r.dist = dist;
    r.pdf_w = pick_pdf * d2 / cos_l;
    r.w = (wsum / (float)m) / pick_ph * cos_l / d2;
    return r;
}

static ES sample_emitter(Ctx *c, V x, V n, uint32_t rd) {
    ES r = {v3(0, 0, 0), v3(0, 1, 0), 0.0f, 0.0f, 0.0f};
    EPick pk = emitter_draw(c, x, n, rd);
    if (pk.p <= 0.0f) return r;
    Em em = emitter_tri_at(c, fb(pk.e[0]));
    float bb[2];
    or_rng2(&c->g, or_rng_sub(rd, 1u), bb);
    float b1 = bb[0], b2 = bb[1];
    if (b1 + b2 > 1.0f) {
        b1 = 1.0f - b1;
        b2 = 1.0f - b2;
    }
    V lp = vadd(vadd(em.v0, vs(em.e1, b1)), vs(em.e2, b2));
    V cr = cross(em.e1, em.e2);
    float area = 0.5f * len(cr);
    V ln = vds(cr, 2.0f * area);
    V dl = vsub(lp, x);
    float d2 = dot(dl, dl);
    float dist = sqrtf(d2);
    V l = vds(dl, dist);
    float cos_l = dot(ln, vneg(l));
    if (!(em.flags & EMITTER_CARD)) cos_l = fabsf(cos_l);
    if (cos_l <= 0.0f) return r;
    r.le = emitter_radiance(c, &em, b1, b2);
    r.l = l;
    r.dist = dist;
    r.pdf_w = pk.p / area * d2 / cos_l;
    r.w = 1.0f / r.pdf_w;
    return r;
}

static V env_light(Ctx *c, const Surface *s, const Lobes *lb, V v, V o) {
    V l = sample_env(c, or_rng_dim(&c->g, D_ENV));
    V f = eval_bsdf(c, s, lb, v, l);
    if (any_pos(f)) {
        float pdf_e = env_pdf(c, l);
        if (pdf_e > 0.0f) {
            V vis = shadow_tr(c, light_origin(c, s, o, l), l, OR_INF, 1); /* sky_tr (975-977) */
            if (any_pos(vis)) {
                V e = vds(vs(vmul(env_radiance(c, l), f), light_mis(c, pdf_e, s, lb, v, l)), pdf_e);
                return glass_on(c) ? vmul(e, vis) : e;
            }
        }
    }
    return v3(0, 0, 0);
}

/* ---- local lights (pathtrace.wgsl:2039-2262) --------------------------------------------------------------- */
#define RIS_EXACT 8u
#define RIS_M 8u
typedef struct {
    V c, l, sl;
    float dist;
} LS;
static inline float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
/* local_light (2042-2104): a point, spot or rect light seen from x: its intensity x attenuation (unshadowed), the
 * direction to its centre (or to the rect's sampled point), and a shadow ray toward a point of it (u: 2 uniforms) */
static LS local_light(const float *li, V x, const float *u) {
    LS r;
    const V pos = ld3(li);
    const float range = li[11];
    if ((uint32_t)li[3] == 3u) { /* the rect: a uniform point on it, luminance x area x cos_light / d^2 */
        const V n = ld3(li + 4), ux = ld3(li + 12);
        const V uy = vs(normalize(cross(n, ux)), li[7]);
        const V lp = vadd(vadd(pos, vs(ux, 2.0f * u[0] - 1.0f)), vs(uy, 2.0f * u[1] - 1.0f));
        const V dl = vsub(lp, x);
        const float d2 = fmaxf(dot(dl, dl), 1e-8f);
        const float dist = sqrtf(d2);
        const V l = vds(dl, dist);
        const float cos_l = dot(n, vneg(l));
        float k = 0.0f;
        if (cos_l > 0.0f) {
            const float area = 4.0f * len(ux) * li[7];
            k = area * cos_l / d2 * smoothstepf(li[18] - 0.04f, li[18] + 0.04f, cos_l);
            if (range > 0.0f) {
                const float q = d2 / (range * range);
                const float win = clampf(1.0f - q * q, 0.0f, 1.0f);
                k *= win * win;
            }
        }
        r.c = vs(ld3(li + 8), k), r.l = l, r.sl = l, r.dist = dist;
        return r;
    }
    V lp = pos;
    const float radius = li[7];
    if (radius > 0.0f) {
        const float z = 1.0f - 2.0f * u[0];
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        or_sincos(TWO_PI * u[1], &sp, &cp);
        lp = vadd(lp, vs(v3(rr * cp, rr * sp, z), radius));
    }
    const V dl = vsub(pos, x);
    const float d2 = fmaxf(dot(dl, dl), 1e-8f);
    const V l = vs(dl, 1.0f / sqrtf(d2));
    float atten = 1.0f / d2;
    if (range > 0.0f) {
        const float q = d2 / (range * range);
        const float win = clampf(1.0f - q * q, 0.0f, 1.0f);
        atten *= win * win;
    }
    if ((uint32_t)li[3] == 2u) { /* the cone: glTF KHR_lights_punctual's angular attenuation */
        const float cone = clampf((dot(vneg(l), ld3(li + 4)) - li[12]) * li[13], 0.0f, 1.0f);
        atten *= cone * cone;
    }
    const V ls = vsub(lp, x);
    const float dist = fmaxf(len(ls), 1e-6f);
    r.c = vs(ld3(li + 8), atten), r.l = l, r.sl = vds(ls, dist), r.dist = dist;
    return r;
}
/* light_vis (908-915): a light that casts no shadow is seen (no ray) */
static V light_vis(Ctx *c, int shadows, V o, V d, float tmax) {
    if (!shadows) return v3(1, 1, 1);
    return shadow_tr(c, o, d, tmax, 0);
}
/* local_lights (2158-2262, offline: no light tiles, no volume): at points seen directly, or with few lights, every
 * unshadowed light and, when few, every shadowed one, exactly (one shadow ray each); otherwise one of RIS_M (3 past
 * the first bounce) uniform candidates resampled by unshadowed contribution, one shadow ray. (The reference skips
 * whole groups of 16 lights a point is outside the reach of: the same sum, every light of such a group being
 * rejected by its own range test below.) */
static V local_lights(Ctx *c, const Surface *s, const Lobes *lb, V v, V o, uint32_t rd) {
    const uint32_t nd = pu(c->P, COUNTS2, 0), nall = pu(c->P, COUNTS, 0);
    const uint32_t nl = nall - nd;
    V sum = v3(0, 0, 0);
    if (nl == 0u) return sum;
    const V x = s->p;
    const int seen = c->g.depth == 0u || (c->g.depth == 1u && c->mirror_path);
    const int glass = glass_on(c);
    uint32_t first = nd, count = nl;
    if (seen || nl <= RIS_EXACT) {
        const uint32_t ns = pu(c->P, COUNTS2, 3), n_sh = nall - ns;
        const uint32_t end = (n_sh <= RIS_EXACT || nl <= RIS_EXACT) ? nall : ns;
        for (uint32_t i = nd; i < end; i++) {
            const float *li = light_at(c, i);
            if (all_nonpos(ld3(li + 8))) continue;
            const V dc = vsub(ld3(li), x);
            const float reach = li[11] + ((uint32_t)li[3] == 3u ? len(ld3(li + 12)) + li[7] : li[7]);
            if (li[11] > 0.0f && dot(dc, dc) > reach * reach) continue;
            if ((uint32_t)li[3] != 3u && dot(dc, s->ng) < -li[7] && !two_sided(c, lb)) continue;
            float u[2];
            or_rng2(&c->g, or_rng_sub(rd, i), u);
            const LS ls = local_light(li, x, u);
            if (all_nonpos(ls.c)) continue;
            const V f = eval_bsdf_k(c, s, lb, v, ls.l, li[17]);
            if (any_pos(f)) {
                const V vis = light_vis(c, li[15] >= 0.5f, light_origin(c, s, o, ls.l), ls.sl, ls.dist * 0.999f);
                if (any_pos(vis)) {
                    const V e = vmul(ls.c, f);
                    sum = vadd(sum, glass ? vmul(e, vis) : e);
                }
            }
        }
        if (end == nall) return sum;
        first = ns;
        count = n_sh;
    }
    const uint32_t m_cand = seen ? RIS_M : 3u;
    float wsum = 0.0f, pick_ph = 0.0f;
    uint32_t pick = 0;
    LS pick_ls;
    memset(&pick_ls, 0, sizeof pick_ls);
    for (uint32_t k = 0; k < m_cand; k++) {
        float r4[4];
        or_rng4(&c->g, or_rng_sub(rd ^ 0x80000000u, k), r4);
        uint32_t j = (uint32_t)(r4[0] * (float)count);
        const uint32_t idx = first + (j < count - 1u ? j : count - 1u);
        const LS ls = local_light(light_at(c, idx), x, r4 + 1);
        float ph = luminance(ls.c);
        if (two_sided(c, lb))
            ph *= fabsf(dot(s->n, ls.l));
        else
            ph *= fmaxf(dot(s->n, ls.l), 0.0f) + 0.05f * fmaxf(dot(s->ng, ls.l), 0.0f);
        if (ph <= 0.0f) continue;
        const float w = ph * 1.0f;
        wsum += w;
        if (r4[3] * wsum < w) pick = idx, pick_ph = ph, pick_ls = ls;
    }
    if (pick_ph > 0.0f) {
        const float *lk = light_at(c, pick);
        const V cc = vmul(pick_ls.c, eval_bsdf_k(c, s, lb, v, pick_ls.l, lk[17]));
        if (any_pos(cc)) {
            const V vis = light_vis(c, lk[15] >= 0.5f, light_origin(c, s, o, pick_ls.l), pick_ls.sl,
                                    pick_ls.dist * 0.999f);
            if (any_pos(vis))
                sum = vadd(sum, vs(glass ? vmul(cc, vis) : cc, wsum * (float)count / ((float)m_cand * pick_ph)));
        }
    }
    return sum;
}

static V direct_light(Ctx *c, const Surface *s, const Lobes *lb, V v) {
    V sum = v3(0, 0, 0);
    V o = offset_origin(s->p, s->ng);
    for (uint32_t i = 0; i < pu(c->P, COUNTS2, 0); i++) {
        const float *li = light_at(c, i);
        float cos_max = li[7];
        V to_light = vneg(ld3(li + 4));
        float u[2];
        or_rng2(&c->g, or_rng_sub(or_rng_dim(&c->g, D_SUN), i), u);
        float cos_t = 1.0f - u[0] * (1.0f - cos_max);
        float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
        float phi = TWO_PI * u[1];
        float sp, cp;
        or_sincos(phi, &sp, &cp);
        M3 tb = basis(to_light);
        V l = mmul(tb, v3(sin_t * cp, sin_t * sp, cos_t));
        V col = ld3(li + 8);
        if (all_nonpos(col)) continue;
        V f = eval_bsdf_k(c, s, lb, v, l, li[17]);
        if (any_pos(f)) {
            V vis = li[15] >= 0.5f ? shadow_tr(c, light_origin(c, s, o, l), l, OR_INF, 0) : v3(1, 1, 1);
            if (any_pos(vis)) {
                float pdf_l = 1.0f / (TWO_PI * (1.0f - cos_max));
                float wgt = light_mis(c, pdf_l, s, lb, v, l);
                V e = vs(vmul(col, f), wgt);
                sum = vadd(sum, glass_on(c) ? vmul(e, vis) : e);
            }
        }
    }
    if (pu(c->P, COUNTS, 0) > pu(c->P, COUNTS2, 0)) sum = vadd(sum, local_lights(c, s, lb, v, o, or_rng_dim(&c->g, D_LOCAL)));
    /* rough glass takes emitters from both sides, as a point without a hemisphere (2488-2491) */
    V ne = two_sided(c, lb) ? v3(0, 0, 0) : s->n;
    c->mis_x = o;
    c->mis_n = ne;
    uint32_t emitters = emitters_on(c) ? emitter_count(c) : 0u;
    if (emitters > 0u) {
        ES es;
        int seen = c->g.depth == 0u || (c->g.depth == 1u && c->mirror_path);
        /* with the tree its candidates are already near the point: P.ltree.w of them at camera hits, one elsewhere */
        if (lt_on(c))
            es = sample_emitter_ris(c, o, ne, c->g.depth == 0u ? (uint32_t)pf(c->P, LTREE, 3) : 1u, s, lb, v,
                                    or_rng_dim(&c->g, D_EMIT));
        else if (seen)
            es = sample_emitter_ris(c, o, ne, c->g.depth == 0u ? EMIT_RIS_M : EMIT_RIS_M / 2u, s, lb, v,
                                    or_rng_dim(&c->g, D_EMIT));
        else if (pf(c->P, LIGHT_LOD, 3) > 0.0f)
            es = sample_emitter_ris(c, o, ne, 2u, s, lb, v, or_rng_dim(&c->g, D_EMIT));
        else
            es = sample_emitter(c, o, ne, or_rng_dim(&c->g, D_EMIT));
        if (es.pdf_w > 0.0f) {
            V f = eval_bsdf(c, s, lb, v, es.l);
            if (any_pos(f)) {
                if (!occluded(c, light_origin(c, s, o, es.l), es.l, es.dist * 0.999f, OR_TRACE_ANY)) {
                    float wgt = light_mis(c, es.pdf_w, s, lb, v, es.l);
                    sum = vadd(sum, vs(vs(vmul(es.le, f), wgt), es.w));
                }
            }
        }
    }
    if (env_sampled(c)) sum = vadd(sum, env_light(c, s, lb, v, o));
    return sum;
}

static float emit_mis(Ctx *c, const or_hit *hit, V d, float pdf) {
    if (!emitters_on(c)) return 1.0f;
    uint32_t rows = inst_at(c, hit->inst)->info[3];
    uint32_t first = row_word(c, rows, hit->part, 2u);
    if (first == NO_EMITTER) return 1.0f;
    uint32_t rank = fb(c->s->geo.leaves[ck(c, hit->tri, c->s->geo.leaf_count)].r[1][3]);
    uint32_t entry = first + rank;
    Em em = emitter_tri_at(c, entry);
    V cr = cross(em.e1, em.e2);
    float area = 0.5f * len(cr);
    float cos_l = fabsf(dot(normalize(cr), d));
    float sel = emitter_pmf(c, entry, &em, c->mis_x, c->mis_n);
    if (!(sel > 0.0f)) return 1.0f;
    return power_heuristic(pdf, sel / fmaxf(area, 1e-12f) * hit->t * hit->t / fmaxf(cos_l, 1e-6f));
}

/* ---- the path ------------------------------------------------------------------------------------------------- */
static inline V clamp_lum(V col, float clamp_v) {
    float lum = luminance(col);
    if (lum > clamp_v) col = vs(col, clamp_v / lum);
    return col;
}

static void path(Ctx *c, uint32_t px, uint32_t py, float *pix) {
    const uint32_t *P = c->P;
    uint32_t w = pu(P, SIZE, 0), h = pu(P, SIZE, 1);
    or_rng_start(&c->g, px, py, pu(P, SIZE, 2), pu(P, SIZE, 3), 0u);
    c->g.depth = 0;
    float cam[4];
    or_rng4(&c->g, or_rng_dim(&c->g, D_CAMERA), cam);
    float ndx = ((float)px + cam[0]) / (float)w * 2.0f - 1.0f;
    float ndy = 1.0f - ((float)py + cam[1]) / (float)h * 2.0f;
    V fwd = v3(pf(P, CAM_FWD, 0), pf(P, CAM_FWD, 1), pf(P, CAM_FWD, 2));
    V right = v3(pf(P, CAM_RIGHT, 0), pf(P, CAM_RIGHT, 1), pf(P, CAM_RIGHT, 2));
    V up = v3(pf(P, CAM_UP, 0), pf(P, CAM_UP, 1), pf(P, CAM_UP, 2));
    float tanh = pf(P, CAM_POS, 3), aspect = pf(P, CAM_FWD, 3);
    V d = normalize(vadd(vadd(fwd, vs(right, ndx * tanh * aspect)), vs(up, ndy * tanh)));
    V o = v3(pf(P, CAM_POS, 0), pf(P, CAM_POS, 1), pf(P, CAM_POS, 2));
    if (pf(P, CAM_RIGHT, 3) > 0.0f) { /* the thin lens, pathtrace.wgsl:4017-4023 */
        V focus = vadd(o, vs(d, pf(P, CAM_UP, 3) / dot(d, fwd)));
        float r = sqrtf(cam[2]) * pf(P, CAM_RIGHT, 3);
        float phi = TWO_PI * cam[3], sp, cp;
        or_sincos(phi, &sp, &cp);
        o = vadd(o, vadd(vs(right, r * cp), vs(up, r * sp)));
        d = normalize(vsub(focus, o));
    }

    V radiance = v3(0, 0, 0), thr = v3(1, 1, 1), albedo = v3(0, 0, 0), normal = v3(0, 0, 0);
    int specular_bounce = 1, behind_glass = 0, bent = 0, straight = 1;
    float last_pdf = 0.0f;
    /* the ray cone (Akenine-Moller et al. 2019) for the textures' level: width at the origin, spread angle
     * (pathtrace.wgsl:4165-4167) */
    float cone_w = 0.0f, cone_a = 2.0f * pf(P, CAM_POS, 3) / lod_height(c);
    float clamp_v = pf(P, MISC, 0);
    c->mirror_path = 0;
    uint32_t last = pu(P, COUNTS, 2);
    for (uint32_t depth = 0; depth <= last; depth++) {
        c->g.depth = depth;
        or_hit hit;
        float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
        if (or_trace(&c->s->geo, of, df, OR_INF, OR_TRACE_CLOSEST, OR_WALK_EXACT, 0, &hit, NULL)) c->failed++;
        if (hit.tri == OR_NO_TRI) {
            V col = vmul(thr, sky_mis(c, d, (depth > 0u && !bent) ? last_pdf : 0.0f));
            if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
            radiance = vadd(radiance, col);
            break;
        }
        const float width = cone_w + cone_a * hit.t;
        Surface s = surface(c, o, d, &hit, width);
        if (depth == 0u) {
            if (s.card) {
                float mx = fmaxf(fmaxf(s.emissive.x, fmaxf(s.emissive.y, s.emissive.z)), 1e-6f);
                albedo = vds(s.emissive, mx);
            } else {
                albedo = s.base;
            }
            normal = s.n;
        }
        if (any_pos(s.emissive) && (s.front || !s.card)) {
            float wgt = 1.0f;
            if (depth > 0u && !specular_bounce && !behind_glass) wgt = emit_mis(c, &hit, d, last_pdf);
            V col = vs(vmul(thr, s.emissive), wgt);
            if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
            radiance = vadd(radiance, col);
        }
        if (s.card) break;
        if (depth == last) break;
        V v = vneg(d);
        float nov = fmaxf(dot(s.n, v), 1e-4f);
        Lobes lb = lobes(c, &s, nov);
        V dl = vmul(thr, direct_light(c, &s, &lb, v));
        if (depth > 0u && !straight) dl = clamp_lum(dl, clamp_v);
        radiance = vadd(radiance, dl);
        V l = sample_bsdf(c, &s, &lb, v);
        /* 4414-4447: through smooth glass a delta lobe, weight tint / its sampling share, no MIS */
        int through = glass_on(c) && c->sampled_trans && !lb.rough_t;
        float pdf = 0.0f;
        V wgt = vds(lb.tint, fmaxf(lb.p_trans, 1e-6f));
        if (!through) {
            pdf = pdf_bsdf(c, &s, &lb, v, l);
            wgt = vds(eval_bsdf(c, &s, &lb, v, l), fmaxf(pdf, 1e-20f));
        }
        if ((!through && pdf <= 0.0f) || all_nonpos(wgt)) break;
        thr = vmul(thr, wgt);
        if (through && dot(l, s.ng) < 0.0f) { /* 4449-4465: glass passed since the last scattering vertex */
            behind_glass = 1;
            bent = !bent;
        } else if (through) {
            last_pdf = 0.0f;
            specular_bounce = 1;
            behind_glass = 0;
            bent = 0;
        } else {
            last_pdf = pdf;
            specular_bounce = 0;
            behind_glass = 0;
            bent = 0;
        }
        straight = straight && through;
        cone_w = width;
        cone_a += 2.0f * s.roughness * s.roughness;
        c->mirror_path = (depth == 0u || c->mirror_path) && s.roughness < MIRROR_ROUGHNESS &&
                         (c->sampled_spec || c->sampled_trans);
        if (depth >= pu(P, COUNTS, 3)) {
            float q = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.95f);
            if (or_rng1(&c->g, or_rng_dim(&c->g, D_RR)) > q) break;
            thr = vds(thr, q);
        }
        o = offset_origin(s.p, dot(l, s.ng) < 0.0f ? vneg(s.ng) : s.ng);
        d = l;
    }
    if (non_finite(radiance)) radiance = v3(0, 0, 0);
    if (non_finite(albedo)) albedo = v3(0, 0, 0);
    if (non_finite(normal)) normal = v3(0, 1, 0);
    float lum = luminance(radiance);
    pix[0] += radiance.x, pix[1] += radiance.y, pix[2] += radiance.z, pix[3] += 1.0f;
    pix[4] += albedo.x, pix[5] += albedo.y, pix[6] += albedo.z, pix[7] += lum;
    pix[8] += normal.x, pix[9] += normal.y, pix[10] += normal.z, pix[11] += lum * lum;
}

int64_t or_render_sample(const or_render_scene *s, const uint32_t *params, uint32_t y0, uint32_t y1, float *accum,
                         int threads) {
    uint32_t w = params[4 * SIZE + 0];
    int64_t failed = 0, oob = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : failed, oob)
#endif
    for (int64_t y = y0; y < (int64_t)y1; y++) {
        Ctx c;
        memset(&c, 0, sizeof c);
        c.s = s;
        c.P = params;
        c.oob = &oob;
        for (uint32_t x = 0; x < w; x++) path(&c, x, (uint32_t)y, accum + 20 * ((uint64_t)y * w + x));
        failed += c.failed;
    }
    (void)threads;
    return failed + oob;
}

/* Test hooks: the emitter draw for given uniforms u[3] at (x, n) (*entry: its CDF entry, -1 none; returns its p), and
 * the pmf of emitter `ei` there (emitter_draw, emitter_pmf above), under `params`. */
static Ctx hook_ctx(const or_render_scene *s, const uint32_t *params, int64_t *oob) {
    Ctx c;
    memset(&c, 0, sizeof(c));
    c.s = s;
    c.P = params;
    c.oob = oob;
    return c;
}
float or_emitter_draw(const or_render_scene *s, const uint32_t *params, const float *x, const float *n,
                      const float *u, int64_t *entry) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, params, &oob);
    EPick pk = emitter_draw_u(&c, v3(x[0], x[1], x[2]), v3(n[0], n[1], n[2]), u);
    *entry = pk.e ? (int64_t)((pk.e - s->emitters) / 4) : -1;
    return oob ? -1.0f : pk.p;
}
float or_emitter_pmf(const or_render_scene *s, const uint32_t *params, const float *x, const float *n, uint32_t ei) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, params, &oob);
    Em em = emitter_tri_at(&c, ei);
    float p = emitter_pmf(&c, ei, &em, v3(x[0], x[1], x[2]), v3(n[0], n[1], n[2]));
    return oob ? -1.0f : p;
}

float or_glass_energy(const or_render_scene *s, float nov, float roughness, float ior, int leaving) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, NULL, &oob);
    float e = glass_energy(&c, nov, roughness, ior, leaving);
    return oob ? -1.0f : e;
}
float or_glass_share(const or_render_scene *s, const uint32_t *params, uint32_t material, float cos_d) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, params, &oob);
    float e = glass_share(&c, material_at(&c, material), cos_d, 1.0f);
    return oob ? -1.0f : e;
}

int or_tex(const or_render_scene *s, int32_t id, float u, float v, float lod, float *out) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, NULL, &oob);
    tex(&c, id, u, v, lod, out);
    return oob ? -1 : 0;
}
int or_env(const or_render_scene *s, const uint32_t *params, const float *d, const float *u, float *radiance,
           float *pdf, float *sample) {
    int64_t oob = 0;
    Ctx c = hook_ctx(s, params, &oob);
    const V dv = v3(d[0], d[1], d[2]);
    const V r = env_radiance(&c, dv);
    radiance[0] = r.x, radiance[1] = r.y, radiance[2] = r.z;
    *pdf = env_pdf(&c, dv);
    const V l = sample_env_u(&c, u);
    sample[0] = l.x, sample[1] = l.y, sample[2] = l.z;
    return oob ? -1 : 0;
}
