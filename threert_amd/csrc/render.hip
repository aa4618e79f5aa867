// The offline path tracer for gfx950: pathtrace.wgsl `main` (LIVE = false) for the features the workloads use,
// over the 4-wide quantised walk of walk.h.
//
// Restated (pathtrace.wgsl lines): camera_ray 4002-4024; main 4106-4557 (offline parts); surface 1320-1425 with
// tri_data 330-360, inst_vec / inst_sign / inst_normal 279-288, inst_attr 406-413, material_of 310-313; lobes
// 1642-1699 with energy_lookup 1492-1505, spec_terms 1521-1525; d_ggx / vis_smith_joint_approx / f_schlick
// 1707-1730; eval_bsdf_k 1733-1764; g1 / pdf_bsdf 1766-1789; light_mis 1797-1817; basis / sample_vndf 1819-1844;
// sample_bsdf 1848-1884; env_uv / env_radiance / env_pdf / lower_bound / sample_env 1889-1941 (a constant sky; a map
// with its CDF: kExt); power_heuristic, luminance,
// offset_origin 1943-1960; sky_mis / sun_disks 1963-1988; emit_mis 2009-2024; emitter_radiance /
// sample_emitter_ris / sample_emitter 2269-2393; env_light 2431-2442; direct_light 2444-2540; emitter_draw /
// emitter_pick_u / emitter_pick_range / lt_imp / emitter_pmf 463-618 (the light tree); the clearcoat; glass (kExt:
// the dielectric 1542-1640 with fresnel_dielectric, lambda_ggx, glass_energy, dielectric_t, pdf_dielectric; its
// parts of lobes, eval_bsdf_k, pdf_bsdf and sample_bsdf; shadow_tr / glass_tr / glass_at / sky_tr 846-977;
// light_origin 2153-2156; the path loop's `through`, behind_glass and bent 4156-4465); textures (kExt: the pages'
// trilinear read tex / tex_level 1107-1117 with srgb_to_linear 1102-1104, mesh_uv / mesh_density / mesh_lod
// 1119-1140, emission 1312-1317, the colour, roughness, metalness, emissive, normal and transmission maps of
// surface 1344-1425 with hit_tangent 366-394, glass_at's maps 950-956, emitter_radiance 2269-2283, the path loop's
// ray cone 4165-4167 and 4451-4452); local lights (kExt: local_light 2042-2104, local_lights 2158-2262 offline, with
// no light tiles or volumes). Not here: media, fog, procedural geometry, cutouts, glass among a shadow caster's
// sections, live-only paths (the upload and the launch refuse scenes and parameters that need them).
// normalize(v) = v / sqrt(v.v), inverseSqrt(x) = 1 / sqrt(x); sin / cos by Moshier's Cephes reduction and
// polynomials in fmaf; every other operation in the reference's order, no contraction: the same bits as the C
// restatement the tests hold this to.
//
// Launch: persistent waves; a wave takes one 8x8 tile at a time from a counter and runs `spp` samples of each of
// its pixels, adding them to the pixel's 80-byte record (records.wgsl Pixel) in sample order.
#include <hip/hip_runtime.h>

#include "render_kernels.h"
#include "walk.h"

namespace trt {
namespace {
using namespace walk;

constexpr float kPi = 3.14159265358979f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kFourPi = 4.0f * kPi;

struct V {
    float x, y, z;
};
__device__ __forceinline__ V v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V vmul(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V vs(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V vds(V a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V vneg(V a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V cross(V a, V b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(V a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V normalize(V a) { return vds(a, sqrtf(dot(a, a))); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ bool any_pos(V a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }
__device__ __forceinline__ bool all_nonpos(V a) { return a.x <= 0.0f && a.y <= 0.0f && a.z <= 0.0f; }
__device__ __forceinline__ float luminance(V c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
__device__ __forceinline__ bool non_finite(V v) {
    return (__float_as_uint(v.x) & 0x7fffffffu) >= 0x7f800000u || (__float_as_uint(v.y) & 0x7fffffffu) >= 0x7f800000u ||
           (__float_as_uint(v.z) & 0x7fffffffu) >= 0x7f800000u;
}
__device__ __forceinline__ V xyz(float4 f) { return {f.x, f.y, f.z}; }

struct M3 {
    V t, b, n;
};
__device__ __forceinline__ V mmul(const M3 &m, V v) { return vadd(vadd(vs(m.t, v.x), vs(m.b, v.y)), vs(m.n, v.z)); }
__device__ __forceinline__ V mtmul(const M3 &m, V v) { return {dot(m.t, v), dot(m.b, v), dot(m.n, v)}; }

// Moshier's Cephes sinf / cosf: octant reduction by pi/4 in three parts, minimax polynomials, fmaf evaluation.
__device__ __forceinline__ void sincos_c(float x, float &s, float &c) {
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    const float FOPI = 1.27323954473516f;
    float sign_s = 1.0f, ax = x;
    if (x < 0.0f) ax = -x, sign_s = -1.0f;
    uint32_t j = static_cast<uint32_t>(ax * FOPI);
    float y = static_cast<float>(j);
    if (j & 1u) j += 1u, y += 1.0f;
    j &= 7u;
    float sign_c = 1.0f;
    if (j > 3u) j -= 4u, sign_s = -sign_s, sign_c = -sign_c;
    if (j > 1u) sign_c = -sign_c;
    float z = fmaf(-y, DP1, ax);
    z = fmaf(-y, DP2, z);
    z = fmaf(-y, DP3, z);
    const float zz = z * z;
    float ps = fmaf(fmaf(-1.9515295891e-4f, zz, 8.3321608736e-3f), zz, -1.6666654611e-1f);
    ps = fmaf(ps * zz, z, z);
    float pc = fmaf(fmaf(2.443315711809948e-5f, zz, -1.388731625493765e-3f), zz, 4.166664568298827e-2f);
    pc = fmaf(pc * zz, zz, fmaf(-0.5f, zz, 1.0f));
    const bool swap = j == 1u || j == 2u;
    s = sign_s * (swap ? pc : ps);
    c = sign_c * (swap ? ps : pc);
}

// ---- sampler (sample.wgsl) ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t owen_rev(uint32_t n, uint32_t seed) {
    n ^= n * 0x3d20adeau;
    n += seed;
    n *= (seed >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}
__device__ __forceinline__ uint32_t hash_hp(uint32_t i) {
    i ^= i >> 16;
    i *= 0x21f0aaadu;
    i ^= i >> 15;
    i *= 0xd35a2d97u;
    i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}
// Direction numbers of Sobol dimensions 1-3 (Joe & Kuo), bit-reversed, the 16 the 16-bit shuffled index uses.
__constant__ uint32_t kSobol[3][16] = {
    {0x00000001u, 0x00000003u, 0x00000005u, 0x0000000fu, 0x00000011u, 0x00000033u, 0x00000055u, 0x000000ffu,
     0x00000101u, 0x00000303u, 0x00000505u, 0x00000f0fu, 0x00001111u, 0x00003333u, 0x00005555u, 0x0000ffffu},
    {0x00000001u, 0x00000003u, 0x00000006u, 0x00000009u, 0x00000017u, 0x0000003au, 0x00000071u, 0x000000a3u,
     0x00000116u, 0x00000339u, 0x00000677u, 0x000009aau, 0x00001601u, 0x00003903u, 0x00007706u, 0x0000aa09u},
    {0x00000001u, 0x00000003u, 0x00000004u, 0x0000000au, 0x0000001fu, 0x0000002eu, 0x00000045u, 0x000000c9u,
     0x0000011bu, 0x000002a4u, 0x0000079au, 0x00000b67u, 0x0000101eu, 0x0000302du, 0x00004041u, 0x0000a0c3u},
};
__device__ __forceinline__ float sobol_owen(uint32_t rev_index, uint32_t dim, uint32_t seed) {
    uint32_t r = 0;
    if (dim == 0) {
        r = __builtin_bitreverse32(rev_index);
    } else {
#pragma unroll
        for (int b = 0; b < 16; b++)
            if (rev_index & (0x80000000u >> b)) r ^= kSobol[dim - 1][b];
    }
    r = __builtin_bitreverse32(owen_rev(r, seed));
    return static_cast<float>(r) * static_cast<float>(1.0 / 4294967808.0);
}

struct Rng {
    uint32_t seed, index, depth;
    __device__ __forceinline__ uint32_t dim(uint32_t slot) const { return (0u << 28) ^ (depth << 20) ^ slot; }
    __device__ __forceinline__ uint32_t sd(uint32_t d) const { return seed ^ hash_hp(d); }
    __device__ __forceinline__ uint32_t idx(uint32_t salt) const {
        return owen_rev(__builtin_bitreverse32(index), salt) & 0xffff0000u;
    }
    __device__ __forceinline__ float r1(uint32_t d) const {
        const uint32_t s = sd(d);
        return sobol_owen(idx(s ^ 0xbff95bfeu), 0, s ^ 0x635c77bdu);
    }
    __device__ __forceinline__ void r2(uint32_t d, float *r) const {
        const uint32_t s = sd(d), i = idx(s ^ 0xf8ade99au);
        r[0] = sobol_owen(i, 0, s ^ 0xe0aaaf76u);
        r[1] = sobol_owen(i, 1, s ^ 0x94964d4eu);
    }
    __device__ __forceinline__ void r3(uint32_t d, float *r) const {
        const uint32_t s = sd(d), i = idx(s ^ 0xcaa726acu);
        r[0] = sobol_owen(i, 0, s ^ 0x9e78e391u);
        r[1] = sobol_owen(i, 1, s ^ 0x67c33241u);
        r[2] = sobol_owen(i, 2, s ^ 0x78c395c5u);
    }
    __device__ __forceinline__ void r4(uint32_t d, float *r) const {
        const uint32_t s = sd(d), i = idx(s ^ 0xc2c1a055u);
        r[0] = sobol_owen(i, 0, s ^ 0x39468210u);
        r[1] = sobol_owen(i, 1, s ^ 0xe9d8a845u);
        r[2] = sobol_owen(i, 2, s ^ 0x5f32b482u);
        r[3] = sobol_owen(i, 3, s ^ 0x1524cc56u);
    }
};
__device__ __forceinline__ uint32_t rng_sub(uint32_t d, uint32_t k) {
    return hash_hp(d ^ hash_hp(k + 0x9e3779b9u)) | 0x80000000u;
}

enum : uint32_t { D_CAMERA = 0, D_BSDF = 1, D_RR = 2, D_SUN = 4, D_LOCAL = 6, D_EMIT = 7, D_ENV = 9 };
constexpr uint32_t kEmitRisM = 6, kGuide = 65536, kNoEmitter = 0xffffffffu, kInstMirrored = 4, kEmitterCard = 1;
constexpr int kMatEmitterCard = 2;
constexpr int kMatUnlit = 4;  // pathtrace.wgsl:54 MAT_UNLIT (three.js MeshBasicMaterial)
constexpr float kMirrorRoughness = 0.3f;
constexpr uint32_t kObjectStride = 3;

// Cephes log2f and exp2f (S. L. Moshier): the mantissa in [sqrt(1/2), sqrt(2)) and a minimax polynomial in it; exp2
// by its integer part and a polynomial on [-1/2, 1/2], in fmaf as the C restatement has them.
__device__ float log2_c(float xx) {
    if (!(xx > 0.0f)) return xx == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (xx == __builtin_inff()) return xx;
    int e;
    float x = frexpf(xx, &e);
    if (x < 0.70710678118654752f) {
        e -= 1;
        x = x + x - 1.0f;
    } else {
        x = x - 1.0f;
    }
    const float z = x * x;
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
    return r + static_cast<float>(e);
}
__device__ float exp2_c(float xx) {
    if (xx != xx) return xx;
    if (xx > 127.0f) return __builtin_inff();
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
    return ldexpf(p, static_cast<int>(i0));
}
// atan2 and acos by Cephes atanf / asinf (S. L. Moshier): the argument reduced to [0, tan(pi/8)] or [0, 1/2] and a
// minimax polynomial there, in fmaf; the oracle has the same steps.
__device__ float atan_red(float x) {  // x >= 0
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
__device__ float atan2_c(float y, float x) {
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 1.5707963267948966f : (y < 0.0f ? -1.5707963267948966f : 0.0f);
    float w = 0.0f;
    if (x < 0.0f) w = y < 0.0f ? -3.141592653589793f : 3.141592653589793f;
    const float q = y / x;
    const float a = q < 0.0f ? -atan_red(-q) : atan_red(q);
    return w + a;
}
__device__ float asin_red(float a) {  // a in [0, 1]
    float x, z;
    bool flag = false;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = true;
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
__device__ float acos_c(float x) {
    if (!(x >= -1.0f && x <= 1.0f)) return __builtin_nanf("");
    if (x < -0.5f) return 3.141592653589793f - 2.0f * asin_red(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin_red(sqrtf(0.5f * (1.0f - x)));
    return 1.5707963267948966f - (x < 0.0f ? -asin_red(-x) : asin_red(x));
}
// WGSL's float to u32: saturating, NaN to 0
__device__ __forceinline__ uint32_t u32_sat(float x) {
    return x > 0.0f ? (x < 4294967040.0f ? static_cast<uint32_t>(x) : 0xffffff00u) : 0u;
}

// ---- the kernel's context ---------------------------------------------------------------------------------------
struct Ctx {
    const RenderArgs *a;
    Stack *st;
    uint32_t *flags;  // kFlag* bits: why a walk or a table read stopped and reported
    uint32_t *traced; // rays this thread traced: [0] closest, [1] shadow and any-hit
    Rng g;
    bool mirror_path;
    bool sampled_spec;
    bool sampled_trans;  // the last BSDF sample went through the surface (pathtrace.wgsl:449)
    V mis_x, mis_n;  // pathtrace.wgsl:438-439: the vertex direct_light last sampled from, for emit_mis
    Reads *cnt;      // what a counting launch's walks read (and cnt->shaded: hits shaded)

    // Every table index is checked against its table: one outside reads entry 0 and reports kFlagTable.
    __device__ __forceinline__ uint64_t ck(uint64_t i, uint64_t n, uint32_t table,
                                           const uint32_t *context = nullptr) const {
        if (i < n) return i;
        *flags |= kFlagTable;
        report_range(a->geo.guard_flag, table, i, n, context);
        return 0;
    }
    __device__ __forceinline__ float pf(int vec, int k) const { return __uint_as_float(a->params[4 * vec + k]); }
    __device__ __forceinline__ uint32_t pu(int vec, int k) const { return a->params[4 * vec + k]; }
    __device__ __forceinline__ const float4 *inst(uint32_t i) const {
        return a->instances + 10 * ck(i, a->geo.instance_count, 16);
    }
    __device__ __forceinline__ const float4 *material(uint32_t i) const {
        return a->materials + 18 * ck(i, a->material_count, 17);
    }
    __device__ __forceinline__ const float4 *light(uint32_t i) const { return a->lights + 5 * ck(i, a->light_count, 18); }
    __device__ __forceinline__ float4 emitter_at(uint32_t i) const { return a->emitters[ck(i, a->emitter_vec4s, 19)]; }
    __device__ __forceinline__ uint32_t row_word(uint32_t first, uint32_t section, uint32_t k) const {
        return a->rows[4 * ck(uint64_t(first) + section, a->row_count, 20) + k];
    }
    __device__ __forceinline__ float obj_power(uint32_t obj) const {
        return a->objects[ck(obj, a->object_count, 21) * kObjectStride].w;
    }
    template <bool kCount>
    __device__ HitRec trace(V o, V d, float tmax, uint32_t mode) const {
        traced[0] += mode == kTraceClosest ? 1u : 0u;
        traced[1] += mode == kTraceClosest ? 0u : 1u;
        return trace_wide<kCount>(a->geo, F3{o.x, o.y, o.z}, F3{d.x, d.y, d.z}, tmax, mode, *st, *flags, *cnt);
    }
};
enum { CAM_POS = 0, CAM_FWD = 1, CAM_RIGHT = 2, CAM_UP = 3, SIZE = 4, COUNTS = 5, ENV = 6, MISC = 7, SKY = 8,
       COUNTS2 = 9, LIGHT_LOD = 18, LTREE = 21, BACKGROUND = 24 };

struct TriAttr {
    float4 n0, n1, n2, uv, x;
};

__device__ TriAttr tri_attr(const Ctx &c, uint32_t ti, uint32_t inst, uint32_t part, V &e1, V &e2) {
    uint32_t base = 0;
    if (inst != kNone) base = __float_as_uint(c.inst(inst)[6].z);
    const uint32_t *ix = c.a->indices + 3 * c.ck(ti, c.a->tri_count, 22);
    const float4 *p0 = c.a->vertices + 3 * c.ck(uint64_t(ix[0]) + base, c.a->vertex_count, 23);
    const float4 *p1 = c.a->vertices + 3 * c.ck(uint64_t(ix[1]) + base, c.a->vertex_count, 23);
    const float4 *p2 = c.a->vertices + 3 * c.ck(uint64_t(ix[2]) + base, c.a->vertex_count, 23);
    const float4 a0 = p0[0], a1 = p1[0], a2 = p2[0], n0 = p0[1], n1 = p1[1], n2 = p2[1];
    e1 = vsub(xyz(a1), xyz(a0));
    e2 = vsub(xyz(a2), xyz(a0));
    TriAttr t;
    t.n0 = make_float4(n0.x, n0.y, n0.z, a0.w);
    t.n1 = make_float4(n1.x, n1.y, n1.z, n0.w);
    t.n2 = make_float4(n2.x, n2.y, n2.z, a1.w);
    t.uv = make_float4(n1.w, a2.w, n2.w, __uint_as_float(part));
    t.x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return t;
}
__device__ __forceinline__ V inst_vec(const float4 *r, V v) { return {dot(xyz(r[0]), v), dot(xyz(r[1]), v), dot(xyz(r[2]), v)}; }
__device__ __forceinline__ float4 inst_normal(const float4 *r, float4 n) {
    const V v = xyz(n);
    const V m = normalize(v3(dot(xyz(r[3]), v), dot(xyz(r[4]), v), dot(xyz(r[5]), v)));
    return make_float4(m.x, m.y, m.z, n.w);
}
__device__ __forceinline__ uint32_t material_of(const Ctx &c, const TriAttr &a, uint32_t inst) {
    if (inst == kNone) return 0;
    return c.row_word(__float_as_uint(c.inst(inst)[6].w), __float_as_uint(a.uv.w), 0) & 0xffffffu;
}

// ---- textures (pathtrace.wgsl:1102-1140; texpack.rs pages: RGBA8, 5 levels a layer) -------------------------------
constexpr uint32_t kTexMips = 5;
constexpr float kTexMaxLevel = 4.0f, kLodBias = 0.5f;
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ float srgb1(float c) {
    return c <= 0.04045f ? c / 12.92f : exp2_c(2.4f * log2_c((c + 0.055f) / 1.055f));
}
__device__ __forceinline__ V srgb_to_linear(V c) { return v3(srgb1(c.x), srgb1(c.y), srgb1(c.z)); }
// vcolor (269-275): the triangle's three sRGB 8-bit corner colours (unpack4x8unorm), decoded and interpolated
__device__ V vcolor(const Ctx &c, uint32_t t, float u, float v) {
    V k[3];
#pragma unroll
    for (uint32_t j = 0; j < 3; j++) {
        const uint32_t w = c.a->vcolors[c.ck(uint64_t(t) * 3u + j, c.a->vcolor_words, 36)];
        k[j] = srgb_to_linear(v3(static_cast<float>(w & 0xffu) / 255.0f, static_cast<float>((w >> 8) & 0xffu) / 255.0f,
                                 static_cast<float>((w >> 16) & 0xffu) / 255.0f));
    }
    return vadd(vadd(vs(k[0], 1.0f - u - v), vs(k[1], u)), vs(k[2], v));
}
// Level `level` of page layer `layer`, bilinear at page uv with the reference's sampler (repeat, linear): texel
// centres at (i + 0.5) / n, four texels weighed by the fractions, RGBA8 as unorm.
__device__ float4 tex_bilinear(const Ctx &c, uint32_t layer, uint32_t level, float pu, float pv) {
    const uint64_t side = c.a->tex_side;
    uint64_t layer_words = 0;
    for (uint32_t l = 0; l < kTexMips; l++) layer_words += (side >> l) * (side >> l);
    uint64_t off = c.ck(layer, c.a->tex_layers, 32) * layer_words;
    for (uint32_t l = 0; l < level; l++) off += (side >> l) * (side >> l);
    const uint32_t n = static_cast<uint32_t>(side >> level);
    const float x = pu * static_cast<float>(n) - 0.5f, y = pv * static_cast<float>(n) - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    const float fx = x - x0f, fy = y - y0f;
    if (!(x0f >= -1048576.0f && x0f <= 1048576.0f)) x0f = 0.0f;
    if (!(y0f >= -1048576.0f && y0f <= 1048576.0f)) y0f = 0.0f;
    const int32_t xi = static_cast<int32_t>(x0f), yi = static_cast<int32_t>(y0f), ni = static_cast<int32_t>(n);
    const uint32_t xa = static_cast<uint32_t>(((xi % ni) + ni) % ni), ya = static_cast<uint32_t>(((yi % ni) + ni) % ni);
    const uint32_t xb = xa + 1u == n ? 0u : xa + 1u, yb = ya + 1u == n ? 0u : ya + 1u;
    const uint32_t *pg = c.a->tex_pages + off;
    const uint32_t w00 = pg[uint64_t(ya) * n + xa], w10 = pg[uint64_t(ya) * n + xb];
    const uint32_t w01 = pg[uint64_t(yb) * n + xa], w11 = pg[uint64_t(yb) * n + xb];
    const float k = 1.0f / 255.0f;
    float o[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const float a00 = static_cast<float>((w00 >> (8 * j)) & 0xffu) * k, a10 = static_cast<float>((w10 >> (8 * j)) & 0xffu) * k;
        const float a01 = static_cast<float>((w01 >> (8 * j)) & 0xffu) * k, a11 = static_cast<float>((w11 >> (8 * j)) & 0xffu) * k;
        o[j] = mixf(mixf(a00, a10, fx), mixf(a01, a11, fx), fy);
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}
// tex_level (1107-1112): the texture's rect in its page (wrapped or clamped uv), trilinear between the two levels
// about `level` (clamped to the pages' 0..4)
__device__ float4 tex_level(const Ctx &c, int32_t id, float u, float v, float level) {
    const float4 a = c.a->tex_info[c.ck(uint64_t(uint32_t(id)) * 2u, c.a->tex_info_count, 33)];
    const float4 b = c.a->tex_info[c.ck(uint64_t(uint32_t(id)) * 2u + 1u, c.a->tex_info_count, 33)];
    float fu, fv;
    if (b.y > 0.5f) {
        fu = fractf(u), fv = fractf(v);
    } else {
        fu = clampf(u, 0.0f, 1.0f), fv = clampf(v, 0.0f, 1.0f);
    }
    const float pu = a.x + fu * a.z, pv = a.y + fv * a.w;
    const uint32_t layer = static_cast<uint32_t>(static_cast<int32_t>(b.x));
    const float lv = clampf(level, 0.0f, kTexMaxLevel);
    const uint32_t l0 = static_cast<uint32_t>(lv);
    const float fl = lv - static_cast<float>(l0);
    float4 r = tex_bilinear(c, layer, l0, pu, pv);
    if (fl > 0.0f) {
        const float4 r1 = tex_bilinear(c, layer, l0 + 1u, pu, pv);
        r = make_float4(mixf(r.x, r1.x, fl), mixf(r.y, r1.y, fl), mixf(r.z, r1.z, fl), mixf(r.w, r1.w, fl));
    }
    return r;
}
// tex (1114-1117): `lod` is log2 of the footprint in texture-UV units; the level adds the texture's own size
__device__ float4 tex(const Ctx &c, int32_t id, float u, float v, float lod) {
    const float4 b = c.a->tex_info[c.ck(uint64_t(uint32_t(id)) * 2u + 1u, c.a->tex_info_count, 33)];
    return tex_level(c, id, u, v, lod + 0.5f * log2_c(b.z * b.w));
}
__device__ __forceinline__ int32_t as_i32(float f) { return static_cast<int32_t>(__float_as_uint(f)); }
__device__ __forceinline__ float lod_height(const Ctx &c) {
    return static_cast<float>(c.pu(COUNTS2, 2) > 0u ? c.pu(COUNTS2, 2) : c.pu(SIZE, 1));
}
// kExt: the kernel has the code of the reference's optional features (its F_GLASS and F_TEX pipelines, local lights);
// which of them a job's scene has is read from its arguments, as the oracle reads it (no glass table: no glass; no
// texture info: no textures). Without kExt the code is compiled out.
template <bool kExt>
__device__ __forceinline__ bool glass_on(const Ctx &c) { return kExt && c.a->glass_energy != nullptr; }
template <bool kExt>
__device__ __forceinline__ bool tex_on(const Ctx &c) { return kExt && c.a->tex_info != nullptr; }
// Vertex colours: the scene's colours (the reference's F_VCOLOR), read for materials that ask (uv.w > 0).
template <bool kExt>
__device__ __forceinline__ bool vcolor_on(const Ctx &c, const float4 *m) {
    return kExt && c.a->vcolors != nullptr && m[5].w > 0.0f;
}
// An environment map: P.env.x, which the launch allows only for a scene that holds a map of P.env.yz's size.
template <bool kExt>
__device__ __forceinline__ bool env_on(const Ctx &c) { return kExt && c.pf(ENV, 0) >= 0.5f; }

struct Surface {
    V p, n, ng, base, emissive;
    float roughness, metallic, specular;
    float coat, coat_rough;     // pathtrace.wgsl:1420-1423 (0 without a coat)
    float transmission, ior;    // pathtrace.wgsl:1413-1419 (0 and 1.5 without glass)
    bool front, card;
};

// kCoat: the launch's scenes hold a clearcoat (the reference's F_COAT pipeline flag); without it the coat terms are
// compiled out, as the reference's shader is specialised. kExt likewise for glass and textures.
// mesh_uv (1119-1122): the hit's uv x the material's repeat + its offset
__device__ __forceinline__ void mesh_uv(const TriAttr &a, const float4 *m, float u, float v, float &uu, float &vv) {
    const float w = 1.0f - u - v;
    const float4 rep = m[5], off = m[6];
    uu = (a.n0.w * w + a.n2.w * u + a.uv.y * v) * rep.x + off.x;
    vv = (a.n1.w * w + a.uv.x * u + a.uv.z * v) * rep.y + off.y;
}
// mesh_density (1127-1137): (UV area x repeat, area) of the triangle as it stands; (repeat, 1) where either has none
__device__ __forceinline__ void mesh_density(V e1, V e2, const TriAttr &a, const float4 *m, float &d0, float &d1) {
    const float d1x = a.n2.w - a.n0.w, d1y = a.uv.x - a.n1.w;
    const float d2x = a.uv.y - a.n0.w, d2y = a.uv.z - a.n1.w;
    const float uv_area = fabsf(d1x * d2y - d2x * d1y);
    const float area = len(cross(e1, e2));
    const float4 rep = m[5];
    const float r = fmaxf(fabsf(rep.x * rep.y), 1e-12f);
    const bool sized = area > 2e-12f && uv_area > 2e-14f;
    d0 = sized ? uv_area * r : r;
    d1 = sized ? area : 1.0f;
}
// mesh_lod (1140)
__device__ __forceinline__ float mesh_lod(float d0, float d1, float fp2, float over) {
    return 0.5f * log2_c(fmaxf(d0 * fp2 / (d1 * over), 1e-30f));
}
// emission (1312-1317): the factor x the emissive map
template <bool kExt>
__device__ V emission(const Ctx &c, const float4 *m, float u, float v, float lod) {
    V em = xyz(m[1]);
    if (all_nonpos(em)) return em;
    if (tex_on<kExt>(c) && as_i32(m[4].x) >= 0) {
        const float4 tx = tex(c, as_i32(m[4].x), u, v, lod);
        em = vmul(em, srgb_to_linear(v3(tx.x, tx.y, tx.z)));
    }
    return em;
}
__device__ __forceinline__ V unit_or_zero(V v) {
    const float dd = dot(v, v);
    if (dd > 0.0f && isfinite(dd)) return normalize(v);
    return v3(0.0f, 0.0f, 0.0f);
}
// hit_tangent (366-394): the vertices' tangents (hand of the first vertex; 0: none, then one from the uv edges, or
// any vector at right angles to the normal where the uvs have no area), placed by the record. `a`, e1, e2: the
// triangle's attributes and edges in its mesh's space.
__device__ float4 hit_tangent(const Ctx &c, const TriAttr &a, V e1, V e2, uint32_t ti, uint32_t inst, float u, float v) {
    uint32_t base = 0;
    if (inst != kNone) base = __float_as_uint(c.inst(inst)[6].z);
    const uint32_t *ix = c.a->indices + 3 * c.ck(ti, c.a->tri_count, 22);
    const float4 t0 = c.a->vertices[3 * c.ck(uint64_t(ix[0]) + base, c.a->vertex_count, 23) + 2];
    const float4 t1 = c.a->vertices[3 * c.ck(uint64_t(ix[1]) + base, c.a->vertex_count, 23) + 2];
    const float4 t2 = c.a->vertices[3 * c.ck(uint64_t(ix[2]) + base, c.a->vertex_count, 23) + 2];
    V t = vadd(vadd(vs(xyz(t0), 1.0f - u - v), vs(xyz(t1), u)), vs(xyz(t2), v));
    float hand = t0.w;
    if (t0.w == 0.0f) {
        const float d1x = a.n2.w - a.n0.w, d1y = a.uv.x - a.n1.w;
        const float d2x = a.uv.y - a.n0.w, d2y = a.uv.z - a.n1.w;
        const float det = d1x * d2y - d2x * d1y;
        if (fabsf(det) < 1e-12f) {
            V any = v3(0.0f, 1.0f, 0.0f);
            if (fabsf(a.n0.x) < 0.9f) any = v3(1.0f, 0.0f, 0.0f);
            t = normalize(cross(xyz(a.n0), any));
            hand = 1.0f;
        } else {
            const float r = 1.0f / det;
            const V tu = vs(vsub(vs(e1, d2y), vs(e2, d1y)), r);
            const float l = len(tu);
            V tn = v3(0.0f, 0.0f, 0.0f);
            if (l > 0.0f && isfinite(l)) tn = vds(tu, l);
            const V bt = vs(vsub(vs(e2, d1x), vs(e1, d2x)), r);
            hand = dot(cross(xyz(a.n0), tn), bt) < 0.0f ? 1.0f : -1.0f;
            t = tn;
        }
    }
    if (inst != kNone) {
        const float4 *r = c.inst(inst);
        t = unit_or_zero(inst_vec(r, t));
        hand = hand * ((__float_as_uint(r[4].w) & kInstMirrored) ? -1.0f : 1.0f);
    }
    return make_float4(t.x, t.y, t.z, hand);
}

// `width`: the ray cone's footprint (m) at the hit, for the textures' level
template <bool kCoat, bool kExt = false>
__device__ Surface surface(const Ctx &c, V o, V d, const HitRec &h, float width = 0.0f) {
    V e1, e2;
    TriAttr a = tri_attr(c, h.tri, h.inst, h.part, e1, e2);
    const TriAttr a_mesh = a;  // the mesh's own attributes and edges (hit_tangent)
    const V e1_mesh = e1, e2_mesh = e2;
    float side = 1.0f;
    if (h.inst != kNone) {
        const float4 *r = c.inst(h.inst);
        e1 = inst_vec(r, e1);
        e2 = inst_vec(r, e2);
        a.n0 = inst_normal(r, a.n0);
        a.n1 = inst_normal(r, a.n1);
        a.n2 = inst_normal(r, a.n2);
        a.x.y = __uint_as_float(__float_as_uint(r[6].y) + __float_as_uint(a.x.y));
        side = (__float_as_uint(r[4].w) & kInstMirrored) ? -1.0f : 1.0f;
    }
    const float w = 1.0f - h.u - h.v;
    V ng = vs(normalize(cross(e1, e2)), side);
    V n = normalize(vadd(vadd(vs(xyz(a.n0), w), vs(xyz(a.n1), h.u)), vs(xyz(a.n2), h.v)));
    const float4 *m = c.material(material_of(c, a, h.inst));
    const bool front = dot(ng, d) < 0.0f;
    if (!front) ng = vneg(ng);
    if (dot(n, ng) < 0.0f) n = vneg(n);
    Surface s;
    s.p = vadd(o, vs(d, h.t));
    s.front = front;
    float uu = 0.0f, vv = 0.0f, lod = 0.0f;
    if (tex_on<kExt>(c)) {  // pathtrace.wgsl:1344-1348
        float d0, d1;
        mesh_uv(a, m, h.u, h.v, uu, vv);
        mesh_density(e1, e2, a, m, d0, d1);
        const float fp = fmaxf(width, 1e-7f);
        lod = mesh_lod(d0, d1, fp * fp, fmaxf(fabsf(dot(ng, d)), 1e-3f)) - kLodBias;
    }
    s.card = static_cast<int>(__float_as_uint(m[4].z)) == kMatEmitterCard;
    const float4 params = m[2];
    V base = xyz(m[0]);
    if (vcolor_on<kExt>(c, m)) base = vmul(base, vcolor(c, h.tri, h.u, h.v));  // 1356
    float rough = params.y, metal = params.x, spec = params.z;
    if (tex_on<kExt>(c)) {  // 1360-1364: the colour (sRGB), roughness (green) and metalness (blue) maps
        const float4 t0 = m[3];
        if (as_i32(t0.x) >= 0) {
            const float4 tx = tex(c, as_i32(t0.x), uu, vv, lod);
            base = vmul(base, srgb_to_linear(v3(tx.x, tx.y, tx.z)));
        }
        if (as_i32(t0.z) >= 0) rough *= tex(c, as_i32(t0.z), uu, vv, lod).y;
        if (as_i32(t0.w) >= 0) metal *= tex(c, as_i32(t0.w), uu, vv, lod).z;
    }
    s.emissive = vs(emission<kExt>(c, m, uu, vv, lod), c.obj_power(__float_as_uint(a.x.y)));
    // 1366-1372: three.js MeshBasicMaterial (tex1.z MAT_UNLIT): its colour to camera rays and mirror-like paths, lit
    // by nothing, lighting nothing
    if (kExt && as_i32(m[4].z) == kMatUnlit) {
        s.emissive = (c.g.depth == 0u || c.mirror_path) ? base : v3(0.0f, 0.0f, 0.0f);
        base = v3(0.0f, 0.0f, 0.0f);
        spec = 0.0f;
        metal = 0.0f;
    }
    // 1376-1387: the normal map where the surface is seen directly or in a mirror-like reflection
    if (tex_on<kExt>(c) && as_i32(m[3].y) >= 0 && (c.g.depth == 0u || c.mirror_path)) {
        const float4 tx = tex(c, as_i32(m[3].y), uu, vv, lod);
        const V tn = v3(tx.x * 2.0f - 1.0f, tx.y * 2.0f - 1.0f, tx.z * 2.0f - 1.0f);
        const float4 tw = hit_tangent(c, a_mesh, e1_mesh, e2_mesh, h.tri, h.inst, h.u, h.v);
        const V twv = xyz(tw);
        const V tp = vsub(twv, vs(n, dot(n, twv)));
        if (dot(tp, tp) > 1e-12f) {
            const V tt = normalize(tp);
            const V b = vs(cross(n, tt), tw.w);
            const V pn = vadd(vadd(vs(tt, tn.x * params.w), vs(b, tn.y * params.w)), vs(n, tn.z));
            if (dot(pn, pn) > 1e-12f && dot(pn, ng) > 0.0f) n = normalize(pn);
        }
    }
    s.base = base;
    s.roughness = clampf(rough, 0.02f, 1.0f);
    s.metallic = clampf(metal, 0.0f, 1.0f);
    s.specular = spec;
    s.ng = ng;
    s.n = n;
    s.coat = 0.0f;
    s.coat_rough = 1.0f;
    if (kCoat) {
        const float4 layer = m[7];
        s.coat = clampf(layer.z, 0.0f, 1.0f);
        s.coat_rough = clampf(layer.w, 0.02f, 1.0f);
    }
    s.ior = 1.5f;
    s.transmission = 0.0f;
    if (glass_on<kExt>(c)) {  // layer.x, layer.y; the transmission map (uv.z: its texture + 1)
        const float4 layer = m[7];
        s.transmission = clampf(layer.x, 0.0f, 1.0f);
        if (tex_on<kExt>(c) && m[5].z > 0.5f) s.transmission *= tex(c, static_cast<int32_t>(m[5].z) - 1, uu, vv, lod).x;
        s.ior = fmaxf(layer.y, 1.0001f);
    }
    return s;
}

// ---- BSDF ---------------------------------------------------------------------------------------------------------
struct Lobes {
    V diffuse, f0, f90, w, e_spec, kd;
    float a2, p_spec, p_coat, p_trans;
    float coat, coat_alpha, e_coat;
    // the dielectric (pathtrace.wgsl:1535-1538): its share x compensation, tint, n(far) / n(view), its directional
    // reflection albedo, and whether its transmission has a density (rough) or is a delta lobe
    V tint;
    float t, eta, e_glass;
    bool rough_t;
};

// ---- the dielectric (pathtrace.wgsl:1542-1640) --------------------------------------------------------------------
constexpr float kGlassDelta = 0.05f;
constexpr uint32_t kGlassEMu = 16, kGlassERough = 16, kGlassEIor = 8;
constexpr float kGlassEZmax = 0.70710678f;
// Unpolarised Fresnel reflectance, c the cosine on the incident side, eta = n(far) / n(incident); 1 at total
// internal reflection (Walter et al. 2007, eq. 22)
__device__ float fresnel_dielectric(float c, float eta) {
    const float g2 = eta * eta - 1.0f + c * c;
    if (g2 <= 0.0f) return 1.0f;
    const float g = sqrtf(g2);
    const float a = (g - c) / (g + c);
    const float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * a * a * (1.0f + b * b);
}
__device__ __forceinline__ float lambda_ggx(float a2, float c) {
    const float c2 = c * c;
    return 0.5f * (sqrtf(1.0f + a2 * (1.0f - c2) / fmaxf(c2, 1e-8f)) - 1.0f);
}
// R + T of the rough interface (the glass energy table: 16 sqrt(cos) x 16 roughness x 8 sqrt((ior - 1) / (ior + 1))
// nodes, entering then leaving, four a float4), trilinear
__device__ __forceinline__ float glass_e_at(const Ctx &c, uint32_t i) {
    const float4 v = c.a->glass_energy[c.ck(i / 4u, 1024, 31)];
    const uint32_t k = i & 3u;
    return k == 0u ? v.x : (k == 1u ? v.y : (k == 2u ? v.z : v.w));
}
__device__ __forceinline__ float glass_e_row(const Ctx &c, uint32_t row, float x) {
    const uint32_t x0 = static_cast<uint32_t>(x);
    const uint32_t x1 = min(x0 + 1u, kGlassEMu - 1u);
    return mixf(glass_e_at(c, row * kGlassEMu + x0), glass_e_at(c, row * kGlassEMu + x1), x - static_cast<float>(x0));
}
__device__ float glass_energy(const Ctx &c, float nov, float roughness, float ior, bool leaving) {
    const float x = clampf(sqrtf(nov), 0.0f, 1.0f) * static_cast<float>(kGlassEMu - 1u);
    const float y = clampf(roughness, 0.0f, 1.0f) * static_cast<float>(kGlassERough - 1u);
    const float z = clampf(sqrtf(fmaxf(ior - 1.0f, 0.0f) / (ior + 1.0f)) / kGlassEZmax, 0.0f, 1.0f) *
                    static_cast<float>(kGlassEIor - 1u);
    const uint32_t y0 = static_cast<uint32_t>(y), y1 = min(y0 + 1u, kGlassERough - 1u);
    const uint32_t side = leaving ? kGlassEIor : 0u, zi = static_cast<uint32_t>(z);
    const uint32_t z0 = side + zi, z1 = side + min(zi + 1u, kGlassEIor - 1u);
    const float fy = y - static_cast<float>(y0);
    const float e0 = mixf(glass_e_row(c, z0 * kGlassERough + y0, x), glass_e_row(c, z0 * kGlassERough + y1, x), fy);
    const float e1 = mixf(glass_e_row(c, z1 * kGlassERough + y0, x), glass_e_row(c, z1 * kGlassERough + y1, x), fy);
    return mixf(e0, e1, z - floorf(z));
}

__device__ __forceinline__ void energy_node(const Ctx &c, uint32_t i, float &e0, float &e1) {
    const float4 v = c.a->energy[c.ck(i >> 1, 512, 24)];
    e0 = (i & 1u) ? v.z : v.x;
    e1 = (i & 1u) ? v.w : v.y;
}
__device__ void energy_lookup(const Ctx &c, float roughness, float nov, float *e) {
    const uint32_t N = 32;
    const float x = sqrtf(clampf(nov, 0.0f, 1.0f)) * static_cast<float>(N - 1);
    const float y = clampf(roughness, 0.0f, 1.0f) * static_cast<float>(N - 1);
    const uint32_t x0 = min(static_cast<uint32_t>(x), N - 2), y0 = min(static_cast<uint32_t>(y), N - 2);
    const uint32_t i = y0 * N + x0;
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    float a0, a1, b0, b1, c0, c1, d0, d1;
    energy_node(c, i, a0, a1);
    energy_node(c, i + 1, b0, b1);
    energy_node(c, i + N, c0, c1);
    energy_node(c, i + N + 1, d0, d1);
    e[0] = mixf(mixf(a0, b0, fx), mixf(c0, d0, fx), fy);
    e[1] = mixf(mixf(a1, b1, fx), mixf(c1, d1, fx), fy);
}

template <bool kExt = false>
__device__ Lobes lobes(const Ctx &c, const Surface &s, float nov) {
    Lobes l;
    const float t = kExt ? s.transmission * (1.0f - s.metallic) : 0.0f;
    l.diffuse = vs(vs(s.base, 1.0f - s.metallic), 1.0f - t);
    float f0 = 0.08f * s.specular;
    if (kExt && t > 0.0f) {  // glass: F0 from the index of refraction (pathtrace.wgsl:1654-1658)
        const float ri = (s.ior - 1.0f) / (s.ior + 1.0f);
        f0 = ri * ri;
    }
    l.f0 = v3(mixf(f0, s.base.x, s.metallic), mixf(f0, s.base.y, s.metallic), mixf(f0, s.base.z, s.metallic));
    float e[2];
    energy_lookup(c, s.roughness, nov, e);
    const float comp = c.pf(MISC, 2);
    const float k503 = 50.0f / 3.0f;
    const float f90 = clampf(dot(l.f0, v3(k503, k503, k503)), 0.0f, 1.0f);
    const float inv = comp * (1.0f / fmaxf(e[0] + e[1], 1e-3f) - 1.0f);
    l.w = v3(1.0f + l.f0.x * inv, 1.0f + l.f0.y * inv, 1.0f + l.f0.z * inv);
    l.e_spec = vmul(vadd(vs(l.f0, e[0]), v3(f90 * e[1], f90 * e[1], f90 * e[1])), l.w);
    l.f90 = v3(f90, f90, f90);
    l.kd = vsub(v3(1.0f, 1.0f, 1.0f), l.e_spec);
    if (kExt && t > 0.0f) {  // 1669-1673: the share that is not the dielectric keeps the specular lobe
        l.w = vs(l.w, 1.0f - t);
        l.e_spec = vs(l.e_spec, 1.0f - t);
    }
    const float r = s.roughness;
    l.a2 = r * r * r * r;
    const V lw = v3(0.3f, 0.6f, 0.1f);
    const float ws = dot(l.e_spec, lw);
    const float wd = dot(vmul(l.diffuse, l.kd), lw);
    l.p_spec = clampf(ws / fmaxf(ws + wd, 1e-6f), 0.05f, 1.0f);
    if (wd <= 0.0f) l.p_spec = 1.0f;
    // pathtrace.wgsl:1675-1683: the coat's directional albedo, a single-scattering lobe of F0 0.04, F90 1
    l.coat = 0.0f;
    l.coat_alpha = 1.0f;
    l.e_coat = 0.0f;
    if (s.coat > 0.0f) {
        float ec[2];
        l.coat = s.coat;
        energy_lookup(c, s.coat_rough, nov, ec);
        l.e_coat = s.coat * (0.04f * ec[0] + ec[1]);
        l.coat_alpha = s.coat_rough * s.coat_rough;
    }
    const float e_coat = l.e_coat, under = 1.0f - e_coat;
    float wt = 0.0f;
    l.t = 0.0f;
    l.eta = 0.0f;
    l.rough_t = false;
    l.tint = v3(0.0f, 0.0f, 0.0f);
    l.e_glass = 0.0f;
    if (kExt && t > 0.0f) {  // 1688-1697
        l.eta = s.front ? s.ior : 1.0f / s.ior;
        l.rough_t = s.roughness > kGlassDelta && fabsf(s.ior - 1.0f) > 1e-3f;
        float gc = 1.0f;
        if (l.rough_t) gc = mixf(1.0f, 1.0f / glass_energy(c, nov, s.roughness, s.ior, !s.front), comp);
        l.t = t * under * gc;
        l.tint = vs(s.base, l.t);
        l.e_glass = t * fresnel_dielectric(nov, l.eta);
        wt = under * (l.e_glass + (t - l.e_glass) * dot(s.base, lw));
    }
    const float total = (ws + wd) * under + e_coat + wt;
    l.p_coat = e_coat / fmaxf(total, 1e-6f);
    l.p_trans = wt / fmaxf(total, 1e-6f);
    return l;
}

__device__ __forceinline__ float d_ggx(float a2, float noh) {
    const float d = (noh * a2 - noh) * noh + 1.0f;
    return a2 / (kPi * d * d);
}
__device__ __forceinline__ float vis_smith(float a2, float nov, float nol) {
    return 0.5f / fmaxf(mixf(2.0f * nol * nov, nol + nov, sqrtf(a2)), 1e-7f);
}
__device__ __forceinline__ V f_schlick(V f0, V f90, float voh) {
    const float c = 1.0f - voh, c2 = c * c, k = c2 * c2 * c;
    return vadd(f0, vs(vsub(f90, f0), k));
}
__device__ __forceinline__ float g1(float a2, float nov) { return 2.0f * nov / (nov + sqrtf(a2 + (1.0f - a2) * nov * nov)); }
// The dielectric's transmission toward l (the far side) x |n.l|, Walter 2007 eq. 21 (pathtrace.wgsl:1610-1622)
__device__ V dielectric_t(const Surface &s, const Lobes &lb, V v, V l, float nov, float nol) {
    if (!lb.rough_t) return v3(0.0f, 0.0f, 0.0f);
    V h = normalize(vadd(v, vs(l, lb.eta)));
    if (dot(h, s.n) < 0.0f) h = vneg(h);
    const float voh = dot(v, h), loh = dot(l, h);
    if (voh <= 0.0f || loh >= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    const float den = voh + lb.eta * loh;
    const float g2 = 1.0f / (1.0f + lambda_ggx(lb.a2, nov) + lambda_ggx(lb.a2, nol));
    const float f = (1.0f - fresnel_dielectric(voh, lb.eta)) * d_ggx(lb.a2, fmaxf(dot(s.n, h), 0.0f)) * g2 * voh *
                    lb.eta * lb.eta * -loh / (nov * fmaxf(den * den, 1e-12f));
    return vs(lb.tint, f);
}
// The density of l under the dielectric's sampling x its share of the samples (1626-1640)
__device__ float pdf_dielectric(const Surface &s, const Lobes &lb, V v, V l) {
    const float nov = fmaxf(dot(s.n, v), 1e-4f);
    if (dot(s.n, l) > 0.0f) {
        const V h = normalize(vadd(v, l));
        const float voh = fmaxf(dot(v, h), 1e-6f);
        const float f = fresnel_dielectric(lb.rough_t ? voh : nov, lb.eta);
        return lb.p_trans * f * g1(lb.a2, nov) * d_ggx(lb.a2, fmaxf(dot(s.n, h), 0.0f)) / (4.0f * nov);
    }
    if (!lb.rough_t) return 0.0f;
    V h = normalize(vadd(v, vs(l, lb.eta)));
    if (dot(h, s.n) < 0.0f) h = vneg(h);
    const float voh = dot(v, h), loh = dot(l, h);
    if (voh <= 0.0f || loh >= 0.0f) return 0.0f;
    const float den = voh + lb.eta * loh;
    return lb.p_trans * (1.0f - fresnel_dielectric(voh, lb.eta)) * g1(lb.a2, nov) * voh *
           d_ggx(lb.a2, fmaxf(dot(s.n, h), 0.0f)) * lb.eta * lb.eta * -loh / (nov * fmaxf(den * den, 1e-12f));
}
// Whether lights on both sides of the surface reach it: rough glass (1606)
template <bool kExt>
__device__ __forceinline__ bool two_sided(const Lobes &lb) { return kExt && lb.rough_t; }

template <bool kExt = false>
__device__ V eval_bsdf_k(const Surface &s, const Lobes &lb, V v, V l, float k) {
    const float nol = dot(s.n, l);
    const float nov = fmaxf(dot(s.n, v), 1e-4f);
    if (two_sided<kExt>(lb) && nol < 0.0f && dot(s.ng, l) < 0.0f) return dielectric_t(s, lb, v, l, nov, -nol);
    if (nol <= 0.0f || dot(s.ng, l) <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
    const V h = normalize(vadd(v, l));
    const float noh = fmaxf(dot(s.n, h), 0.0f);
    const float voh = fmaxf(dot(v, h), 0.0f);
    const float dv = d_ggx(lb.a2, noh) * vis_smith(lb.a2, nov, nol);
    V spec = vs(vmul(vs(f_schlick(lb.f0, lb.f90, voh), dv), lb.w), k);
    // the dielectric's reflection (1747-1748): lb.t holds the coat's share already, so it joins after the coat
    float glass = 0.0f;
    if (kExt && lb.t > 0.0f) glass = lb.t * fresnel_dielectric(voh, lb.eta) * dv * k;
    V diff = vmul(vs(lb.diffuse, kInvPi), lb.kd);
    if (lb.coat > 0.0f) {  // pathtrace.wgsl:1750-1755
        const float ca2 = lb.coat_alpha * lb.coat_alpha;
        const float cf = f_schlick(v3(0.04f, 0.04f, 0.04f), v3(1.0f, 1.0f, 1.0f), voh).x;
        const float coat = d_ggx(ca2, noh) * vis_smith(ca2, nov, nol) * cf * lb.coat * k;
        diff = vs(diff, 1.0f - lb.e_coat);
        spec = vadd(vs(spec, 1.0f - lb.e_coat), v3(coat, coat, coat));
    }
    if (kExt) spec = vadd(spec, v3(glass, glass, glass));
    return vs(vadd(diff, spec), nol);
}
template <bool kExt = false>
__device__ __forceinline__ V eval_bsdf(const Surface &s, const Lobes &lb, V v, V l) {
    return eval_bsdf_k<kExt>(s, lb, v, l, 1.0f);
}
template <bool kExt = false>
__device__ float pdf_bsdf(const Surface &s, const Lobes &lb, V v, V l) {
    const float nol = dot(s.n, l);
    float pg = 0.0f;
    if (kExt && lb.p_trans > 0.0f) pg = pdf_dielectric(s, lb, v, l);
    if (nol <= 0.0f) return pg;
    const float nov = fmaxf(dot(s.n, v), 1e-4f);
    const V h = normalize(vadd(v, l));
    const float noh = fmaxf(dot(s.n, h), 0.0f);
    const float voh = fmaxf(dot(v, h), 1e-6f);
    const float dv = g1(lb.a2, nov) * voh * d_ggx(lb.a2, noh) / nov;
    const float ps = dv / (4.0f * voh);
    const float pd = nol / kPi;
    float pc = 0.0f;
    if (lb.coat > 0.0f) {  // pathtrace.wgsl:1783-1786
        const float ca2 = lb.coat_alpha * lb.coat_alpha;
        pc = g1(ca2, nov) * d_ggx(ca2, noh) / (4.0f * nov);
    }
    return pg + lb.p_coat * pc + (1.0f - lb.p_coat - lb.p_trans) * (lb.p_spec * ps + (1.0f - lb.p_spec) * pd);
}
__device__ __forceinline__ float power_heuristic(float a, float b) {
    const float a2 = a * a;
    return a2 / fmaxf(a2 + b * b, 1e-20f);
}
template <bool kExt = false>
__device__ __forceinline__ float light_mis(float pdf_l, const Surface &s, const Lobes &lb, V v, V l) {
    return power_heuristic(pdf_l, pdf_bsdf<kExt>(s, lb, v, l));
}
__device__ __forceinline__ M3 basis(V n) {
    const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sgn + n.z);
    const float b = n.x * n.y * a;
    M3 m;
    m.t = v3(1.0f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x);
    m.b = v3(b, sgn + n.y * n.y * a, -n.y);
    m.n = n;
    return m;
}
__device__ V sample_vndf(V v, const M3 &tbn, float alpha, float u1, float u2) {
    const V vl = mtmul(tbn, v);
    const V vh = normalize(v3(alpha * vl.x, alpha * vl.y, vl.z));
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V t1 = lensq > 0.0f ? vds(v3(-vh.y, vh.x, 0.0f), sqrtf(lensq)) : v3(1.0f, 0.0f, 0.0f);
    const V t2 = cross(vh, t1);
    const float r = sqrtf(u1);
    const float phi = kTwoPi * u2;
    float sp, cp;
    sincos_c(phi, sp, cp);
    const float p1 = r * cp;
    float p2 = r * sp;
    const float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * sqrtf(1.0f - p1 * p1) + s * p2;
    const V nh = vadd(vadd(vs(t1, p1), vs(t2, p2)), vs(vh, sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2))));
    const V hl = normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(0.0f, nh.z)));
    return mmul(tbn, hl);
}
__device__ __forceinline__ V reflect(V e1, V e2) { return vsub(e1, vs(e2, 2.0f * dot(e2, e1))); }
// WGSL refract: the zero vector at total internal reflection
__device__ __forceinline__ V refract(V e1, V e2, float eta) {
    const float d = dot(e2, e1);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3(0.0f, 0.0f, 0.0f);
    return vsub(vs(e1, eta), vs(e2, eta * d + sqrtf(k)));
}

template <bool kExt = false>
__device__ V sample_bsdf(Ctx &c, const Surface &s, const Lobes &lb, V v) {
    const M3 tbn = basis(s.n);
    float r[4];
    c.g.r4(c.g.dim(D_BSDF), r);
    c.sampled_trans = false;
    if (r[0] < lb.p_coat) {  // pathtrace.wgsl:1853-1857: the coat's lobe (p_coat is 0 without one)
        c.sampled_spec = true;
        return reflect(vneg(v), sample_vndf(v, tbn, lb.coat_alpha, r[2], r[3]));
    }
    if (kExt && r[0] < lb.p_coat + lb.p_trans) {
        // 1858-1874, the dielectric: a visible microfacet normal (smooth glass refracts through the shading normal),
        // reflected or refracted by the Fresnel reflectance there; a reflection under the surface or a refraction
        // above it is no sample (the zero vector)
        const V h = sample_vndf(v, tbn, s.roughness * s.roughness, r[2], r[3]);
        const V m = lb.rough_t ? h : s.n;
        const V tr = refract(vneg(v), m, 1.0f / lb.eta);
        c.sampled_spec = r[0] - lb.p_coat < fresnel_dielectric(fmaxf(dot(v, m), 1e-4f), lb.eta) * lb.p_trans ||
                         (tr.x == 0.0f && tr.y == 0.0f && tr.z == 0.0f);
        if (c.sampled_spec) {
            const V rl = reflect(vneg(v), h);
            return dot(rl, s.n) <= 0.0f ? v3(0.0f, 0.0f, 0.0f) : rl;
        }
        c.sampled_trans = true;
        return dot(tr, s.n) >= 0.0f ? v3(0.0f, 0.0f, 0.0f) : tr;
    }
    c.sampled_spec = r[1] < lb.p_spec;
    if (c.sampled_spec) {
        const V h = sample_vndf(v, tbn, s.roughness * s.roughness, r[2], r[3]);
        return reflect(vneg(v), h);
    }
    const float rr = sqrtf(r[2]);
    const float phi = kTwoPi * r[3];
    float sp, cp;
    sincos_c(phi, sp, cp);
    return mmul(tbn, v3(rr * cp, rr * sp, sqrtf(fmaxf(0.0f, 1.0f - rr * rr))));
}

// ---- light --------------------------------------------------------------------------------------------------------
__device__ __forceinline__ V offset_origin(V p, V n) {
    const float m = fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
    return vadd(p, vs(n, 1e-4f * m));
}
template <bool kCount>
__device__ __forceinline__ bool occluded(const Ctx &c, V o, V d, float tmax, uint32_t mode) {
    return c.trace<kCount>(o, d, tmax, mode).tri != kNone;
}
__device__ __forceinline__ V sky(const Ctx &c) { return v3(c.pf(SKY, 0), c.pf(SKY, 1), c.pf(SKY, 2)); }
// ---- the environment (pathtrace.wgsl:1889-1941): a constant sky, or (kExt) an equirectangular map with its CDF --------
__device__ __forceinline__ float half_f(uint32_t h) {  // IEEE binary16 to binary32, exactly
    const uint32_t s = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1fu) return __uint_as_float(s | 0x7f800000u | (m << 13));
    if (e == 0u) return __uint_as_float(__float_as_uint(static_cast<float>(m) * 5.9604644775390625e-8f) | s);
    return __uint_as_float(s | ((e + 112u) << 23) | (m << 13));
}
// env_uv (1889-1893)
__device__ __forceinline__ void env_uv(V d, float &u, float &v) {
    u = atan2_c(d.x, -d.z) / kTwoPi + 0.5f;
    v = acos_c(clampf(d.y, -1.0f, 1.0f)) / kPi;
}
__device__ __forceinline__ float env_cdf_at(const Ctx &c, uint64_t i) { return c.a->env_cdf[c.ck(i, c.a->env_cdf_count, 35)]; }
// env_radiance (1895-1898): the map read at level 0 as the reference's sampler defines it (repeat in u and v,
// linear), with float weights
template <bool kExt>
__device__ V env_radiance(const Ctx &c, V d) {
    if (!env_on<kExt>(c)) return sky(c);
    float u, v;
    env_uv(d, u, v);
    const uint32_t w = c.a->env_w, h = c.a->env_h;
    const float x = u * static_cast<float>(w) - 0.5f, y = v * static_cast<float>(h) - 0.5f;
    float x0f = floorf(x), y0f = floorf(y);
    const float fx = x - x0f, fy = y - y0f;
    if (!(x0f >= -1048576.0f && x0f <= 1048576.0f)) x0f = 0.0f;
    if (!(y0f >= -1048576.0f && y0f <= 1048576.0f)) y0f = 0.0f;
    const int32_t xi = static_cast<int32_t>(x0f), yi = static_cast<int32_t>(y0f);
    const int32_t wi = static_cast<int32_t>(w), hi = static_cast<int32_t>(h);
    const uint32_t xa = static_cast<uint32_t>(((xi % wi) + wi) % wi), ya = static_cast<uint32_t>(((yi % hi) + hi) % hi);
    const uint32_t xb = xa + 1u == w ? 0u : xa + 1u, yb = ya + 1u == h ? 0u : ya + 1u;
    const uint64_t n = uint64_t(w) * h;
    const uint2 t00 = c.a->env[c.ck(uint64_t(ya) * w + xa, n, 34)], t10 = c.a->env[c.ck(uint64_t(ya) * w + xb, n, 34)];
    const uint2 t01 = c.a->env[c.ck(uint64_t(yb) * w + xa, n, 34)], t11 = c.a->env[c.ck(uint64_t(yb) * w + xb, n, 34)];
    const float r = mixf(mixf(half_f(t00.x & 0xffffu), half_f(t10.x & 0xffffu), fx),
                         mixf(half_f(t01.x & 0xffffu), half_f(t11.x & 0xffffu), fx), fy);
    const float g = mixf(mixf(half_f(t00.x >> 16), half_f(t10.x >> 16), fx),
                         mixf(half_f(t01.x >> 16), half_f(t11.x >> 16), fx), fy);
    const float b = mixf(mixf(half_f(t00.y & 0xffffu), half_f(t10.y & 0xffffu), fx),
                         mixf(half_f(t01.y & 0xffffu), half_f(t11.y & 0xffffu), fx), fy);
    return v3(r, g, b);
}
// env_pdf (1900-1913): the rows' and the row's CDF steps of the texel d falls in, per solid angle
template <bool kExt>
__device__ float env_pdf(const Ctx &c, V d) {
    if (!env_on<kExt>(c)) return 1.0f / kFourPi;
    const uint32_t w = c.a->env_w, h = c.a->env_h;
    float u, v;
    env_uv(d, u, v);
    const uint32_t x = min(u32_sat(u * static_cast<float>(w)), w - 1u), y = min(u32_sat(v * static_cast<float>(h)), h - 1u);
    const float row_p = env_cdf_at(c, y) - (y > 0u ? env_cdf_at(c, y - 1u) : 0.0f);
    const uint64_t base = uint64_t(h) + uint64_t(y) * w;
    const float col_p = env_cdf_at(c, base + x) - (x > 0u ? env_cdf_at(c, base + x - 1u) : 0.0f);
    float sin_t, cos_t;
    sincos_c((static_cast<float>(y) + 0.5f) / static_cast<float>(h) * kPi, sin_t, cos_t);
    return row_p * col_p * static_cast<float>(w) * static_cast<float>(h) / (2.0f * kPi * kPi * fmaxf(sin_t, 1e-4f));
}
// lower_bound (1915-1923)
__device__ uint32_t env_lower_bound(const Ctx &c, uint64_t start, uint32_t count, float x) {
    uint32_t lo = 0u, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (env_cdf_at(c, start + mid) < x)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return min(lo, count - 1u);
}
__device__ __forceinline__ bool env_sampled(const Ctx &c) {
    return c.pu(COUNTS2, 1) != 0u && (c.pf(ENV, 0) > 0.5f || luminance(sky(c)) > 0.0f);
}
// sample_env (1925-1941): the uniform sphere for a constant sky; the map's row by its CDF, the texel in it by the
// row's, a uniform point of the texel in (u, v)
template <bool kExt>
__device__ V sample_env(Ctx &c, uint32_t rd) {
    float u[4];
    c.g.r4(rd, u);
    if (!env_on<kExt>(c)) {
        const float z = 1.0f - 2.0f * u[0];
        const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = kTwoPi * u[1];
        float sp, cp;
        sincos_c(phi, sp, cp);
        return v3(r * cp, z, r * sp);
    }
    const uint32_t w = c.a->env_w, h = c.a->env_h;
    const uint32_t y = env_lower_bound(c, 0u, h, u[0]);
    const uint32_t x = env_lower_bound(c, uint64_t(h) + uint64_t(y) * w, w, u[1]);
    const float uu = (static_cast<float>(x) + u[2]) / static_cast<float>(w);
    const float vv = (static_cast<float>(y) + u[3]) / static_cast<float>(h);
    const float theta = vv * kPi, phi = (uu - 0.5f) * 2.0f * kPi;
    float st, ct, sp, cp;
    sincos_c(theta, st, ct);
    sincos_c(phi, sp, cp);
    return v3(st * sp, ct, -st * cp);
}
// sky_mis (1963-1974); with kExt, a camera ray (or a path bent straight through glass) that leaves the scene sees
// three.js's scene.background where P.background.w says there is one (the reference's F_BACKGROUND)
template <bool kExt>
__device__ V sky_mis(const Ctx &c, V d, float pdf) {
    if (kExt && pdf <= 0.0f && c.pf(BACKGROUND, 3) > 0.5f)
        return v3(c.pf(BACKGROUND, 0), c.pf(BACKGROUND, 1), c.pf(BACKGROUND, 2));
    float wgt = 1.0f;
    if (pdf > 0.0f && c.pu(COUNTS2, 1) != 0u) wgt = power_heuristic(pdf, env_pdf<kExt>(c, d));
    V col = vs(env_radiance<kExt>(c, d), wgt);
    if (pdf > 0.0f) {
        V sum = v3(0.0f, 0.0f, 0.0f);
        for (uint32_t i = 0; i < c.pu(COUNTS2, 0); i++) {
            const float4 *li = c.light(i);
            const float4 dir = li[1];
            if (dot(d, vneg(xyz(dir))) >= dir.w) {
                const float omega = kTwoPi * (1.0f - dir.w);
                sum = vadd(sum, vs(vds(xyz(li[2]), omega), power_heuristic(pdf, 1.0f / omega)));
            }
        }
        col = vadd(col, sum);
    }
    return col;
}

// ---- glass along shadow rays (pathtrace.wgsl:846-977) -------------------------------------------------------------
// The share of a ray a glass material lets through at |cos| cs: transmission x (1 - the specular lobe's directional
// albedo there, spec_terms 1521-1525 at the glass's F0); 0: not glass (glass_at 955-969)
__device__ float glass_share(const Ctx &c, const float4 *m, float cs, float tr_map) {
    const float4 params = m[2], layer = m[7];
    float tr = clampf(layer.x, 0.0f, 1.0f) * (1.0f - clampf(params.x, 0.0f, 1.0f));
    if (tr <= 0.0f) return 0.0f;
    tr = tr * tr_map;  // the transmission map's red, 1 without one
    const float ior = fmaxf(layer.y, 1.0001f);
    const float ri = (ior - 1.0f) / (ior + 1.0f);
    const float f0 = ri * ri;
    float e[2];
    energy_lookup(c, clampf(params.y, 0.02f, 1.0f), cs, e);
    const float k503 = 50.0f / 3.0f;
    const float f90 = clampf(dot(v3(f0, f0, f0), v3(k503, k503, k503)), 0.0f, 1.0f);
    const float w = 1.0f + f0 * (c.pf(MISC, 2) * (1.0f / fmaxf(e[0] + e[1], 1e-3f) - 1.0f));
    return tr * (1.0f - (f0 * e[0] + f90 * e[1]) * w);
}
// glass_tr (885-906): past every opaque caster, what the glass a shadow ray crosses lets through (up to 8 surfaces,
// the ray kept straight: no refraction); toward the sky an odd number of glass surfaces lets nothing through
template <bool kCount, bool kExt>
__device__ V glass_tr(const Ctx &c, V o, V d, float tmax, bool to_sky) {
    float left = tmax;
    V tr = v3(1.0f, 1.0f, 1.0f);
    bool odd = false;
    float t0 = 0.0f;  // the distance from the ray's start (the maps' level, as a camera ray's footprint there)
    for (uint32_t k = 0; k < 8u; k++) {
        const HitRec h = c.trace<kCount>(o, d, left, kTraceClosest);
        if (h.tri == kNone) break;
        t0 += h.t;
        // glass_at (930-970): the hit's geometric normal, its material's share at this angle and its tint
        V e1, e2;
        TriAttr a = tri_attr(c, h.tri, h.inst, h.part, e1, e2);
        float side = 1.0f;
        if (h.inst != kNone) {
            const float4 *r = c.inst(h.inst);
            e1 = inst_vec(r, e1);
            e2 = inst_vec(r, e2);
            side = (__float_as_uint(r[4].w) & kInstMirrored) ? -1.0f : 1.0f;
        }
        const float4 *m = c.material(material_of(c, a, h.inst));
        const V ng = vs(normalize(cross(e1, e2)), side);
        V tint = xyz(m[0]);
        if (vcolor_on<kExt>(c, m)) tint = vmul(tint, vcolor(c, h.tri, h.u, h.v));  // 949
        float tr_map = 1.0f;
        if (tex_on<kExt>(c) && clampf(m[7].x, 0.0f, 1.0f) * (1.0f - clampf(m[2].x, 0.0f, 1.0f)) > 0.0f &&
            (as_i32(m[3].x) >= 0 || m[5].z > 0.5f)) {  // 950-956
            float uu, vv, d0, d1;
            mesh_uv(a, m, h.u, h.v, uu, vv);
            const float fp = fmaxf(t0 * 2.0f * c.pf(CAM_POS, 3) / lod_height(c), 1e-7f);
            mesh_density(e1, e2, a, m, d0, d1);
            const float lod = mesh_lod(d0, d1, fp * fp, 1.0f) - kLodBias;
            if (as_i32(m[3].x) >= 0) {
                const float4 tx = tex(c, as_i32(m[3].x), uu, vv, lod);
                tint = vmul(tint, srgb_to_linear(v3(tx.x, tx.y, tx.z)));
            }
            if (m[5].z > 0.5f) tr_map = tex(c, static_cast<int32_t>(m[5].z) - 1, uu, vv, lod).x;
        }
        const float share = glass_share(c, m, fabsf(dot(ng, d)), tr_map);
        if (share > 0.0f) {
            tr = vmul(tr, vs(tint, share));
            odd = !odd;
        }
        o = offset_origin(vadd(o, vs(d, h.t)), dot(ng, d) < 0.0f ? vneg(ng) : ng);
        left -= h.t;
        if (left <= 0.0f || all_nonpos(tr)) break;
    }
    if (to_sky && odd) return v3(0.0f, 0.0f, 0.0f);
    return tr;
}
// shadow_tr (846-855): 0 past a shadow caster; else what the glass lets through (objects that cast no shadow are
// glass); sky_tr (975-977) with to_sky
template <bool kCount, bool kExt>
__device__ V shadow_tr(const Ctx &c, V o, V d, float tmax, bool to_sky) {
    if (c.trace<kCount>(o, d, tmax, kTraceShadow).tri != kNone) return v3(0.0f, 0.0f, 0.0f);
    return glass_tr<kCount, kExt>(c, o, d, tmax, to_sky);
}
// light_origin (2153-2156): a shadow ray's origin on the side l leaves the surface from
__device__ __forceinline__ V light_origin(const Surface &s, V o, V l, bool glass) {
    if (glass && dot(l, s.ng) < 0.0f) return offset_origin(s.p, vneg(s.ng));
    return o;
}

__device__ __forceinline__ uint32_t emitter_count(const Ctx &c) { return c.pu(COUNTS, 1); }
__device__ __forceinline__ float4 emitter_totals(const Ctx &c) {
    return c.emitter_at(emitter_count(c) + kGuide / 4u + 1u);
}
__device__ __forceinline__ bool emitters_on(const Ctx &c) { return emitter_count(c) != 0u && emitter_totals(c).x > 0.0f; }
__device__ __forceinline__ uint32_t emitter_guide(const Ctx &c, uint32_t k) {
    const float4 g = c.emitter_at(emitter_count(c) + k / 4u);
    const uint32_t j = k % 4u;
    return __float_as_uint(j == 0 ? g.x : (j == 1 ? g.y : (j == 2 ? g.z : g.w)));
}
__device__ float4 emitter_pick_u(const Ctx &c, float u) {
    const uint32_t k = min(static_cast<uint32_t>(u * static_cast<float>(kGuide)), kGuide - 1u);
    uint32_t lo = emitter_guide(c, k), hi = emitter_guide(c, k + 1u);
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (c.emitter_at(mid).y < u)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return c.emitter_at(min(lo, emitter_count(c) - 1u));
}

// ---- the light tree (pathtrace.wgsl:489-618) -----------------------------------------------------------------
constexpr uint32_t kLtLeaf = 0x80000000u, kLtPathBits = 26;
__device__ __forceinline__ bool lt_on(const Ctx &c) { return c.pf(LTREE, 3) > 0.5f; }
__device__ __forceinline__ float4 lt_node(const Ctx &c, uint32_t i) {
    return c.a->light_tree[c.ck(i, c.a->light_tree_count, 30)];
}
// unpack2x16float's halves: IEEE binary16 to binary32, exactly
__device__ __forceinline__ float half_bits(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1fu) return __uint_as_float(sign | 0x7f800000u | (m << 13));
    if (e == 0u) return __uint_as_float(__float_as_uint(static_cast<float>(m) * 5.9604644775390625e-8f) | sign);
    return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}
// lt_imp (504-524): power x the best cosine any point of the node's sphere has at x / max(d^2, r^2) (offline: no
// distance cut)
__device__ float lt_imp(float4 nd, V x, V n) {
    const uint32_t a = __float_as_uint(nd.x), b = __float_as_uint(nd.y);
    const V dv = vsub(v3(half_bits(a & 0xffffu), half_bits(a >> 16), half_bits(b & 0xffffu)), x);
    const float r = half_bits(b >> 16);
    const float d2 = dot(dv, dv);
    const float r2 = r * r;
    float cb = 1.0f;
    if (d2 > r2 && dot(n, n) > 0.0f) {
        const float inv = 1.0f / sqrtf(d2);
        const float cos_t = dot(n, dv) * inv;
        const float sin_b = r * inv;
        const float cos_b = sqrtf(fmaxf(1.0f - sin_b * sin_b, 0.0f));
        if (cos_t < cos_b) cb = fmaxf(cos_t * cos_b + sqrtf(fmaxf(1.0f - cos_t * cos_t, 0.0f)) * sin_b, 0.0f);
    }
    return nd.z * cb / fmaxf(d2, r2);
}
// emitter_pick_range (527-540): an entry by weight within [lo, hi)
__device__ float4 emitter_pick_range(const Ctx &c, uint32_t lo, uint32_t hi, float r) {
    const float p0 = lo > 0u ? c.emitter_at(max(lo, 1u) - 1u).y : 0.0f;
    const float p1 = c.emitter_at(hi - 1u).y;
    const float u = fminf(p0 + r * (p1 - p0), p1);
    const uint32_t k = min(static_cast<uint32_t>(u * static_cast<float>(kGuide)), kGuide - 1u);
    uint32_t l = emitter_guide(c, k), h = emitter_guide(c, k + 1u);
    while (l < h) {
        const uint32_t mid = (l + h) / 2u;
        if (c.emitter_at(mid).y < u)
            l = mid + 1u;
        else
            h = mid;
    }
    return c.emitter_at(min(max(l, lo), hi - 1u));
}
struct EPick {
    float4 e;
    float p, w;
};
// emitter_draw (547-585): without the tree the CDF; with it the tail by its share, else the tree descended by
// importance at (x, n) to a leaf and one of its entries by weight
// kTree: the launch has a scene whose Params ask for the light tree; without it the tree's code is compiled out (the
// reference's pipelines are specialised to the features a scene uses), which keeps its registers from the others.
template <bool kTree>
__device__ EPick emitter_draw(const Ctx &c, V x, V n, uint32_t rd) {
    float r3[3];
    c.g.r3(rd, r3);
    const float4 tot = emitter_totals(c);
    const EPick none{make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, 0.0f};
    if (!(tot.x > 0.0f)) return none;
    if (!kTree || !lt_on(c)) {
        const float4 e = emitter_pick_u(c, r3[0]);
        return {e, e.z / tot.x, e.z};
    }
    if (r3[0] < tot.y) {
        const float4 e = emitter_pick_range(c, c.pu(LTREE, 1), emitter_count(c), r3[1]);
        return {e, e.z / tot.x, e.z};
    }
    float p = 1.0f - tot.y;
    float u = r3[1];
    uint32_t w = __float_as_uint(lt_node(c, 0u).w);
    for (uint32_t k = 0u; k <= kLtPathBits && (w & kLtLeaf) == 0u; k++) {
        const float4 l = lt_node(c, w), r = lt_node(c, w + 1u);
        const float il = lt_imp(l, x, n), ir = lt_imp(r, x, n);
        if (!(il + ir > 0.0f)) return none;
        const float pl = il / (il + ir);
        if (u < pl) {
            w = __float_as_uint(l.w);
            p *= pl;
            u = u / pl;
        } else {
            w = __float_as_uint(r.w);
            p *= 1.0f - pl;
            u = (u - pl) / (1.0f - pl);
        }
        u = fminf(fmaxf(u, 0.0f), 0.99999994f);
    }
    if ((w & kLtLeaf) == 0u) return none;
    const uint32_t lo = w & 0x1fffffffu, hi = lo + ((w >> 29) & 3u) + 1u;
    float wsum = 0.0f;
    for (uint32_t j = lo; j < hi; j++) wsum += c.emitter_at(j).z;
    float t = r3[2] * wsum;
    float4 e = c.emitter_at(hi - 1u);
    for (uint32_t j = lo; j < hi; j++) {
        const float4 cc = c.emitter_at(j);
        if (t < cc.z) {
            e = cc;
            break;
        }
        t -= cc.z;
    }
    return {e, p * e.z / wsum, e.z};
}

struct Em {
    V v0, e1, e2;
    uint32_t tri, inst, code, part, flags;
};
// emitter_pmf (590-618): the probability emitter_draw at (x, n) picks emitter `ei`
template <bool kTree>
__device__ float emitter_pmf(const Ctx &c, uint32_t ei, const Em &em, V x, V n) {
    const float4 tot = emitter_totals(c);
    if (!(tot.x > 0.0f)) return 0.0f;
    const uint32_t code = em.code & ~0x80000000u;
    if (!kTree || (code >> kLtPathBits) == 0u) return c.emitter_at(code).z / tot.x;
    const uint32_t depth = (code >> kLtPathBits) - 1u;
    float p = 1.0f - tot.y;
    uint32_t w = __float_as_uint(lt_node(c, 0u).w);
    for (uint32_t k = 0u; k < depth; k++) {
        const float4 l = lt_node(c, w), r = lt_node(c, w + 1u);
        const float il = lt_imp(l, x, n), ir = lt_imp(r, x, n);
        if (!(il + ir > 0.0f)) return 0.0f;
        const bool right = ((code >> k) & 1u) == 1u;
        p *= (right ? ir : il) / (il + ir);
        w = __float_as_uint(right ? r.w : l.w);
    }
    const uint32_t lo = w & 0x1fffffffu, hi = lo + ((w >> 29) & 3u) + 1u;
    float wsum = 0.0f, we = 0.0f;
    for (uint32_t j = lo; j < hi; j++) {
        const float4 cc = c.emitter_at(j);
        wsum += cc.z;
        if (__float_as_uint(cc.x) == ei) we = cc.z;
    }
    return p * we / wsum;
}
// site: which caller reads the table (25 RIS candidate, 27 RIS pick, 28 single draw, 29 a hit's own emitter)
__device__ __forceinline__ Em emitter_tri_at(const Ctx &c, uint32_t i, uint32_t site,
                                             const uint32_t *context = nullptr) {
    const float4 *r = c.a->emitter_tris + 4 * c.ck(i, c.a->emitter_tri_count, site, context);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    return {xyz(r0), xyz(r1), xyz(r2), __float_as_uint(r0.w), __float_as_uint(r1.w), __float_as_uint(r2.w),
            __float_as_uint(r3.x), __float_as_uint(r3.y)};
}
// emitter_radiance (2269-2283): the emission at the sample (b1, b2), its map read at a 5 cm footprint
template <bool kExt>
__device__ V emitter_radiance(const Ctx &c, const Em &em, float b1, float b2) {
    V e1, e2;
    TriAttr a = tri_attr(c, em.tri, em.inst, em.part, e1, e2);
    if (em.inst != kNone) a.x.y = __uint_as_float(__float_as_uint(c.inst(em.inst)[6].y) + __float_as_uint(a.x.y));
    const float4 *m = c.material(material_of(c, a, em.inst));
    float uu = 0.0f, vv = 0.0f, lod = 0.0f;
    if (tex_on<kExt>(c)) {
        float d0, d1;
        mesh_uv(a, m, b1, b2, uu, vv);
        mesh_density(em.e1, em.e2, a, m, d0, d1);
        lod = mesh_lod(d0, d1, 0.0025f, 1.0f);
    }
    return vs(emission<kExt>(c, m, uu, vv, lod), c.obj_power(__float_as_uint(a.x.y)));
}

struct ES {
    V le, l;
    float dist, pdf_w, w;
};

template <bool kTree, bool kExt = false>
__device__ ES sample_emitter_ris(Ctx &c, V x, V n, uint32_t m, const Surface &sf, const Lobes &lb, V v, uint32_t rd) {
    ES r{v3(0.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f), 0.0f, 0.0f, 0.0f};
    float wsum = 0.0f, pick_b1 = 0.0f, pick_b2 = 0.0f, pick_ph = 0.0f, pick_pdf = 0.0f;
    uint32_t pick_i = 0;
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t dk = rng_sub(rd, k);
        const EPick pk = emitter_draw<kTree>(c, x, n, dk);
        const float4 e = pk.e;
        const float pk_p = pk.p, pk_w = pk.w;
        float bb[3];
        c.g.r3(rng_sub(dk, 1u), bb);
        float b1 = bb[0], b2 = bb[1];
        if (pk_p <= 0.0f) continue;
        if (b1 + b2 > 1.0f) b1 = 1.0f - b1, b2 = 1.0f - b2;
        const uint32_t ei = __float_as_uint(e.x);
        const Em em = emitter_tri_at(c, ei, 25);
        const V lp = vadd(vadd(em.v0, vs(em.e1, b1)), vs(em.e2, b2));
        const V cr = cross(em.e1, em.e2);
        const float a2 = len(cr);
        if (!(a2 > 0.0f)) continue;
        const V ln = vds(cr, a2);
        const V dl = vsub(lp, x);
        const float d2 = fmaxf(dot(dl, dl), 1e-8f);
        const V l = vs(dl, 1.0f / sqrtf(d2));
        float g = fabsf(dot(ln, l)) * (dot(n, n) == 0.0f ? 1.0f : fmaxf(dot(n, l), 0.0f)) / d2;
        if (m > 1u) g = fabsf(dot(ln, l)) * luminance(eval_bsdf<kExt>(sf, lb, v, l)) / d2;
        const float ph = pk_w / (e.w * 1.0f) * g;
        const float pdf_a = pk_p / (0.5f * a2);
        const float w = ph / pdf_a;
        if (!(w > 0.0f)) continue;
        wsum += w;
        if (bb[2] * wsum < w) pick_b1 = b1, pick_b2 = b2, pick_ph = ph, pick_i = ei, pick_pdf = pdf_a;
    }
    if (pick_ph <= 0.0f) return r;
    const Em em = emitter_tri_at(c, pick_i, 27);
    const V lp = vadd(vadd(em.v0, vs(em.e1, pick_b1)), vs(em.e2, pick_b2));
    const V ln = normalize(cross(em.e1, em.e2));
    const V dl = vsub(lp, x);
    const float d2 = dot(dl, dl);
    const float dist = sqrtf(d2);
    const V l = vds(dl, dist);
    float cos_l = dot(ln, vneg(l));
    if (!(em.flags & kEmitterCard)) cos_l = fabsf(cos_l);
    if (cos_l <= 0.0f) return r;
    r.le = emitter_radiance<kExt>(c, em, pick_b1, pick_b2);
    r.l = l;
    r.dist = dist;
    r.pdf_w = pick_pdf * d2 / cos_l;
    r.w = (wsum / static_cast<float>(m)) / pick_ph * cos_l / d2;
    return r;
}

template <bool kTree, bool kExt = false>
__device__ ES sample_emitter(Ctx &c, V x, V n, uint32_t rd) {
    ES r{v3(0.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f), 0.0f, 0.0f, 0.0f};
    const EPick pk = emitter_draw<kTree>(c, x, n, rd);
    const float4 e = pk.e;
    const float pk_p = pk.p;
    if (pk_p <= 0.0f) return r;
    const Em em = emitter_tri_at(c, __float_as_uint(e.x), 28);
    float bb[2];
    c.g.r2(rng_sub(rd, 1u), bb);
    float b1 = bb[0], b2 = bb[1];
    if (b1 + b2 > 1.0f) b1 = 1.0f - b1, b2 = 1.0f - b2;
    const V lp = vadd(vadd(em.v0, vs(em.e1, b1)), vs(em.e2, b2));
    const V cr = cross(em.e1, em.e2);
    const float area = 0.5f * len(cr);
    const V ln = vds(cr, 2.0f * area);
    const V dl = vsub(lp, x);
    const float d2 = dot(dl, dl);
    const float dist = sqrtf(d2);
    const V l = vds(dl, dist);
    float cos_l = dot(ln, vneg(l));
    if (!(em.flags & kEmitterCard)) cos_l = fabsf(cos_l);
    if (cos_l <= 0.0f) return r;
    r.le = emitter_radiance<kExt>(c, em, b1, b2);
    r.l = l;
    r.dist = dist;
    r.pdf_w = pk_p / area * d2 / cos_l;
    r.w = 1.0f / r.pdf_w;
    return r;
}

// Where the wavefront's direct light goes: contribution j as if visible, and (when it waits on one) the shadow ray
// that decides it in ray slot j + 1. The shade kernel of the next bounce adds the visible ones, in j's order.
struct WfShadows {
    const WfArgs *w;
    uint32_t p;        // the path
    uint32_t n = 0;    // contributions so far
    uint32_t present = 0, traced = 0;  // bit j: contribution j is there; it counts only if shadow ray j + 1 misses
    __device__ __forceinline__ void put(V col, bool trace, V o, V d, float tmax, uint32_t mode) {
        if (n >= kWfShadow) return;  // the host sends jobs with more than kWfMaxSuns suns to the megakernel
        const uint64_t st = w->stride;
        w->contrib[n * st + p] = make_float4(col.x, col.y, col.z, 0.0f);
        if (trace) {
            float4 *r = w->rays + 2 * ((n + 1) * st + p);
            r[0] = make_float4(o.x, o.y, o.z, tmax);
            r[1] = make_float4(d.x, d.y, d.z, __uint_as_float(mode));
            traced |= 1u << n;
        }
        present |= 1u << n;
        n++;
    }
};

// ---- local lights (pathtrace.wgsl:2039-2262) ---------------------------------------------------------------------
constexpr uint32_t kRisExact = 8, kRisM = 8;
struct LS {
    V c, l, sl;
    float dist;
};
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
    const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// local_light (2042-2104): a point, spot or rect light seen from x: its intensity x attenuation (unshadowed), the
// direction to its centre (or to the rect's sampled point), and a shadow ray toward a point of it
__device__ LS local_light(const float4 *li, V x, float u0, float u1) {
    LS r;
    const float4 p4 = li[0], d4 = li[1], c4 = li[2], e4 = li[3];
    const V pos = xyz(p4);
    const float range = c4.w;
    if (static_cast<uint32_t>(p4.w) == 3u) {  // the rect: a uniform point on it, luminance x area x cos_light / d^2
        const V n = xyz(d4), ux = xyz(e4);
        const V uy = vs(normalize(cross(n, ux)), d4.w);
        const V lp = vadd(vadd(pos, vs(ux, 2.0f * u0 - 1.0f)), vs(uy, 2.0f * u1 - 1.0f));
        const V dl = vsub(lp, x);
        const float d2 = fmaxf(dot(dl, dl), 1e-8f);
        const float dist = sqrtf(d2);
        const V l = vds(dl, dist);
        const float cos_l = dot(n, vneg(l));
        float k = 0.0f;
        if (cos_l > 0.0f) {
            const float area = 4.0f * len(ux) * d4.w;
            const float bd = li[4].z;
            k = area * cos_l / d2 * smoothstepf(bd - 0.04f, bd + 0.04f, cos_l);
            if (range > 0.0f) {
                const float q = d2 / (range * range);
                const float win = clampf(1.0f - q * q, 0.0f, 1.0f);
                k *= win * win;
            }
        }
        r.c = vs(xyz(c4), k), r.l = l, r.sl = l, r.dist = dist;
        return r;
    }
    V lp = pos;
    const float radius = d4.w;
    if (radius > 0.0f) {
        const float z = 1.0f - 2.0f * u0;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float sp, cp;
        sincos_c(kTwoPi * u1, sp, cp);
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
    if (static_cast<uint32_t>(p4.w) == 2u) {  // the cone: glTF KHR_lights_punctual's angular attenuation
        const float cone = clampf((dot(vneg(l), xyz(d4)) - e4.x) * e4.y, 0.0f, 1.0f);
        atten *= cone * cone;
    }
    const V ls = vsub(lp, x);
    const float dist = fmaxf(len(ls), 1e-6f);
    r.c = vs(xyz(c4), atten), r.l = l, r.sl = vds(ls, dist), r.dist = dist;
    return r;
}
// local_lights (2158-2262, offline: no light tiles, no volume): where the point is seen directly, or with few lights,
// every unshadowed light and, when few, every shadowed one exactly (one shadow ray each); otherwise one of kRisM (3
// past the first bounce) uniform candidates resampled by unshadowed contribution, one shadow ray. (The reference skips
// whole groups of 16 lights a point is outside the reach of: the same sum, every light of such a group being
// rejected by its own range test here.)
template <bool kCount, bool kExt>
__device__ V local_lights(Ctx &c, const Surface &s, const Lobes &lb, V v, V o, uint32_t rd) {
    const uint32_t nd = c.pu(COUNTS2, 0), nall = c.pu(COUNTS, 0);
    const uint32_t nl = nall - nd;
    V sum = v3(0.0f, 0.0f, 0.0f);
    const V x = s.p;
    const bool seen = c.g.depth == 0u || (c.g.depth == 1u && c.mirror_path);
    const bool glass = glass_on<kExt>(c);
    uint32_t first = nd, count = nl;
    if (seen || nl <= kRisExact) {
        const uint32_t ns = c.pu(COUNTS2, 3), n_sh = nall - ns;
        const uint32_t end = (n_sh <= kRisExact || nl <= kRisExact) ? nall : ns;
        for (uint32_t i = nd; i < end; i++) {
            const float4 *li = c.light(i);
            const float4 p4 = li[0], d4 = li[1], c4 = li[2], e4 = li[3];
            if (all_nonpos(xyz(c4))) continue;
            const V dc = vsub(xyz(p4), x);
            const float reach = c4.w + (static_cast<uint32_t>(p4.w) == 3u ? len(xyz(e4)) + d4.w : d4.w);
            if (c4.w > 0.0f && dot(dc, dc) > reach * reach) continue;
            if (static_cast<uint32_t>(p4.w) != 3u && dot(dc, s.ng) < -d4.w && !two_sided<kExt>(lb)) continue;
            float u[2];
            c.g.r2(rng_sub(rd, i), u);
            const LS ls = local_light(li, x, u[0], u[1]);
            if (all_nonpos(ls.c)) continue;
            const V f = eval_bsdf_k<kExt>(s, lb, v, ls.l, li[4].y);
            if (any_pos(f)) {
                const V lo = light_origin(s, o, ls.l, glass);
                const V vis = e4.w >= 0.5f ? (glass ? shadow_tr<kCount, kExt>(c, lo, ls.sl, ls.dist * 0.999f, false)
                                                    : (occluded<kCount>(c, lo, ls.sl, ls.dist * 0.999f, kTraceShadow)
                                                           ? v3(0.0f, 0.0f, 0.0f)
                                                           : v3(1.0f, 1.0f, 1.0f)))
                                           : v3(1.0f, 1.0f, 1.0f);
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
    const uint32_t m_cand = seen ? kRisM : 3u;
    float wsum = 0.0f, pick_ph = 0.0f;
    uint32_t pick = 0;
    LS pick_ls{v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f), 0.0f};
    for (uint32_t k = 0; k < m_cand; k++) {
        float r4[4];
        c.g.r4(rng_sub(rd ^ 0x80000000u, k), r4);
        const uint32_t idx = first + min(static_cast<uint32_t>(r4[0] * static_cast<float>(count)), count - 1u);
        const LS ls = local_light(c.light(idx), x, r4[1], r4[2]);
        float ph = luminance(ls.c);
        if (two_sided<kExt>(lb))
            ph *= fabsf(dot(s.n, ls.l));
        else
            ph *= fmaxf(dot(s.n, ls.l), 0.0f) + 0.05f * fmaxf(dot(s.ng, ls.l), 0.0f);
        if (ph <= 0.0f) continue;
        const float w = ph * 1.0f;
        wsum += w;
        if (r4[3] * wsum < w) pick = idx, pick_ph = ph, pick_ls = ls;
    }
    if (pick_ph > 0.0f) {
        const float4 *lk = c.light(pick);
        const V cc = vmul(pick_ls.c, eval_bsdf_k<kExt>(s, lb, v, pick_ls.l, lk[4].y));
        if (any_pos(cc)) {
            const V lo = light_origin(s, o, pick_ls.l, glass);
            const float tmax = pick_ls.dist * 0.999f;
            const V vis = lk[3].w >= 0.5f ? (glass ? shadow_tr<kCount, kExt>(c, lo, pick_ls.sl, tmax, false)
                                                   : (occluded<kCount>(c, lo, pick_ls.sl, tmax, kTraceShadow)
                                                          ? v3(0.0f, 0.0f, 0.0f)
                                                          : v3(1.0f, 1.0f, 1.0f)))
                                          : v3(1.0f, 1.0f, 1.0f);
            if (any_pos(vis))
                sum = vadd(sum, vs(glass ? vmul(cc, vis) : cc,
                                   wsum * static_cast<float>(count) / (static_cast<float>(m_cand) * pick_ph)));
        }
    }
    return sum;
}

// pathtrace.wgsl direct_light: the sum of what the suns, one emitter sample and one environment sample send toward v.
// kDefer (the wavefront): each contribution goes to `q` with its shadow ray, untraced, and the sum returned is 0.
// kExt (the megakernel alone): with glass, shadow rays leave from the far side toward lights behind rough glass, the
// suns' and the sky's pass glass at its share (shadow_tr), and rough glass samples emitters from both sides.
template <bool kTree, bool kCount, bool kDefer = false, bool kExt = false>
__device__ V direct_light(Ctx &c, const Surface &s, const Lobes &lb, V v, WfShadows *q = nullptr) {
    static_assert(!(kDefer && kExt), "the wavefront does not render glass, textures or local lights");
    const bool glass = glass_on<kExt>(c);
    V sum = v3(0.0f, 0.0f, 0.0f);
    const V o = offset_origin(s.p, s.ng);
    for (uint32_t i = 0; i < c.pu(COUNTS2, 0); i++) {
        const float4 *li = c.light(i);
        const float4 dir = li[1];
        const float cos_max = dir.w;
        const V to_light = vneg(xyz(dir));
        float u[2];
        c.g.r2(rng_sub(c.g.dim(D_SUN), i), u);
        const float cos_t = 1.0f - u[0] * (1.0f - cos_max);
        const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
        const float phi = kTwoPi * u[1];
        float sp, cp;
        sincos_c(phi, sp, cp);
        const M3 tb = basis(to_light);
        const V l = mmul(tb, v3(sin_t * cp, sin_t * sp, cos_t));
        const V col = xyz(li[2]);
        if (all_nonpos(col)) continue;
        const V f = eval_bsdf_k<kExt>(s, lb, v, l, li[4].y);
        if (kDefer && any_pos(f)) {
            const float pdf_l = 1.0f / (kTwoPi * (1.0f - cos_max));
            q->put(vs(vmul(col, f), light_mis(pdf_l, s, lb, v, l)), li[3].w >= 0.5f, o, l, kInf, kTraceShadow);
        } else if (glass && any_pos(f)) {
            const V vis = li[3].w >= 0.5f ? shadow_tr<kCount, kExt>(c, light_origin(s, o, l, glass), l, kInf, false)
                                          : v3(1.0f, 1.0f, 1.0f);
            if (any_pos(vis)) {
                const float pdf_l = 1.0f / (kTwoPi * (1.0f - cos_max));
                const float wgt = light_mis<kExt>(pdf_l, s, lb, v, l);
                sum = vadd(sum, vmul(vs(vmul(col, f), wgt), vis));
            }
        } else if (any_pos(f)) {
            const bool vis = li[3].w >= 0.5f ? !occluded<kCount>(c, o, l, kInf, kTraceShadow) : true;
            if (vis) {
                const float pdf_l = 1.0f / (kTwoPi * (1.0f - cos_max));
                const float wgt = light_mis(pdf_l, s, lb, v, l);
                sum = vadd(sum, vs(vmul(col, f), wgt));
            }
        }
    }
    if (kExt && c.pu(COUNTS, 0) > c.pu(COUNTS2, 0))
        sum = vadd(sum, local_lights<kCount, kExt>(c, s, lb, v, o, c.g.dim(D_LOCAL)));
    // rough glass takes emitters from both sides, as a point without a hemisphere (2488-2491)
    const V ne = two_sided<kExt>(lb) ? v3(0.0f, 0.0f, 0.0f) : s.n;
    c.mis_x = o;
    c.mis_n = ne;
    if (emitters_on(c)) {
        ES es;
        const bool seen = c.g.depth == 0u || (c.g.depth == 1u && c.mirror_path);
        // with the tree its candidates are already near the point: P.ltree.w of them at camera hits, one elsewhere
        if (kTree && lt_on(c))
            es = sample_emitter_ris<kTree, kExt>(
                c, o, ne, c.g.depth == 0u ? static_cast<uint32_t>(c.pf(LTREE, 3)) : 1u, s, lb, v, c.g.dim(D_EMIT));
        else if (seen)
            es = sample_emitter_ris<kTree, kExt>(c, o, ne, c.g.depth == 0u ? kEmitRisM : kEmitRisM / 2u, s, lb,
                                                         v, c.g.dim(D_EMIT));
        else if (c.pf(LIGHT_LOD, 3) > 0.0f)
            es = sample_emitter_ris<kTree, kExt>(c, o, ne, 2u, s, lb, v, c.g.dim(D_EMIT));
        else
            es = sample_emitter<kTree, kExt>(c, o, ne, c.g.dim(D_EMIT));
        if (es.pdf_w > 0.0f) {
            const V f = eval_bsdf<kExt>(s, lb, v, es.l);
            if (kDefer && any_pos(f)) {
                q->put(vs(vs(vmul(es.le, f), light_mis(es.pdf_w, s, lb, v, es.l)), es.w), true, o, es.l,
                       es.dist * 0.999f, kTraceAny);
            } else if (any_pos(f)) {
                // emitter_tr (1066-1070): any surface stops it, glass too
                if (!occluded<kCount>(c, light_origin(s, o, es.l, glass), es.l, es.dist * 0.999f, kTraceAny)) {
                    const float wgt = light_mis<kExt>(es.pdf_w, s, lb, v, es.l);
                    sum = vadd(sum, vs(vs(vmul(es.le, f), wgt), es.w));
                }
            }
        }
    }
    if (env_sampled(c)) {
        const V l = sample_env<kExt>(c, c.g.dim(D_ENV));
        const V f = eval_bsdf<kExt>(s, lb, v, l);
        if (kDefer && any_pos(f)) {
            const float pdf_e = env_pdf<kExt>(c, l);
            if (pdf_e > 0.0f)
                q->put(vds(vs(vmul(sky(c), f), light_mis(pdf_e, s, lb, v, l)), pdf_e), true, o, l, kInf, kTraceShadow);
        } else if (glass && any_pos(f)) {
            const float pdf_e = env_pdf<kExt>(c, l);
            if (pdf_e > 0.0f) {
                const V vis = shadow_tr<kCount, kExt>(c, light_origin(s, o, l, glass), l, kInf, true);
                if (any_pos(vis))
                    sum = vadd(sum, vmul(vds(vs(vmul(env_radiance<kExt>(c, l), f), light_mis<kExt>(pdf_e, s, lb, v, l)),
                                             pdf_e), vis));
            }
        } else if (any_pos(f)) {
            const float pdf_e = env_pdf<kExt>(c, l);
            if (pdf_e > 0.0f && !occluded<kCount>(c, o, l, kInf, kTraceShadow))
                sum = vadd(sum, vds(vs(vmul(env_radiance<kExt>(c, l), f), light_mis(pdf_e, s, lb, v, l)), pdf_e));
        }
    }
    return sum;
}

template <bool kTree>
__device__ float emit_mis(const Ctx &c, const HitRec &hit, V d, float pdf) {
    if (!emitters_on(c)) return 1.0f;
    const uint32_t rows = __float_as_uint(c.inst(hit.inst)[6].w);
    const uint32_t first = c.row_word(rows, hit.part, 2u);
    if (first == kNoEmitter) return 1.0f;
    const uint32_t rank = __float_as_uint(c.a->geo.leaves[3 * c.ck(hit.tri, c.a->geo.leaf_count, 26) + 1].w);
    const uint32_t entry = first + rank;
    const uint32_t context[3] = {hit.tri, hit.part, first};
    const Em em = emitter_tri_at(c, entry, 29, context);
    const V cr = cross(em.e1, em.e2);
    const float area = 0.5f * len(cr);
    const float cos_l = fabsf(dot(normalize(cr), d));
    const float sel = emitter_pmf<kTree>(c, entry, em, c.mis_x, c.mis_n);
    if (!(sel > 0.0f)) return 1.0f;
    return power_heuristic(pdf, sel / fmaxf(area, 1e-12f) * hit.t * hit.t / fmaxf(cos_l, 1e-6f));
}

__device__ __forceinline__ V clamp_lum(V col, float clamp_v) {
    const float lum = luminance(col);
    if (lum > clamp_v) col = vs(col, clamp_v / lum);
    return col;
}

// camera_ray (pathtrace.wgsl:4002-4024) for sample `index` at pixel (px, py), with the pixel's sampler set up.
__device__ __forceinline__ void camera_ray(Ctx &c, uint32_t px, uint32_t py, uint32_t index, V &o_out, V &d_out) {
    const uint32_t w = c.pu(SIZE, 0), h = c.pu(SIZE, 1);
    c.g.seed = hash_hp(px ^ hash_hp(py ^ hash_hp(c.pu(SIZE, 3))));
    c.g.index = index;
    c.g.depth = 0;
    float cam[4];
    c.g.r4(c.g.dim(D_CAMERA), cam);
    const float ndx = (static_cast<float>(px) + cam[0]) / static_cast<float>(w) * 2.0f - 1.0f;
    const float ndy = 1.0f - (static_cast<float>(py) + cam[1]) / static_cast<float>(h) * 2.0f;
    const V fwd = v3(c.pf(CAM_FWD, 0), c.pf(CAM_FWD, 1), c.pf(CAM_FWD, 2));
    const V right = v3(c.pf(CAM_RIGHT, 0), c.pf(CAM_RIGHT, 1), c.pf(CAM_RIGHT, 2));
    const V up = v3(c.pf(CAM_UP, 0), c.pf(CAM_UP, 1), c.pf(CAM_UP, 2));
    const float tanh = c.pf(CAM_POS, 3), aspect = c.pf(CAM_FWD, 3);
    V d = normalize(vadd(vadd(fwd, vs(right, ndx * tanh * aspect)), vs(up, ndy * tanh)));
    V o = v3(c.pf(CAM_POS, 0), c.pf(CAM_POS, 1), c.pf(CAM_POS, 2));
    if (c.pf(CAM_RIGHT, 3) > 0.0f) {  // the thin lens (pathtrace.wgsl:4017-4023): cam_right.w radius, cam_up.w focus
        const V focus = vadd(o, vs(d, c.pf(CAM_UP, 3) / dot(d, fwd)));
        const float r = sqrtf(cam[2]) * c.pf(CAM_RIGHT, 3);
        const float phi = kTwoPi * cam[3];
        float sp, cp;
        sincos_c(phi, sp, cp);
        o = vadd(o, vadd(vs(right, r * cp), vs(up, r * sp)));
        d = normalize(vsub(focus, o));
    }
    o_out = o, d_out = d;
}

// One path of sample `index` at pixel (px, py): its radiance, the camera hit's albedo and normal.
template <bool kCoat, bool kTree, bool kCount, bool kExt>
__device__ void path(Ctx &c, uint32_t px, uint32_t py, uint32_t index, V &radiance, V &albedo, V &normal) {
    V o, d;
    camera_ray(c, px, py, index, o, d);
    radiance = v3(0.0f, 0.0f, 0.0f);
    albedo = v3(0.0f, 0.0f, 0.0f);
    normal = v3(0.0f, 0.0f, 0.0f);
    V thr = v3(1.0f, 1.0f, 1.0f);
    bool specular_bounce = true, straight = true;
    // glass passed since the last scattering vertex (pathtrace.wgsl:4156-4162): any, and an odd number of surfaces
    bool behind_glass = false, bent = false;
    float last_pdf = 0.0f;
    // the ray cone (Akenine-Moller et al. 2019) for the textures' level: width at the origin, spread angle (4165-4167)
    float cone_w = 0.0f, cone_a = kExt ? 2.0f * c.pf(CAM_POS, 3) / lod_height(c) : 0.0f;
    const float clamp_v = c.pf(MISC, 0);
    c.mirror_path = false;
    const uint32_t last = c.pu(COUNTS, 2);
    for (uint32_t depth = 0; depth <= last; depth++) {
        c.g.depth = depth;
        const HitRec hit = c.trace<kCount>(o, d, kInf, kTraceClosest);
        if (hit.tri == kNone) {
            V col = vmul(thr, sky_mis<kExt>(c, d, (depth > 0u && !(kExt && bent)) ? last_pdf : 0.0f));
            if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
            radiance = vadd(radiance, col);
            break;
        }
        const float width = kExt ? cone_w + cone_a * hit.t : 0.0f;
        const Surface s = surface<kCoat, kExt>(c, o, d, hit, width);
        if (kCount) c.cnt->shaded++;
        if (depth == 0u) {
            if (s.card) {
                const float mx = fmaxf(fmaxf(s.emissive.x, fmaxf(s.emissive.y, s.emissive.z)), 1e-6f);
                albedo = vds(s.emissive, mx);
            } else {
                albedo = s.base;
            }
            normal = s.n;
        }
        if (any_pos(s.emissive) && (s.front || !s.card)) {
            float wgt = 1.0f;
            if (depth > 0u && !specular_bounce && !(kExt && behind_glass)) wgt = emit_mis<kTree>(c, hit, d, last_pdf);
            V col = vs(vmul(thr, s.emissive), wgt);
            if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
            radiance = vadd(radiance, col);
        }
        if (s.card) break;
        if (depth == last) break;
        const V v = vneg(d);
        const float nov = fmaxf(dot(s.n, v), 1e-4f);
        const Lobes lb = lobes<kExt>(c, s, nov);
        V dl = vmul(thr, direct_light<kTree, kCount, false, kExt>(c, s, lb, v));
        if (depth > 0u && !straight) dl = clamp_lum(dl, clamp_v);
        radiance = vadd(radiance, dl);
        const V l = sample_bsdf<kExt>(c, s, lb, v);
        if (kExt) {
            // 4414-4465: through smooth glass a delta lobe, weight tint / its sampling share, no MIS
            const bool through = c.sampled_trans && !lb.rough_t;
            float pdf = 0.0f;
            V wgt = vds(lb.tint, fmaxf(lb.p_trans, 1e-6f));
            if (!through) {
                pdf = pdf_bsdf<kExt>(s, lb, v, l);
                wgt = vds(eval_bsdf<kExt>(s, lb, v, l), fmaxf(pdf, 1e-20f));
            }
            if ((!through && pdf <= 0.0f) || all_nonpos(wgt)) break;
            thr = vmul(thr, wgt);
            if (through && dot(l, s.ng) < 0.0f) {
                behind_glass = true;
                bent = !bent;
            } else {
                last_pdf = through ? 0.0f : pdf;
                specular_bounce = through;
                behind_glass = false;
                bent = false;
            }
            straight = straight && through;
            if (kExt) {
                cone_w = width;
                cone_a += 2.0f * s.roughness * s.roughness;
            }
        } else {
            const float pdf = pdf_bsdf(s, lb, v, l);
            const V wgt = vds(eval_bsdf(s, lb, v, l), fmaxf(pdf, 1e-20f));
            if (pdf <= 0.0f || all_nonpos(wgt)) break;
            thr = vmul(thr, wgt);
            last_pdf = pdf;
            specular_bounce = false;
            straight = false;
            if (kExt) {
                cone_w = width;
                cone_a += 2.0f * s.roughness * s.roughness;
            }
        }
        c.mirror_path = (depth == 0u || c.mirror_path) && s.roughness < kMirrorRoughness &&
                        (c.sampled_spec || (kExt && c.sampled_trans));
        if (depth >= c.pu(COUNTS, 3)) {
            const float q = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.95f);
            if (c.g.r1(c.g.dim(D_RR)) > q) break;
            thr = vds(thr, q);
        }
        o = offset_origin(s.p, dot(l, s.ng) < 0.0f ? vneg(s.ng) : s.ng);
        d = l;
    }
    if (non_finite(radiance)) radiance = v3(0.0f, 0.0f, 0.0f);
    if (non_finite(albedo)) albedo = v3(0.0f, 0.0f, 0.0f);
    if (non_finite(normal)) normal = v3(0.0f, 1.0f, 0.0f);
}

// Every job of the launch is a run of 8x8 tiles (RenderJob::tile_first on); a wave takes the next tile of the whole
// launch, finds its job (a binary search over the prefix, the same for the whole wave) and renders it with that
// job's scene, parameters and accumulator. Jobs of any size share the launch: what one job leaves idle, another fills.
// The path loop needs about 200 registers, which leaves two waves a SIMD; held to three (168 registers, about 50
// spilled to scratch, mostly in the shading code), the kernel hides more of its memory latency: the 720p city
// frame 2.46 -> 1.95 ms, the 3M-triangle scene 1.36x faster, the Cornell box 5 % (four waves: 2.00 ms, 1.34x, 11 %).
// kCount: a counting launch, which adds what its walks read and the hits it shades to reads[0..3] (nodes,
// triangles, units, shaded hits).
template <bool kCoat, bool kTree, bool kCount, bool kExt>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_render(
    const RenderJob *jobs, uint32_t njobs, uint32_t total, uint32_t *counter, uint32_t *spill, uint32_t spill_stride,
    uint32_t spill_entries, uint32_t *guard, unsigned long long *rays, unsigned long long *reads) {
    __shared__ uint32_t lds[kLdsStack * kRenderBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st{(TRT_LDS uint32_t *)(lds + threadIdx.x), spill, gtid, spill_stride, 0, kLdsStack + spill_entries};
    uint32_t flags = 0;
    uint32_t traced[2] = {0, 0};
    Reads cnt;
    const ShardSet sh{counter, nullptr, total, 0};
    uint32_t shard = blockIdx.x % kShards, closed = 0;
    for (;;) {
        uint32_t req = 0;
        if (lane == 0) req = atomicAdd(sh.counter(shard), 1u);
        uint32_t tile, t1;
        if (!shard_take<false>(sh, shard, closed, req, 1u, lane, tile, t1)) break;
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (jobs[mid].tile_first <= tile)
                lo = mid;
            else
                hi = mid;
        }
        const RenderJob *job = jobs + lo;
        const RenderArgs *a = &job->args;
        const uint32_t t = tile - job->tile_first;
        const uint32_t w = a->params[16], h = a->params[17], tiles_x = (w + 7) / 8;
        Ctx c{a, &st, &flags, traced, Rng{0, 0, 0}, false, false, false, v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f), &cnt};
        const uint32_t px = (t % tiles_x) * 8 + (lane & 7u), py = (t / tiles_x) * 8 + (lane >> 3);
        if (px < w && py < h) {
            float4 *pix = a->accum + 5 * (uint64_t(py) * w + px);
            float4 col = pix[0], alb = pix[1], nrm = pix[2];
            for (uint32_t s = 0; s < a->spp; s++) {
                V rad, al, nm;
                path<kCoat, kTree, kCount, kExt>(c, px, py, a->first_sample + s, rad, al, nm);
                const float lum = luminance(rad);
                col.x += rad.x, col.y += rad.y, col.z += rad.z, col.w += 1.0f;
                alb.x += al.x, alb.y += al.y, alb.z += al.z, alb.w += lum;
                nrm.x += nm.x, nrm.y += nm.y, nrm.z += nm.z, nrm.w += lum * lum;
            }
            pix[0] = col;
            pix[1] = alb;
            pix[2] = nrm;
        }
    }
    if (flags) atomicOr(guard, flags);
    // the wave's ray counts, summed across lanes, one atomic each
    for (int k = 32; k >= 1; k >>= 1) {
        traced[0] += __shfl_xor(traced[0], k, 64);
        traced[1] += __shfl_xor(traced[1], k, 64);
    }
    if (lane == 0) {
        unsigned long long *r = rays + (blockIdx.x % kShards) * kRayShardGap;
        atomicAdd(r, (unsigned long long)traced[0]);
        atomicAdd(r + 1, (unsigned long long)traced[1]);
    }
    if (kCount) {
        atomicAdd(reads + 0, (unsigned long long)cnt.nodes);
        atomicAdd(reads + 1, (unsigned long long)cnt.tris);
        atomicAdd(reads + 2, (unsigned long long)cnt.units);
        atomicAdd(reads + 3, (unsigned long long)cnt.shaded);
    }
}

// ---- the wavefront path tracer ---------------------------------------------------------------------------------
// The path loop above, cut at its traces: k_wf_begin queues every path's camera ray, the wide walk (k_trace_dyn)
// traces the queued rays of a bounce, and k_wf_shade resolves the previous bounce's shadow rays, shades the hit and
// queues the next rays. A path's state between kernels: throughput and last pdf (thr), radiance and these flag bits
// (rad.w), the camera hit's albedo and normal, direct light's last vertex (emit_mis), and the pending direct light.
enum : uint32_t {
    kWfDepth = 0xffu,        // the bounce of the ray in slot 0
    kWfHasRay = 1u << 8,     // slot 0 holds a traced closest ray to shade
    kWfSpecular = 1u << 9,   // the path's specular_bounce
    kWfStraight = 1u << 10,  // the path's straight
    kWfMirror = 1u << 11,    // the path's mirror_path
    kWfDl = 1u << 12,        // a bounce's direct light waits: its visible contributions x thr_dl go to the radiance
    kWfClampDl = 1u << 13,   // ... clamped (clamp_lum)
    kWfPresentShift = 14,    // 5 bits: contribution j is there
    kWfTracedShift = 19,     // 5 bits: contribution j counts only if shadow ray j + 1 missed
};
static_assert(kWfShadow == 5, "the flag words hold 5 contributions");

struct WfPath {
    uint32_t job, px, py, index;
};
// path p of the chunk: its run (a binary search over the runs' first paths), pixel and sample
__device__ __forceinline__ WfPath wf_path(const WfArgs &a, uint32_t p) {
    uint32_t lo = 0, hi = a.nsubs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (a.subs[mid].path_first <= p)
            lo = mid;
        else
            hi = mid;
    }
    const WfSub &s = a.subs[lo];
    const uint32_t off = p - s.path_first;
    const uint32_t pixel = s.pixel_first + off % s.count;
    return {s.job, pixel % s.width, pixel / s.width, s.sample_first + off / s.count};
}
__device__ __forceinline__ Ctx wf_ctx(const RenderArgs *ra, uint32_t *flags, uint32_t *traced, Reads *cnt) {
    return Ctx{ra, nullptr, flags, traced, Rng{0, 0, 0}, false, false, false, v3(0.0f, 0.0f, 0.0f), v3(0.0f, 0.0f, 0.0f),
               cnt};
}
// The wave's lanes take their own jobs one job at a time, so that a job's arguments are the same for the whole wave
// (scalar loads) however the wave's paths are spread over jobs. `body(ra)` runs once in every lane with `valid`.
template <typename F>
__device__ __forceinline__ void wf_by_job(const WfArgs &a, bool valid, uint32_t job, F &&body) {
    bool done = !valid;
    for (uint64_t todo = __ballot(!done); todo != 0; todo = __ballot(!done)) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(__shfl(job, __ffsll((unsigned long long)todo) - 1, 64));
        if (!done && job == j) {
            body(&a.jobs[j].args);
            done = true;
        }
    }
}
// The queues are kShards segments (shards.h), each with its own counts: a path stays in its segment (p / cap) from
// bounce to bounce, the shade kernel's blocks of segment s (blockIdx % kShards == s) take its active paths and append
// to it, and the walk kernel takes the segments as its shards.
__device__ __forceinline__ uint32_t *wf_count(const WfArgs &a, uint32_t ring, uint32_t which, uint32_t seg) {
    return a.counts + ((ring * 2 + which) * kShards + seg) * kShardGap;  // which: 0 the ray list, 1 the active paths
}
// The wave's queued rays (bit s of `slots`: ray slot s of path p) appended to segment `seg` of ray list (ring, parity),
// one atomic a wave, and its paths still running to the segment of the active list.
__device__ __forceinline__ void wf_append(const WfArgs &a, uint32_t ring, uint32_t parity, uint32_t seg, uint32_t p,
                                          uint32_t slots, bool alive, uint32_t lane, unsigned long long &closest,
                                          unsigned long long &shadow) {
    const uint64_t below = (uint64_t{1} << lane) - 1;
    uint32_t total = 0;
    for (uint32_t s = 0; s < kWfSlots; s++) {
        const uint32_t m = __popcll(__ballot((slots >> s) & 1u));
        total += m;
        if (s == 0) closest += m;
        else shadow += m;
    }
    uint32_t first = 0;
    if (lane == 0 && total) first = atomicAdd(wf_count(a, ring, 0, seg), total);
    first = __builtin_amdgcn_readfirstlane(__shfl(first, 0, 64));
    const uint64_t st = a.stride;
    uint32_t *list = a.list[parity] + uint64_t(seg) * kWfSlots * a.cap;
    for (uint32_t s = 0; s < kWfSlots; s++) {
        const uint64_t b = __ballot((slots >> s) & 1u);
        if ((slots >> s) & 1u) list[first + __popcll(b & below)] = uint32_t(s * st + p);
        first += __popcll(b);
    }
    const uint64_t live = __ballot(alive);
    uint32_t afirst = 0;
    if (lane == 0 && live) afirst = atomicAdd(wf_count(a, ring, 1, seg), uint32_t(__popcll(live)));
    afirst = __builtin_amdgcn_readfirstlane(__shfl(afirst, 0, 64));
    if (alive) a.active[parity][uint64_t(seg) * a.cap + afirst + __popcll(live & below)] = p;
}
__device__ __forceinline__ void wf_flush(const WfArgs &a, uint32_t flags, uint32_t lane, unsigned long long closest,
                                         unsigned long long shadow) {
    if (flags) atomicOr(a.guard, flags);
    if (lane == 0 && (closest | shadow)) {
        unsigned long long *r = a.rays_traced + (blockIdx.x % kShards) * kRayShardGap;
        atomicAdd(r, closest);
        atomicAdd(r + 1, shadow);
    }
}

// Every path's camera ray in slot 0 and its state reset; ray list 0 and active list 0 hold every path, path p in
// segment p / cap.
__global__ __launch_bounds__(kRenderBlock) void k_wf_begin(WfArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    if (blockIdx.x == 0 && threadIdx.x < kShards) {
        const uint32_t seg = threadIdx.x, lo = seg * a.cap;
        const uint32_t n = a.npaths > lo ? min(a.npaths - lo, a.cap) : 0u;
        *wf_count(a, 0, 0, seg) = *wf_count(a, 0, 1, seg) = n;
        *wf_count(a, 1, 0, seg) = *wf_count(a, 1, 1, seg) = 0;
        if (seg == 0) atomicAdd(a.rays_traced, (unsigned long long)a.npaths);  // the camera rays
    }
    uint32_t flags = 0, traced[2] = {0, 0};
    Reads cnt;
    const uint32_t waves = gridDim.x * (kRenderBlock / 64);
    for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) / 64 * 64; base < a.npaths; base += waves * 64) {
        const uint32_t p = base + lane;
        const bool valid = p < a.npaths;
        const WfPath id = valid ? wf_path(a, p) : WfPath{0, 0, 0, 0};
        wf_by_job(a, valid, id.job, [&](const RenderArgs *ra) {
            Ctx c = wf_ctx(ra, &flags, traced, &cnt);
            V o, d;
            camera_ray(c, id.px, id.py, id.index, o, d);
            a.rays[2 * uint64_t(p)] = make_float4(o.x, o.y, o.z, kInf);
            a.rays[2 * uint64_t(p) + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(kTraceClosest));
            a.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            a.rad[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kWfHasRay | kWfSpecular | kWfStraight));
            a.alb[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            a.nrm[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const uint32_t seg = p / a.cap;
            a.list[0][uint64_t(seg) * kWfSlots * a.cap + (p - seg * a.cap)] = p;
            a.active[0][p] = p;
        });
    }
    wf_flush(a, flags, lane, 0, 0);
}

// One bounce of path p (the loop body of `path`), from what the walk left in its ray slots: the pending direct light
// resolved first, then the hit in slot 0 shaded. Returns the ray slots it queued (bit s: slot s); `alive` says
// whether the path waits on anything.
template <bool kCoat, bool kTree>
__device__ uint32_t wf_bounce(Ctx &c, const WfArgs &a, uint32_t p, const WfPath &id, bool &alive) {
    const uint64_t st = a.stride;
    const float4 rd = a.rad[p];
    uint32_t f = __float_as_uint(rd.w);
    V radiance = xyz(rd);
    const float clamp_v = c.pf(MISC, 0);
    if (f & kWfDl) {
        V sum = v3(0.0f, 0.0f, 0.0f);
        for (uint32_t j = 0; j < kWfShadow; j++) {
            if (!((f >> (kWfPresentShift + j)) & 1u)) continue;
            if ((f >> (kWfTracedShift + j)) & 1u) {
                const uint32_t tri = a.hits[6 * ((j + 1) * st + p) + 1];
                if (tri != kNone) continue;
            }
            sum = vadd(sum, xyz(a.contrib[j * st + p]));
        }
        V dl = vmul(xyz(a.thr_dl[p]), sum);
        if (f & kWfClampDl) dl = clamp_lum(dl, clamp_v);
        radiance = vadd(radiance, dl);
        f &= ~(kWfDl | kWfClampDl | (0x1fu << kWfPresentShift) | (0x1fu << kWfTracedShift));
    }
    uint32_t slots = 0;
    if (f & kWfHasRay) {
        f &= ~kWfHasRay;
        const uint32_t depth = f & kWfDepth;
        c.g.seed = hash_hp(id.px ^ hash_hp(id.py ^ hash_hp(c.pu(SIZE, 3))));
        c.g.index = id.index;
        c.g.depth = depth;
        c.mirror_path = (f & kWfMirror) != 0;
        const bool specular_bounce = (f & kWfSpecular) != 0, straight = (f & kWfStraight) != 0;
        const float4 t4 = a.thr[p];
        V thr = xyz(t4);
        const float last_pdf = t4.w;
        const float4 r0 = a.rays[2 * uint64_t(p)], r1 = a.rays[2 * uint64_t(p) + 1];
        const V o = xyz(r0), d = xyz(r1);
        const uint32_t *hw = a.hits + 6 * uint64_t(p);
        HitRec hit;
        hit.t = __uint_as_float(hw[0]);
        hit.tri = hw[1];
        hit.u = __uint_as_float(hw[2]);
        hit.v = __uint_as_float(hw[3]);
        hit.inst = hw[4];
        hit.part = hw[5];
        const uint32_t last = c.pu(COUNTS, 2);
        if (hit.tri == kNone) {
            V col = vmul(thr, sky_mis<false>(c, d, depth > 0u ? last_pdf : 0.0f));
            if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
            radiance = vadd(radiance, col);
        } else {
            const Surface s = surface<kCoat>(c, o, d, hit);
            if (depth == 0u) {
                V albedo;
                if (s.card) {
                    const float mx = fmaxf(fmaxf(s.emissive.x, fmaxf(s.emissive.y, s.emissive.z)), 1e-6f);
                    albedo = vds(s.emissive, mx);
                } else {
                    albedo = s.base;
                }
                a.alb[p] = make_float4(albedo.x, albedo.y, albedo.z, 0.0f);
                a.nrm[p] = make_float4(s.n.x, s.n.y, s.n.z, 0.0f);
            }
            if (any_pos(s.emissive) && (s.front || !s.card)) {
                float wgt = 1.0f;
                if (depth > 0u && !specular_bounce) {
                    c.mis_x = xyz(a.misx[p]);
                    c.mis_n = xyz(a.misn[p]);
                    wgt = emit_mis<kTree>(c, hit, d, last_pdf);
                }
                V col = vs(vmul(thr, s.emissive), wgt);
                if (depth > 1u && !straight) col = clamp_lum(col, clamp_v);
                radiance = vadd(radiance, col);
            }
            if (!s.card && depth != last) {
                const V v = vneg(d);
                const float nov = fmaxf(dot(s.n, v), 1e-4f);
                const Lobes lb = lobes(c, s, nov);
                WfShadows q{&a, p};
                direct_light<kTree, false, true>(c, s, lb, v, &q);
                a.thr_dl[p] = make_float4(thr.x, thr.y, thr.z, 0.0f);
                a.misx[p] = make_float4(c.mis_x.x, c.mis_x.y, c.mis_x.z, 0.0f);
                a.misn[p] = make_float4(c.mis_n.x, c.mis_n.y, c.mis_n.z, 0.0f);
                f |= kWfDl | (depth > 0u && !straight ? kWfClampDl : 0u) | (q.present << kWfPresentShift) |
                     (q.traced << kWfTracedShift);
                slots |= q.traced << 1;
                const V l = sample_bsdf(c, s, lb, v);
                const float pdf = pdf_bsdf(s, lb, v, l);
                const V wgt = vds(eval_bsdf(s, lb, v, l), fmaxf(pdf, 1e-20f));
                bool go = !(pdf <= 0.0f || all_nonpos(wgt));
                if (go) {
                    thr = vmul(thr, wgt);
                    const bool mirror = (depth == 0u || c.mirror_path) && s.roughness < kMirrorRoughness &&
                                        c.sampled_spec;
                    if (depth >= c.pu(COUNTS, 3)) {
                        const float qq = fminf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.95f);
                        if (c.g.r1(c.g.dim(D_RR)) > qq)
                            go = false;
                        else
                            thr = vds(thr, qq);
                    }
                    if (go) {
                        const V no = offset_origin(s.p, dot(l, s.ng) < 0.0f ? vneg(s.ng) : s.ng);
                        a.rays[2 * uint64_t(p)] = make_float4(no.x, no.y, no.z, kInf);
                        a.rays[2 * uint64_t(p) + 1] = make_float4(l.x, l.y, l.z, __uint_as_float(kTraceClosest));
                        a.thr[p] = make_float4(thr.x, thr.y, thr.z, pdf);
                        f = (f & ~(kWfDepth | kWfSpecular | kWfStraight | kWfMirror)) | (depth + 1u) | kWfHasRay |
                            (mirror ? kWfMirror : 0u);
                        slots |= 1u;
                    }
                }
            }
        }
    }
    a.rad[p] = make_float4(radiance.x, radiance.y, radiance.z, __uint_as_float(f));
    alive = (f & (kWfHasRay | kWfDl)) != 0;
    return slots;
}

// Bounce k's shading: the paths of active list k % 2 (its counts in ring entry k % 3); the rays and paths it queues go
// to the lists (k + 1) % 2, counted in ring entry (k + 1) % 3, which the launch of bounce k - 1 cleared. Segment s is
// the work of the blocks with blockIdx % kShards == s (the grid has at least kShards blocks).
template <bool kCoat, bool kTree>
__global__ __launch_bounds__(kRenderBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_wf_shade(WfArgs a,
                                                                                                   uint32_t k) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t in = k % 3u, out = (k + 1u) % 3u, clear = (k + 2u) % 3u;
    if (blockIdx.x == 0 && threadIdx.x < 2 * kShards) *wf_count(a, clear, threadIdx.x / kShards, threadIdx.x % kShards) = 0;
    const uint32_t seg = blockIdx.x % kShards;
    const uint32_t nb = (gridDim.x - seg + kShards - 1) / kShards, bi = blockIdx.x / kShards;
    const uint32_t n = *wf_count(a, in, 1, seg);
    const uint32_t *act = a.active[k & 1u] + uint64_t(seg) * a.cap;
    uint32_t flags = 0, traced[2] = {0, 0};
    Reads cnt;
    unsigned long long closest = 0, shadow = 0;
    for (uint32_t base = bi * kRenderBlock + threadIdx.x / 64 * 64; base < n; base += nb * kRenderBlock) {
        const uint32_t i = base + lane;
        const bool valid = i < n;
        const uint32_t p = valid ? act[i] : 0u;
        const WfPath id = valid ? wf_path(a, p) : WfPath{0, 0, 0, 0};
        uint32_t slots = 0;
        bool alive = false;
        wf_by_job(a, valid, id.job, [&](const RenderArgs *ra) {
            Ctx c = wf_ctx(ra, &flags, traced, &cnt);
            slots = wf_bounce<kCoat, kTree>(c, a, p, id, alive);
        });
        wf_append(a, out, (k + 1u) & 1u, seg, p, slots, alive, lane, closest, shadow);
    }
    wf_flush(a, flags, lane, closest, shadow);
}

// Every path's radiance, albedo and normal added to its pixel's record, as k_render adds a pixel's samples: in the
// order of their sample index (a run holds a pixel's samples `count` paths apart; the runs of one pixel come in
// sample order, chunk after chunk).
__global__ __launch_bounds__(kRenderBlock) void k_wf_accum(WfArgs a, uint32_t pixels) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = a.nsubs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (a.subs[mid].acc_first <= i)
                lo = mid;
            else
                hi = mid;
        }
        const WfSub s = a.subs[lo];
        const RenderArgs *ra = &a.jobs[s.job].args;
        const uint32_t j = i - s.acc_first, pixel = s.pixel_first + j;
        float4 *pix = ra->accum + 5 * uint64_t(pixel);
        float4 col = pix[0], alb = pix[1], nrm = pix[2];
        for (uint32_t r = 0; r < s.rounds; r++) {
            const uint32_t p = s.path_first + r * s.count + j;
            V rad = xyz(a.rad[p]), al = xyz(a.alb[p]), nm = xyz(a.nrm[p]);
            if (non_finite(rad)) rad = v3(0.0f, 0.0f, 0.0f);
            if (non_finite(al)) al = v3(0.0f, 0.0f, 0.0f);
            if (non_finite(nm)) nm = v3(0.0f, 1.0f, 0.0f);
            const float lum = luminance(rad);
            col.x += rad.x, col.y += rad.y, col.z += rad.z, col.w += 1.0f;
            alb.x += al.x, alb.y += al.y, alb.z += al.z, alb.w += lum;
            nrm.x += nm.x, nrm.y += nm.y, nrm.z += nm.z, nrm.w += lum * lum;
        }
        pix[0] = col;
        pix[1] = alb;
        pix[2] = nrm;
    }
}

__global__ void k_reset(uint32_t *c) { c[threadIdx.x * kShardGap] = 0; }

}  // namespace

int render_occupancy(int *blocks_per_cu) {
    int b = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_render<false, false, false, false>, kRenderBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int wf_shade_occupancy(int *blocks_per_cu) {
    int b = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_shade<false, false>, kRenderBlock, 0);
    *blocks_per_cu = b;
    return e == hipSuccess ? 0 : -1;
}

int launch_wf_begin(const WfArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_wf_begin, dim3(grid), dim3(kRenderBlock), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_wf_shade(const WfArgs &a, uint32_t k, bool coat, bool tree, int grid, hipStream_t s) {
    if (coat && tree)
        hipLaunchKernelGGL((k_wf_shade<true, true>), dim3(grid), dim3(kRenderBlock), 0, s, a, k);
    else if (coat)
        hipLaunchKernelGGL((k_wf_shade<true, false>), dim3(grid), dim3(kRenderBlock), 0, s, a, k);
    else if (tree)
        hipLaunchKernelGGL((k_wf_shade<false, true>), dim3(grid), dim3(kRenderBlock), 0, s, a, k);
    else
        hipLaunchKernelGGL((k_wf_shade<false, false>), dim3(grid), dim3(kRenderBlock), 0, s, a, k);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_wf_accum(const WfArgs &a, uint32_t pixels, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_wf_accum, dim3(grid), dim3(kRenderBlock), 0, s, a, pixels);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_render(const RenderJob *d_jobs, uint32_t njobs, uint32_t total_tiles, uint32_t *counter,
                  uint32_t *spill, uint32_t spill_stride, uint32_t spill_entries, uint32_t *guard,
                  unsigned long long *rays, bool coat, bool tree, bool ext, unsigned long long *reads, int grid,
                  hipStream_t s) {
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(kShards), 0, s, counter);
#define TRT_RENDER(C, T, N, E)                                                                                       \
    hipLaunchKernelGGL((k_render<C, T, N, E>), dim3(grid), dim3(kRenderBlock), 0, s, d_jobs, njobs, total_tiles,     \
                       counter, spill, spill_stride, spill_entries, guard, rays, reads)
#define TRT_RENDER_N(C, T, E)          \
    if (reads)                         \
        TRT_RENDER(C, T, true, E);     \
    else                               \
        TRT_RENDER(C, T, false, E)
    // Scenes with glass, textures or local lights take the kernel with their code and the coat's terms (the same
    // bits for materials without a coat); which of them a job's scene has, the kernel reads from its arguments.
    if (ext && tree) {
        TRT_RENDER_N(true, true, true);
    } else if (ext) {
        TRT_RENDER_N(true, false, true);
    } else if (coat && tree) {
        TRT_RENDER_N(true, true, false);
    } else if (coat) {
        TRT_RENDER_N(true, false, false);
    } else if (tree) {
        TRT_RENDER_N(false, true, false);
    } else {
        TRT_RENDER_N(false, false, false);
    }
#undef TRT_RENDER_N
#undef TRT_RENDER
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace trt
