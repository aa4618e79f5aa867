/*
 * Oracle (test infrastructure only): the kernel's sampler, sample.wgsl restated (sample.rs is its CPU twin).
 *   sample.wgsl:5-64   sobol_d1..d3 (Joe & Kuo dimensions 1-3, bit-reversed, 16 bits)
 *   sample.wgsl:67-75  owen_rev;  77-85 hash_hp;  88-101 sobol_owen
 *   sample.wgsl:110-111 rng_dim / rng_sub;  144-169 rng_seed / rng_index / rng1..rng4;  173-177 rng_start
 * Integer arithmetic (exact); one u32 -> f32 conversion (round to nearest even) per number.
 */
#include <stdint.h>

#include "oracle_render.h"

/* sample.rs:5-32 SOBOL_REV, computed from the Joe-Kuo initial numbers (dimension: s, a, m0). */
static uint32_t g_dirs[3][32];
static int g_dirs_ready = 0;

static uint32_t rev32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}

static void dirs(int s, uint32_t a, const uint32_t *m0, uint32_t *out) {
    uint32_t m[32];
    for (int i = 0; i < 32; i++) {
        if (i < s) {
            m[i] = m0[i];
        } else {
            uint32_t v = m[i - s] ^ (m[i - s] << s);
            for (int k = 1; k < s; k++) v ^= ((a >> (s - 1 - k)) & 1u) * (m[i - k] << k);
            m[i] = v;
        }
    }
    for (int i = 0; i < 32; i++) out[i] = rev32(m[i] << (31 - i));
}

static void init_dirs(void) {
    if (g_dirs_ready) return;
    const uint32_t m1[3] = {1, 0, 0}, m2[3] = {1, 3, 0}, m3[3] = {1, 3, 1};
    dirs(1, 0, m1, g_dirs[0]);
    dirs(2, 1, m2, g_dirs[1]);
    dirs(3, 1, m3, g_dirs[2]);
    g_dirs_ready = 1;
}

void or_sobol_dirs(uint32_t *out96) {
    init_dirs();
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < 32; i++) out96[32 * d + i] = g_dirs[d][i];
}

uint32_t or_reverse_bits(uint32_t x) { return rev32(x); }

uint32_t or_owen_rev(uint32_t n, uint32_t seed) {
    n ^= n * 0x3d20adeau;
    n += seed;
    n *= (seed >> 16) | 1u;
    n ^= n * 0x05526c56u;
    n ^= n * 0x53a22864u;
    return n;
}

uint32_t or_hash_hp(uint32_t i) {
    i ^= i >> 16;
    i *= 0x21f0aaadu;
    i ^= i >> 15;
    i *= 0xd35a2d97u;
    i ^= i >> 15;
    return i ^ 0xe6fe3bebu;
}

/* sample.wgsl:5-64: dimension k's 16 top index bits, each selecting a direction number. */
static uint32_t sobol_dk(int k, uint32_t x) {
    init_dirs();
    uint32_t r = 0;
    for (int b = 0; b < 16; b++)
        if (x & (0x80000000u >> b)) r ^= g_dirs[k - 1][b];
    return r;
}

float or_sobol_owen(uint32_t rev_index, uint32_t dim, uint32_t seed) {
    uint32_t r = dim == 0 ? rev32(rev_index) : sobol_dk((int)dim, rev_index);
    r = rev32(or_owen_rev(r, seed));
    return (float)r * (float)(1.0 / 4294967808.0);
}

uint32_t or_rng_dim(const or_rng *g, uint32_t slot) { return (g->path << 28) ^ (g->depth << 20) ^ slot; }
uint32_t or_rng_sub(uint32_t d, uint32_t k) { return or_hash_hp(d ^ or_hash_hp(k + 0x9e3779b9u)) | 0x80000000u; }
static uint32_t rng_seed(const or_rng *g, uint32_t d) { return g->seed ^ or_hash_hp(d); }
static uint32_t rng_index(const or_rng *g, uint32_t salt) {
    return or_owen_rev(rev32(g->index), salt) & 0xffff0000u;
}

float or_rng1(const or_rng *g, uint32_t d) {
    uint32_t s = rng_seed(g, d);
    return or_sobol_owen(rng_index(g, s ^ 0xbff95bfeu), 0, s ^ 0x635c77bdu);
}
void or_rng2(const or_rng *g, uint32_t d, float *r) {
    uint32_t s = rng_seed(g, d), i = rng_index(g, s ^ 0xf8ade99au);
    r[0] = or_sobol_owen(i, 0, s ^ 0xe0aaaf76u);
    r[1] = or_sobol_owen(i, 1, s ^ 0x94964d4eu);
}
void or_rng3(const or_rng *g, uint32_t d, float *r) {
    uint32_t s = rng_seed(g, d), i = rng_index(g, s ^ 0xcaa726acu);
    r[0] = or_sobol_owen(i, 0, s ^ 0x9e78e391u);
    r[1] = or_sobol_owen(i, 1, s ^ 0x67c33241u);
    r[2] = or_sobol_owen(i, 2, s ^ 0x78c395c5u);
}
void or_rng4(const or_rng *g, uint32_t d, float *r) {
    uint32_t s = rng_seed(g, d), i = rng_index(g, s ^ 0xc2c1a055u);
    r[0] = or_sobol_owen(i, 0, s ^ 0x39468210u);
    r[1] = or_sobol_owen(i, 1, s ^ 0xe9d8a845u);
    r[2] = or_sobol_owen(i, 2, s ^ 0x5f32b482u);
    r[3] = or_sobol_owen(i, 3, s ^ 0x1524cc56u);
}
void or_rng_start(or_rng *g, uint32_t px, uint32_t py, uint32_t index, uint32_t seed, uint32_t path) {
    g->seed = or_hash_hp(px ^ or_hash_hp(py ^ or_hash_hp(seed)));
    g->index = index;
    g->path = path;
    g->depth = 0;
}

/* rng4 for a pixel seed (RNG_SEED), a sample index and a dimension id: sample.rs:78-82, for the tests. */
void or_rng4_at(uint32_t pixel_seed, uint32_t index, uint32_t d, float *r) {
    or_rng g = {pixel_seed, index, 0, 0};
    uint32_t s = pixel_seed ^ or_hash_hp(d);
    uint32_t i = or_owen_rev(rev32(index), s ^ 0xc2c1a055u) & 0xffff0000u;
    (void)g;
    r[0] = or_sobol_owen(i, 0, s ^ 0x39468210u);
    r[1] = or_sobol_owen(i, 1, s ^ 0xe9d8a845u);
    r[2] = or_sobol_owen(i, 2, s ^ 0x5f32b482u);
    r[3] = or_sobol_owen(i, 3, s ^ 0x1524cc56u);
}
